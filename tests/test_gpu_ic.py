"""Device-built initial conditions (igr_build_ic) against the host
restatements of the reference's builders, their amplitude derivatives
against central differences of the device builder, and the state they leave
driving a run exactly as an uploaded one does."""
import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs

pytestmark = pytest.mark.gpu

# exp, sin, cos and pow of the device and of numpy may differ in the last place
RTOL = 1e-13


def _close(dev, host):
    sc = max(np.max(np.abs(host)), 1.0)
    np.testing.assert_allclose(dev, host, rtol=RTOL, atol=RTOL * sc)


def _ctx2(Nex, Ney, Lx=1.0, Ly=1.0, p=3):
    return igr.Context(igr.CartesianMesh2D(Nex, Ney, Lx, Ly, igr.DGBasis(p)))


@pytest.mark.parametrize("Lx,Ly", [(1.0, 1.0), (1.0, 1.3)])
def test_2d_builders_match_host(Lx, Ly):
    print(f"[case] 2D IC builders Lx={Lx} Ly={Ly}", flush=True)
    Nex, Ney, g = 9, 7, 1.4
    c = _ctx2(Nex, Ney, Lx, Ly)
    c.build_ic("taylor_green", 0.3, g)
    _close(c.get_state()[0], configs.taylor_green2(3, Nex, Ney, 0.3, g, Lx, Ly))
    c.build_ic("sedov_blast", 5.0, g, prm=(0.5, 0.12))
    _close(c.get_state()[0], configs.sedov_blast2(3, Nex, Ney, 5.0, g, 0.5, 0.12, Lx, Ly))
    c.build_ic("sedov_blast", 5.0, g)          # keyword defaults p_bg = 1, sigma = 0.1
    _close(c.get_state()[0], configs.sedov_blast2(3, Nex, Ney, 5.0, g, 1.0, 0.1, Lx, Ly))
    c.build_ic("isentropic_vortex", 0.0, g, prm=(4.0, 0.5, -0.25))
    _close(c.get_state()[0], configs.isentropic_vortex2(3, Nex, Ney, g, 4.0, 0.5, -0.25, Lx, Ly))
    c.build_ic("constant", 0.0, g, prm=(1.2, 0.3, -0.7, 2.0))
    _close(c.get_state()[0], configs.constant2(3, Nex, Ney, g, 1.2, 0.3, -0.7, 2.0))
    c.close()


def test_oseen_pair_blast_matches_host():
    print("[case] Oseen pair blast 16x16", flush=True)
    N, g = 16, 1.4
    c = _ctx2(N, N)
    x, y = configs.coords2(3, N, N)
    # the blast-vortex workload's parameters
    prm = (0.05, 0.5, 0.25, 0.50, 0.03, 0.45, 0.60, 1.0, 0.45, 0.40, -1.0)
    c.build_ic("oseen_pair_blast", 100.0, g, prm=prm)
    _close(c.get_state()[0], configs.oseen_pair_blast(x, y, g, A=100.0, sigma=0.03))
    # the reference's keyword defaults (sigma = 0.09)
    c.build_ic("oseen_pair_blast", 2.0, g)
    _close(c.get_state()[0], configs.oseen_pair_blast(x, y, g, A=2.0, sigma=0.09))
    # a single vortex
    c.build_ic("oseen_pair_blast", 1.0, 1.67, prm=(0.07, 0.3, 0.7, 0.2, 0.05, 0.5, 0.5, -1.0))
    _close(c.get_state()[0], configs.oseen_pair_blast(x, y, 1.67, A=1.0, sigma=0.05, a_core=0.07, M_star=0.3,
                                                      centers=((0.5, 0.5, -1.0),), blast=(0.7, 0.2)))
    c.close()


@pytest.mark.parametrize("kind", ["acoustic", "entropy", "density", "pressure", "velocity"])
def test_1d_pulses_and_seeds_match_host(kind):
    print(f"[case] 1D pulse {kind}", flush=True)
    Ne, g, A = 24, 1.4, 0.2
    c = igr.Context(igr.PeriodicMesh1D(Ne, 1.0, igr.DGBasis(3)))
    name = "entropy_wave" if kind == "entropy" else kind + "_pulse"
    dq = c.build_ic(name, A, g, dA=True)
    _close(c.get_state()[0], configs.pulse1(kind, 3, Ne, A, g))
    _close(dq, configs.pulse1_dA(kind, 3, Ne, A, g))
    c.close()


def test_1d_init_builders_match_host():
    print("[case] 1D init builders", flush=True)
    Ne, g = 20, 1.4
    c = igr.Context(igr.PeriodicMesh1D(Ne, 1.0, igr.DGBasis(3)))
    c.build_ic("init_entropy_wave", 0.0, g)
    _close(c.get_state()[0], configs.entropy_wave_init1(3, Ne, g))
    c.build_ic("constant", 0.0, g, prm=(1.5, -0.3, 0.8))
    q = c.get_state()[0]
    _close(q, np.broadcast_to(np.array([1.5 * -0.3, 1.5, 0.8 / 0.4 + 0.5 * 1.5 * 0.09])[:, None, None], q.shape))
    c.close()


@pytest.mark.parametrize("ic,A,prm", [("taylor_green", 0.3, None), ("sedov_blast", 4.0, (0.5, 0.1)),
                                      ("oseen_pair_blast", 50.0, None)])
def test_2d_amplitude_seed_is_the_derivative(ic, A, prm):
    print(f"[case] 2D dA seed {ic}", flush=True)
    c = _ctx2(8, 8)
    dq = c.build_ic(ic, A, 1.4, prm, dA=True)
    h = 1e-3 * A
    c.build_ic(ic, A + h, 1.4, prm)
    qp = c.get_state()[0]
    c.build_ic(ic, A - h, 1.4, prm)
    qm = c.get_state()[0]
    fd = (qp - qm) / (2 * h)
    # E is linear (Sedov, Oseen) or quadratic (Taylor-Green) in A: the
    # central difference is exact up to rounding
    sc = np.max(np.abs(dq))
    assert sc > 0
    np.testing.assert_allclose(dq, fd, rtol=0, atol=1e-8 * sc + 1e-11 * np.max(np.abs(qp)) / h)
    c.close()


def test_built_state_runs_like_uploaded_state():
    print("[case] device IC drives run_forward", flush=True)
    N, g = 12, 1.4
    c1, c2 = _ctx2(N, N), _ctx2(N, N)
    c1.build_ic("sedov_blast", 3.0, g)
    c2.set_state(configs.sedov_blast2(3, N, N, 3.0, g))
    q_dev, _ = c1.get_state()
    c2.set_state(q_dev)          # the same bits on both, so the runs must agree exactly
    dt, T = 2e-4, 1e-3
    c1.run_forward(dt, T, g, 1e-3, 20, "pcg")
    c2.run_forward(dt, T, g, 1e-3, 20, "pcg")
    np.testing.assert_array_equal(c1.get_state()[0], c2.get_state()[0])
    c1.close(); c2.close()


def test_unknown_or_wrong_dimension_ic_is_rejected():
    c = _ctx2(4, 4)
    with pytest.raises(ValueError):
        c.build_ic("no_such_ic", 1.0, 1.4)
    with pytest.raises((ValueError, igr.IGRError)):
        c.build_ic("acoustic_pulse", 1.0, 1.4)       # a 1D builder on a 2D mesh
    with pytest.raises((ValueError, igr.IGRError)):
        c.build_ic(99, 1.0, 1.4)
    with pytest.raises((ValueError, igr.IGRError)):
        c.build_ic("sedov_blast", 1.0, 1.4, prm=(1.0,))   # wrong parameter count
    c.close()
