"""Continuous adjoint on the GPU (igr_run_forward_store, igr_run_adjoint_conservative)
against the oracle's restatement of run_forward_store (IGR_adjoints_2d/forward.jl:111-158,
IGR_adjoints_1d/forward.jl:137-187) and run_adjoint_conservative
(IGR_adjoints_2d/adjoint_pde.jl:522-611, IGR_adjoints_1d/adjoint_pde.jl:466-552).

The adjoint runs on the oracle's own snapshots (uploaded with igr_set_snapshots)
so that its comparison does not inherit forward rounding; the forward store is
compared record by record.  Tolerance 1e-11 relative to the field's max (the
state) and 1e-10 for Sigma, the output of a fixed-iteration solve that is not
converged, where the GPU's different dot-product summation order is amplified."""
import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SOLVERS = {"pcg": O.PCG, "jacobi": O.JACOBI, "chebyshev": O.CHEBYSHEV}


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def setup2(p=3, N=12, seed=7, steps=3, cfl=0.3):
    q0 = configs.smooth_random2(p, N, N, seed=seed, amp=0.3)
    dt = cfl * (1.0 / N) / ((2 * p + 1) * configs.max_wavespeed2(q0, 1.4))
    return q0, dt, steps * dt * 0.999


def setup1(p=3, Ne=24, seed=7, steps=4, cfl=0.2):
    q0 = configs.smooth_random1(p, Ne, seed=seed, amp=0.3)
    dt = cfl * (1.0 / Ne) / ((2 * p + 1) * configs.max_wavespeed1(q0, 1.4))
    return q0, dt, steps * dt * 0.999


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
@pytest.mark.parametrize("alpha", [0.0, 2e-3])
def test_forward_store_2d(solver, alpha):
    print(f"[case] forward store 2D {solver} alpha={alpha}", flush=True)
    p, N = 3, 12
    q0, dt, T = setup2(p, N)
    c = igr.Context(igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(p)))
    c.set_state(q0)
    n = c.run_forward_store(dt, T, 1.4, alpha, 15, solver)
    ref, nr = O.Grid2D(p, N, N).run_forward_store(1.4, alpha, 15, SOLVERS[solver], dt, T, q0)
    assert n == nr == 3 and c.snapshot_count() == 3
    for k in range(n + 1):
        q, S = c.get_snapshot(k)
        assert rel(q, ref[k, :4]) < 1e-12, k
        if alpha > 0:
            assert rel(S, ref[k, 4]) < 1e-10, k
        else:
            assert not np.any(S)
    np.testing.assert_array_equal(c.get_state()[0], c.get_snapshot(n)[0])
    c.close()


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
@pytest.mark.parametrize("alpha", [0.0, 2e-3])
@pytest.mark.parametrize("obj", range(7))
def test_adjoint_pde_2d(solver, alpha, obj):
    print(f"[case] adjoint PDE 2D {solver} alpha={alpha} obj={obj}", flush=True)
    p, N = 3, 12
    q0, dt, T = setup2(p, N)
    g = O.Grid2D(p, N, N)
    prm = (0.45, 0.55, 0.2, 0.0)
    snaps, n = g.run_forward_store(1.4, alpha, 15, SOLVERS[solver], dt, T, q0)
    ref = g.run_adjoint_conservative(snaps, n, dt, T, 1.4, alpha, 15, SOLVERS[solver], obj, prm)
    c = igr.Context(igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(p)))
    c.set_snapshots(snaps)
    a = c.run_adjoint_conservative(obj, dt, T, 1.4, alpha, 15, solver, prm=prm)
    assert rel(a, ref) < 1e-11
    c.close()


def test_adjoint_pde_2d_own_snapshots_and_graphs():
    """forward store + adjoint on the GPU's own snapshots: graphs on and off
    bit-equal; tile shapes 4x4 / 8x4 / 2x2 within 1e-12 (the kinetic-energy
    and mass sums of the terminal condition are grid reductions over the
    tile's workgroups, so their order follows the shape); all within 1e-10 of
    the oracle's whole pipeline."""
    print("[case] adjoint PDE 2D pipeline", flush=True)
    p, N = 3, 20
    q0, dt, T = setup2(p, N, steps=4)
    g = O.Grid2D(p, N, N)
    snaps, n = g.run_forward_store(1.4, 3e-3, 20, O.PCG, dt, T, q0)
    ref = g.run_adjoint_conservative(snaps, n, dt, T, 1.4, 3e-3, 20, O.PCG, O.OBJ_KINETIC_ENERGY)
    outs = []
    for graphs, htile in [(1, 404), (0, 404), (1, 804), (1, 202)]:
        c = igr.Context(igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(p)))
        c.set_option(igr.core.OPT_GRAPHS, graphs)
        c.set_option(igr.core.OPT_HTILE, htile)
        c.set_state(q0)
        assert c.run_forward_store(dt, T, 1.4, 3e-3, 20, "pcg") == n
        outs.append(c.run_adjoint_conservative(O.OBJ_KINETIC_ENERGY, dt, T, 1.4, 3e-3, 20, "pcg"))
        c.close()
    np.testing.assert_array_equal(outs[1], outs[0])
    for o in outs[2:]:
        assert rel(o, outs[0]) < 1e-12
    assert rel(outs[0], ref) < 1e-10


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
@pytest.mark.parametrize("alpha", [0.0, 2e-3])
def test_forward_store_1d(solver, alpha):
    print(f"[case] forward store 1D {solver} alpha={alpha}", flush=True)
    p, Ne = 3, 24
    q0, dt, T = setup1(p, Ne)
    c = igr.Context(igr.PeriodicMesh1D(Ne, 1.0, igr.DGBasis(p)))
    c.set_state(q0)
    n = c.run_forward_store(dt, T, 1.4, alpha, 15, solver)
    ref, nr = O.Grid1D(p, Ne).run_forward_store(1.4, alpha, 15, SOLVERS[solver], dt, T, q0)
    assert n == nr
    for k in range(n + 1):
        q, S = c.get_snapshot(k)
        assert rel(q, ref[k, :3]) < 1e-12, k
        if alpha > 0:
            assert rel(S, ref[k, 3]) < 1e-10, k
    c.close()


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
@pytest.mark.parametrize("alpha", [0.0, 2e-3])
@pytest.mark.parametrize("obj", [0, 1, 2, 4])
def test_adjoint_pde_1d(solver, alpha, obj):
    print(f"[case] adjoint PDE 1D {solver} alpha={alpha} obj={obj}", flush=True)
    p, Ne = 3, 24
    q0, dt, T = setup1(p, Ne)
    g = O.Grid1D(p, Ne)
    snaps, n = g.run_forward_store(1.4, alpha, 15, SOLVERS[solver], dt, T, q0)
    ref = g.run_adjoint_conservative(snaps, n, dt, T, 1.4, alpha, 15, SOLVERS[solver], obj)
    c = igr.Context(igr.PeriodicMesh1D(Ne, 1.0, igr.DGBasis(p)))
    c.set_snapshots(snaps)
    a = c.run_adjoint_conservative(obj, dt, T, 1.4, alpha, 15, solver)
    assert rel(a, ref) < 1e-11
    c.close()


@pytest.mark.parametrize("p", [1, 2, 4])
def test_adjoint_pde_other_degrees(p):
    print(f"[case] adjoint PDE degree {p}", flush=True)
    N, Ne = 8, 16
    q0, dt, T = setup2(p, N)
    g = O.Grid2D(p, N, N)
    snaps, n = g.run_forward_store(1.4, 2e-3, 15, O.PCG, dt, T, q0)
    ref = g.run_adjoint_conservative(snaps, n, dt, T, 1.4, 2e-3, 15, O.PCG, O.OBJ_L2_PRESSURE)
    c = igr.Context(igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(p)))
    c.set_state(q0)
    assert c.run_forward_store(dt, T, 1.4, 2e-3, 15, "pcg") == n
    assert rel(c.run_adjoint_conservative(O.OBJ_L2_PRESSURE, dt, T, 1.4, 2e-3, 15, "pcg"), ref) < 1e-10
    c.close()
    q1, dt1, T1 = setup1(p, Ne)
    g1 = O.Grid1D(p, Ne)
    s1, n1 = g1.run_forward_store(1.4, 2e-3, 15, O.PCG, dt1, T1, q1)
    r1 = g1.run_adjoint_conservative(s1, n1, dt1, T1, 1.4, 2e-3, 15, O.PCG, O.OBJ_KINETIC_ENERGY)
    c1 = igr.Context(igr.PeriodicMesh1D(Ne, 1.0, igr.DGBasis(p)))
    c1.set_state(q1)
    assert c1.run_forward_store(dt1, T1, 1.4, 2e-3, 15, "pcg") == n1
    assert rel(c1.run_adjoint_conservative(O.OBJ_KINETIC_ENERGY, dt1, T1, 1.4, 2e-3, 15, "pcg"), r1) < 1e-10
    c1.close()


def test_quasi1d_adjoint_pde_gpu():
    """The reference's quasi-1D test 3 (run_quasi1d_test.jl:131-143) on the GPU:
    2D y-uniform == 1D to 1e-9, y-nonuniformity below 1e-11."""
    print("[case] quasi-1D adjoint PDE", flush=True)
    gam, L, p, Ne, A, T, CFL, Ney, Ly = 1.4, 1.0, 3, 12, 0.4, 0.08, 0.5, 4, 0.5
    q1 = configs.pulse1("velocity", p, Ne, A, gam)
    x2, _ = configs.coords2(p, Ne, Ney, L, Ly)
    u2 = A * np.sin(2 * np.pi * x2 / L)
    q2 = np.zeros((4,) + x2.shape)
    q2[0], q2[2], q2[3] = u2, 1.0, 1.0 / (gam - 1) + 0.5 * u2 ** 2
    dt = CFL * (L / Ne) / ((2 * p + 1) * configs.max_wavespeed1(q1, gam))
    ns = int(np.ceil(T / dt))
    dt = T / ns
    alpha = 3.0 * (L / Ne) ** 2
    snaps1, n1 = igr.run_forward_store(q1, dt, T, None, igr.PeriodicMesh1D(Ne, L, igr.DGBasis(p)), gam, alpha, 400)
    a1 = igr.run_adjoint_conservative(snaps1, n1, dt, T, None, None, gam, alpha, 400)
    snaps2, n2 = igr.run_forward_store(q2, dt, T, None, igr.CartesianMesh2D(Ne, Ney, L, Ly, igr.DGBasis(p)), gam,
                                       alpha, 400)
    a2 = igr.run_adjoint_conservative(snaps2, n2, dt, T, None, None, gam, alpha, 400)
    assert n1 == n2 == ns
    for f2, f1 in ((a2[0], a1[0]), (a2[2], a1[1]), (a2[3], a1[2])):
        assert np.max(np.abs(f2 - f1[None, :, None, :])) < 1e-9
        assert np.max(np.abs(f2 - f2[:1, :, :1, :])) < 1e-11
    assert np.max(np.abs(a2[1])) < 1e-11
    assert len(snaps1) == n1 + 1 and snaps2[-1][0].shape == (4,) + x2.shape


def test_adjoint_pde_errors():
    print("[case] adjoint PDE errors", flush=True)
    c = igr.Context(igr.CartesianMesh2D(4, 4, 1.0, 1.0, igr.DGBasis(3)))
    c.set_state(configs.taylor_green2(3, 4, 4))
    with pytest.raises(ValueError):
        c.run_adjoint_conservative(0, 1e-3, 1e-3, 1.4, 0.0, 1)        # no snapshots yet
    n = c.run_forward_store(1e-3, 2e-3, 1.4, 0.0, 1)
    assert n == 2
    with pytest.raises(ValueError):
        c.run_adjoint_conservative(0, 1e-3, 2e-3, 1.4, 0.0, 1, n_steps=3)   # more steps than stored
    with pytest.raises(ValueError):
        c.run_adjoint_conservative(9, 1e-3, 2e-3, 1.4, 0.0, 1)        # unknown objective
    with pytest.raises(ValueError):
        c.get_snapshot(3)
    c.close()
    c1 = igr.Context(igr.PeriodicMesh1D(8, 1.0, igr.DGBasis(3)))
    c1.set_state(configs.smooth_random1(3, 8, seed=1))
    c1.run_forward_store(1e-3, 2e-3, 1.4, 0.0, 1)
    with pytest.raises(ValueError):
        c1.run_adjoint_conservative(3, 1e-3, 2e-3, 1.4, 0.0, 1)       # windowed KE is 2D only
    c1.close()


@pytest.mark.parametrize("ic,A,prm", [("oseen_pair_blast", 2.0, None), ("sedov_blast", 3.0, (1.0, 0.1))])
def test_paper_script_path_device_ic(ic, A, prm):
    """The OtD-refinement scripts' path (IGR_adjoints_2d/paper_scripts/
    generate_otd_refinement_{oseen,sedov}_data.jl): initial condition from the
    reference's builder, run_forward_store, run_adjoint_conservative of the
    energy objective.  Here the IC is built on the device; the oracle runs the
    same steps from the same bits."""
    print(f"[case] paper-script path {ic}", flush=True)
    p, N, gam, alpha = 3, 12, 1.4, 1e-3
    c = igr.Context(igr.CartesianMesh2D(N, N, 1.0, 1.0, igr.DGBasis(p)))
    c.build_ic(ic, A, gam, prm)
    q0, _ = c.get_state()
    dt = 0.3 * (1.0 / N) / ((2 * p + 1) * configs.max_wavespeed2(q0, gam))
    T = 3 * dt * 0.999
    n = c.run_forward_store(dt, T, gam, alpha, 15, "pcg")
    g = O.Grid2D(p, N, N)
    ref, nr = g.run_forward_store(gam, alpha, 15, O.PCG, dt, T, q0)
    assert n == nr == 3
    q, S = c.get_snapshot(n)
    assert rel(q, ref[n, :4]) < 1e-12
    assert rel(S, ref[n, 4]) < 1e-10
    obj = igr.OBJECTIVES["kinetic_energy"]
    a = c.run_adjoint_conservative(obj, dt, T, gam, alpha, 15, "pcg")
    a_ref = g.run_adjoint_conservative(ref, n, dt, T, gam, alpha, 15, O.PCG, obj, (0.0, 0.0, 1.0, 0.0))
    # the GPU adjoint runs on its own snapshots here, so forward rounding is inherited
    assert rel(a, a_ref) < 1e-9
    c.close()
