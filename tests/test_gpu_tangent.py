"""GPU forward tangent (sensitivity) equations against the oracle's
restatement of run_sensitivity (IGR_adjoints_2d/sensitivity.jl:353-413,
IGR_adjoints_1d/sensitivity.jl:475-515) on the same inputs, and against
central finite differences of the GPU forward map (restating
IGR_adjoints_2d/tests/run_sensitivity_test.jl, combined L2 < 5e-4)."""
import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs
from test_oracle_tangent import tg_case, _pulse1

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("solver,alpha", [("chebyshev", 1.5e-3), ("pcg", 1.5e-3), ("jacobi", 1.5e-3), ("pcg", 0.0)])
def test_2d_tangent_matches_oracle(orc, solver, alpha):
    print(f"[case] 2D tangent {solver} alpha={alpha}", flush=True)
    g, q0, s0, dt, gam = tg_case(orc)
    code = igr.SOLVERS[solver]
    T = 20 * dt
    qr, Sr, sr, vsr, nsr = g.run_sensitivity(gam, alpha, 30, code, dt, T, q0, s0)
    c = igr.Context(igr.CartesianMesh2D(8, 8, 1.0, 1.0, igr.DGBasis(3)))
    c.set_state(q0)
    s, vs, ns = c.run_sensitivity(s0, dt, T, gam, alpha, 30, solver)
    q, S = c.get_state()
    assert ns == nsr
    assert rel(q, qr) < 1e-9 and rel(s[0], sr) < 1e-9
    if alpha > 0:
        assert rel(S, Sr) < 1e-8 and rel(vs[0], vsr) < 1e-8


@pytest.mark.parametrize("solver,alpha", [("chebyshev", 1.5e-3), ("pcg", 0.0)])
def test_2d_gamma_tangent_matches_oracle(orc, solver, alpha):
    """A d/dgamma tangent in 2D (sgam = 1, an extension: the reference's 2D
    tangent has no gamma source) against the oracle, whose own gamma tangent
    is pinned to central differences in gamma (test_oracle_tangent.py)."""
    print(f"[case] 2D gamma tangent {solver} alpha={alpha}", flush=True)
    g, q0, _, dt, gam = tg_case(orc, A=0.3)
    mx, my, r, E = q0
    s0 = np.zeros_like(q0)
    s0[3] = -(gam - 1) * (E - (mx * mx + my * my) / (2 * r)) / (gam - 1) ** 2
    T = 12 * dt
    code = igr.SOLVERS[solver]
    qr, Sr, sr, vsr, nsr = g.run_sensitivity(gam, alpha, 30, code, dt, T, q0, s0, sgam=1.0)
    c = igr.Context(igr.CartesianMesh2D(8, 8, 1.0, 1.0, igr.DGBasis(3)))
    c.set_state(q0)
    s, vs, ns = c.run_sensitivity(s0, dt, T, gam, alpha, 30, solver, sgam=1.0)
    assert ns == nsr
    assert rel(s[0], sr) < 1e-9
    if alpha > 0:
        assert rel(vs[0], vsr) < 1e-8
    # in a batch with the reference's tangent (sgam = 0) beside it
    c2 = igr.Context(igr.CartesianMesh2D(8, 8, 1.0, 1.0, igr.DGBasis(3)))
    c.set_state(q0)
    c2.set_state(q0)
    out = igr.run_sensitivity_batch([dict(ctx=c, s0=s0[None], dt=dt, T=T, gamma=gam, alpha=alpha, sgamma=1.0),
                                     dict(ctx=c2, s0=s0[None], dt=dt, T=T, gamma=gam, alpha=alpha)],
                                    n_iter=30, solver=solver)
    np.testing.assert_array_equal(out[0][0], s)
    assert np.max(np.abs(out[1][0] - s)) > 1e-6 * np.max(np.abs(s))
    c.close()
    c2.close()


def test_2d_multiple_tangents_equal_single():
    print("[case] 2D four tangents", flush=True)
    q0 = configs.taylor_green2(3, 8, 8, A=0.3)
    rng = np.random.default_rng(5)
    S4 = rng.standard_normal((4,) + q0.shape) * 0.1
    dt = 0.4 * (1 / 8) / (7 * configs.max_wavespeed2(q0, 1.4))
    c = igr.Context(igr.CartesianMesh2D(8, 8, 1.0, 1.0, igr.DGBasis(3)))
    c.set_state(q0)
    s_all, v_all, _ = c.run_sensitivity(S4, dt, 10 * dt, 1.4, 2e-3, 20, "pcg")
    for t in range(4):
        c.set_state(q0)
        s1, v1, _ = c.run_sensitivity(S4[t], dt, 10 * dt, 1.4, 2e-3, 20, "pcg")
        np.testing.assert_array_equal(s1[0], s_all[t])
        np.testing.assert_array_equal(v1[0], v_all[t])


def test_2d_tangent_vs_gpu_fd(orc):
    print("[case] 2D tangent vs FD", flush=True)
    g, q0, s0, dt, gam = tg_case(orc)
    c = igr.Context(igr.CartesianMesh2D(8, 8, 1.0, 1.0, igr.DGBasis(3)))
    T, al = 0.1, 1.5e-3
    c.set_state(q0)
    s, _, _ = c.run_sensitivity(s0, dt, T, gam, al, 100, "chebyshev")
    h = 1e-5
    outs = []
    for A in (0.4 + h, 0.4 - h):
        c.set_state(configs.taylor_green2(3, 8, 8, A=A))
        c.run_forward(dt, T, gam, al, 100, "chebyshev")
        outs.append(c.get_state()[0])
    fd = (outs[0] - outs[1]) / (2 * h)
    assert np.linalg.norm(s[0] - fd) / np.linalg.norm(fd) < 5e-4


@pytest.mark.parametrize("form,sgam,solver", [(1, 0.0, "pcg"), (2, 0.0, "pcg"), (2, 1.0, "jacobi"), (1, 0.0, "chebyshev")])
def test_1d_tangent_matches_oracle(orc, form, sgam, solver):
    print(f"[case] 1D tangent form={form} sgam={sgam} {solver}", flush=True)
    g, ic, s0, dt, gam = _pulse1(orc)
    q0 = ic(0.05)
    T = 0.05
    code = igr.SOLVERS[solver]
    qr, Sr, sr, sSr, nsr = g.run_sensitivity(gam, 2e-3, 40, code, dt, T, q0, s0, form=form, sgam=sgam)
    c = igr.Context(igr.PeriodicMesh1D(16, 1.0, igr.DGBasis(3)))
    c.set_state(q0)
    s, sS, ns = c.run_sensitivity(s0, dt, T, gam, 2e-3, 40, solver, form=form, sgam=sgam)
    q, S = c.get_state()
    assert ns == nsr
    assert rel(q, qr) < 1e-10 and rel(s[0], sr) < 1e-9
    assert rel(S, Sr) < 1e-8 and rel(sS[0], sSr) < 1e-8
