"""GPU 1D discrete adjoint (igr_gradient on a 1D context) against the
oracle's exact discrete JVP of the restated 1D map (dual numbers through
IGR_adjoints_1d/: SSP-RK3 loop of enzyme_adjoint.jl:34-63 and the
objectives :65-71, :110-117): the dot-product identity <J^T v, u> = <v, J u>
for the initial state, gamma and alpha, for every solver.  This stands in for
the reference's Enzyme reverse mode (Enzyme == ForwardDiff to 1e-8 in
IGR_adjoints_1d/tests/run_enzyme_test.jl:87)."""
import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs

pytestmark = pytest.mark.gpu


def close(a, b, tol):
    assert abs(a - b) <= tol * max(abs(a), abs(b), 1e-300), (a, b, abs(a - b) / max(abs(a), abs(b)))


def case(ne, seed):
    q0 = configs.smooth_random1(3, ne, seed=seed, amp=0.3)
    gam = 1.4
    dt = 0.3 * (1.0 / ne) / (7 * configs.max_wavespeed1(q0, gam))
    return q0, gam, dt


@pytest.mark.parametrize("solver,obj,alpha,n_iter", [
    ("pcg", 0, 3e-3, 12), ("pcg", 1, 3e-3, 0), ("pcg", 2, 3e-3, 1), ("pcg", 4, 3e-3, 25),
    ("jacobi", 0, 3e-3, 10), ("jacobi", 1, 3e-3, 1), ("chebyshev", 2, 3e-3, 8), ("chebyshev", 1, 3e-3, 1),
    ("pcg", 1, 0.0, 5)])
def test_gradient_1d_dot_product(orc, solver, obj, alpha, n_iter):
    print(f"[case] 1D gradient {solver} obj={obj} alpha={alpha} n_iter={n_iter}", flush=True)
    ne, ns = 10, 3
    q0, gam, dt = case(ne, 300 + obj)
    c = igr.Context(igr.PeriodicMesh1D(ne, 1.0, igr.DGBasis(3)))
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(obj, ns, dt, gam, alpha, n_iter, solver)
    g = orc.Grid1D(3, ne)
    rng = np.random.default_rng(310 + obj)
    for trial in range(2):
        u = rng.standard_normal(q0.shape)
        ug, ua = rng.standard_normal(), rng.standard_normal()
        Jr, dJ, _, _ = g.jvp_objective_run(gam, ug, alpha, ua * alpha, n_iter, igr.SOLVERS[solver], ns, dt, obj,
                                           q0, u)
        close(J, Jr, 1e-12)
        close(dJ, np.vdot(qb, u) + gb * ug + ab * ua * alpha, 1e-10)


@pytest.mark.parametrize("solver", ["pcg", "jacobi"])
def test_gradient_1d_checkpoint_schedules_bitwise_equal(solver):
    print(f"[case] 1D checkpoints {solver}", flush=True)
    ne, ns = 12, 9
    q0, gam, dt = case(ne, 330)
    res = []
    for slots in (2, 3, 5, 10):
        c = igr.Context(igr.PeriodicMesh1D(ne, 1.0, igr.DGBasis(3)))
        c.set_state(q0)
        res.append(c.gradient(1, ns, dt, gam, 2e-3, 9, solver, ckpt_slots=slots))
    for r in res[1:]:
        assert r[0] == res[0][0]
        np.testing.assert_array_equal(r[1], res[0][1])
        assert r[2] == res[0][2] and r[3] == res[0][3]


def test_enzyme_adjoint_ic_1d_primitive_form():
    """The 1D mirror returns the reference's primitive adjoints
    (IGR_adjoints_1d/enzyme_adjoint.jl:306-311) from the conservative ones."""
    print("[case] 1D primitive adjoints", flush=True)
    ne, ns = 10, 2
    q0, gam, dt = case(ne, 340)
    basis = igr.DGBasis(3)
    mesh = igr.PeriodicMesh1D(ne, 1.0, basis)
    J, ar, au, ap = igr.enzyme_adjoint_ic(q0, ns, dt, basis, mesh, gam, 2e-3, n_iter=6, solver="pcg",
                                          objective="l2_pressure")
    c = igr.Context(mesh)
    c.set_state(q0)
    J2, qb, _, _ = c.gradient(1, ns, dt, gam, 2e-3, 6, "pcg")
    u0 = q0[0] / q0[1]
    assert J == J2
    np.testing.assert_allclose(ar, qb[1] + u0 * qb[0] + u0 ** 2 / 2 * qb[2], rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(au, q0[1] * qb[0] + q0[1] * u0 * qb[2], rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(ap, qb[2] / (gam - 1), rtol=1e-14, atol=1e-300)
    with pytest.raises(ValueError):
        c.gradient(3, ns, dt, gam, 2e-3, 6, "pcg")     # the windowed objectives are 2D only


@pytest.mark.parametrize("p", [1, 2, 4])
def test_gradient_1d_other_degrees(orc, p):
    ne, ns = 8, 2
    q0 = configs.smooth_random1(p, ne, seed=350 + p, amp=0.3)
    gam = 1.4
    dt = 0.3 / ne / ((2 * p + 1) * configs.max_wavespeed1(q0, gam))
    c = igr.Context(igr.PeriodicMesh1D(ne, 1.0, igr.DGBasis(p)))
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(0, ns, dt, gam, 3e-3, 9, "chebyshev")
    u = np.random.default_rng(360 + p).standard_normal(q0.shape)
    Jr, dJ, _, _ = orc.Grid1D(p, ne).jvp_objective_run(gam, 0.2, 3e-3, 0.5 * 3e-3, 9, orc.CHEBYSHEV, ns, dt, 0,
                                                       q0, u)
    close(J, Jr, 1e-12)
    close(dJ, np.vdot(qb, u) + 0.2 * gb + 0.5 * 3e-3 * ab, 1e-10)
