"""GPU discrete adjoint against the oracle's exact discrete JVP (dual
numbers through the restated reference map): the dot-product identity
<v, J u> = <J^T v, u> for random u, v, component by component and for the
whole gradient, plus central finite differences of the GPU forward map.
This stands in for the reference's Enzyme checks (Enzyme == ForwardDiff to
1e-8 in IGR_adjoints_1d/tests/run_enzyme_test.jl:87; vs FD to 1e-3 in
IGR_adjoints_2d/tests/run_enzyme_test.jl:88)."""
import ctypes as C

import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs

pytestmark = pytest.mark.gpu


def ctx2(p, nx, ny, lx=1.0, ly=1.0):
    print(f"[case] 2D p={p} {nx}x{ny}", flush=True)
    return igr.Context(igr.CartesianMesh2D(nx, ny, lx, ly, igr.DGBasis(p)))


def close(a, b, tol):
    assert abs(a - b) <= tol * max(abs(a), abs(b), 1e-300), (a, b, abs(a - b) / max(abs(a), abs(b)))


@pytest.mark.parametrize("nx,ny", [(6, 5), (8, 8), (3, 1)])
def test_hyperbolic_vjp_dot_product(orc, nx, ny):
    c = ctx2(3, nx, ny, 1.0, 1.2)
    g = orc.Grid2D(3, nx, ny, 1.0, 1.2)
    q = configs.smooth_random2(3, nx, ny, seed=21, amp=0.3)
    rng = np.random.default_rng(22)
    S = 0.01 * rng.standard_normal(g.shape)
    u = rng.standard_normal(q.shape)
    uS = rng.standard_normal(g.shape)
    ug = 0.7
    v = rng.standard_normal(q.shape)
    _, Ju = g.jvp_hyperbolic_rhs(1.4, ug, q, u, S, uS)
    qb, sb, gb = c.hyperbolic_vjp(1.4, q, S, v)
    close(np.vdot(v, Ju), np.vdot(qb, u) + np.vdot(sb, uS) + gb * ug, 1e-12)


@pytest.mark.parametrize("nx,ny", [(6, 5), (8, 8), (2, 3)])
def test_apply_sip_vjp_dot_product(orc, nx, ny):
    c = ctx2(3, nx, ny, 1.0, 0.8)
    g = orc.Grid2D(3, nx, ny, 1.0, 0.8)
    rng = np.random.default_rng(nx + 10 * ny)
    rho = 1.0 + 0.5 * rng.random(g.shape)
    x = rng.standard_normal(g.shape)
    dx, drho, da = rng.standard_normal(g.shape), rng.standard_normal(g.shape), 0.3
    yb = rng.standard_normal(g.shape)
    a = 4e-3
    _, Ju = g.jvp_apply_sip(a, da, x, dx, rho, drho)
    xb, rb, ab = c.apply_sip_vjp(a, x, rho, yb)
    close(np.vdot(yb, Ju), np.vdot(xb, dx) + np.vdot(rb, drho) + ab * da, 1e-12)


SOLVE_CASES = [("pcg", 0), ("pcg", 1), ("pcg", 5), ("pcg", 30), ("jacobi", 0), ("jacobi", 1), ("jacobi", 7),
               ("jacobi", 30), ("chebyshev", 0), ("chebyshev", 1), ("chebyshev", 2), ("chebyshev", 6),
               ("chebyshev", 25)]


@pytest.mark.parametrize("solver,n_iter", SOLVE_CASES)
def test_solve_vjp_dot_product(orc, solver, n_iter):
    c = ctx2(3, 6, 5)
    g = orc.Grid2D(3, 6, 5)
    rng = np.random.default_rng(30 + n_iter)
    q = configs.smooth_random2(3, 6, 5, seed=31, amp=0.3)
    S0 = 0.01 * rng.standard_normal(g.shape)
    dq, dS0, da = rng.standard_normal(q.shape), rng.standard_normal(g.shape), 0.2
    Sb = rng.standard_normal(g.shape)
    a = 5e-3
    _, dS = g.jvp_solve_elliptic(a, da * a, n_iter, igr.SOLVERS[solver], q, dq, S0, dS0)
    c.set_state(q)
    qb, s0b, ab = c.solve_elliptic_vjp(a, n_iter, solver, S0, Sb)
    close(np.vdot(Sb, dS), np.vdot(qb, dq) + np.vdot(s0b, dS0) + ab * da * a, 1e-10)


OBJS = [(0, None), (1, None), (2, None), (3, (0.55, 0.5, 0.15, 0.0)), (4, None), (5, (0.4, 0.6, 0.1, 0.0)),
        (6, (0.5, 0.45, 0.3, 0.0))]


@pytest.mark.parametrize("obj,prm", OBJS)
def test_gradient_dot_product(orc, obj, prm):
    nx, ny, ns = 6, 5, 3
    c = ctx2(3, nx, ny)
    g = orc.Grid2D(3, nx, ny)
    q0 = configs.smooth_random2(3, nx, ny, seed=40 + obj, amp=0.3)
    gam, a = 1.4, 4e-3
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, gam))
    p4 = prm if prm is not None else (0.0, 0.0, 1.0, 0.0)
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(obj, ns, dt, gam, a, 20, "pcg", prm=p4)
    rng = np.random.default_rng(50 + obj)
    for trial in range(2):
        u = rng.standard_normal(q0.shape)
        ug, ua = rng.standard_normal(), rng.standard_normal()
        Jr, dJ, _, _ = g.jvp_objective_run(gam, ug, a, ua * a, 20, orc.PCG, ns, dt, obj, q0, u, p4)
        close(J, Jr, 1e-12)
        close(dJ, np.vdot(qb, u) + gb * ug + ab * ua * a, 1e-10)


@pytest.mark.parametrize("solver,obj,prm", [("jacobi", 1, None), ("jacobi", 3, (0.55, 0.5, 0.15, 0.0)),
                                            ("chebyshev", 0, None), ("chebyshev", 6, (0.5, 0.45, 0.3, 0.0))])
def test_gradient_dot_product_jacobi_chebyshev(orc, solver, obj, prm):
    nx, ny, ns = 6, 5, 3
    c = ctx2(3, nx, ny)
    g = orc.Grid2D(3, nx, ny)
    q0 = configs.smooth_random2(3, nx, ny, seed=90 + obj, amp=0.3)
    gam, a = 1.4, 4e-3
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, gam))
    p4 = prm if prm is not None else (0.0, 0.0, 1.0, 0.0)
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(obj, ns, dt, gam, a, 12, solver, prm=p4)
    rng = np.random.default_rng(95 + obj)
    for trial in range(2):
        u = rng.standard_normal(q0.shape)
        ug, ua = rng.standard_normal(), rng.standard_normal()
        Jr, dJ, _, _ = g.jvp_objective_run(gam, ug, a, ua * a, 12, igr.SOLVERS[solver], ns, dt, obj, q0, u, p4)
        close(J, Jr, 1e-12)
        close(dJ, np.vdot(qb, u) + gb * ug + ab * ua * a, 1e-10)


def test_gradient_alpha_zero_and_deterministic(orc):
    nx, ny, ns = 5, 4, 4
    q0 = configs.smooth_random2(3, nx, ny, seed=60, amp=0.3)
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, 1.4))
    c = ctx2(3, nx, ny)
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(1, ns, dt, 1.4, 0.0, 1, "jacobi")
    g = orc.Grid2D(3, nx, ny)
    u = np.random.default_rng(61).standard_normal(q0.shape)
    Jr, dJ, _, _ = g.jvp_objective_run(1.4, 0.5, 0.0, 0.0, 1, orc.JACOBI, ns, dt, 1, q0, u)
    close(J, Jr, 1e-12)
    close(dJ, np.vdot(qb, u) + 0.5 * gb, 1e-10)
    assert ab == 0.0


@pytest.mark.parametrize("tile,htile,rtile", [(808, 804, 808), (808, 804, 804), (404, 404, 808), (202, 202, 202)])
def test_gradient_dot_product_tiles(orc, tile, htile, rtile):
    nx, ny, ns = 12, 10, 2
    c = ctx2(3, nx, ny)
    c.set_option(igr.core.OPT_TILE, tile)
    c.set_option(igr.core.OPT_HTILE, htile)
    c.set_option(igr.core.OPT_RTILE, rtile)
    assert c.get_option(igr.core.OPT_RTILE) == rtile
    g = orc.Grid2D(3, nx, ny)
    q0 = configs.smooth_random2(3, nx, ny, seed=120, amp=0.3)
    gam, a = 1.4, 4e-3
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, gam))
    p4 = (0.5, 0.5, 0.2, 0.0)
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(3, ns, dt, gam, a, 15, "pcg", prm=p4)
    u = np.random.default_rng(121).standard_normal(q0.shape)
    Jr, dJ, _, _ = g.jvp_objective_run(gam, 0.4, a, 0.3 * a, 15, orc.PCG, ns, dt, 3, q0, u, p4)
    close(J, Jr, 1e-12)
    close(dJ, np.vdot(qb, u) + 0.4 * gb + 0.3 * a * ab, 1e-10)


def test_reverse_tile_defaults_and_rejects():
    # the reverse kernels take the forward SIP family's shape by default
    for n, want in ((32, (404, 404)), (128, (804, 804)), (256, (804, 804))):
        c = ctx2(3, n, n)
        assert (c.get_option(igr.core.OPT_TILE), c.get_option(igr.core.OPT_RTILE)) == want
    c = ctx2(3, 8, 8)
    with pytest.raises(ValueError):
        c.set_option(igr.core.OPT_RTILE, 303)
    assert c.get_option(igr.core.OPT_RTILE) == 404


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
def test_gradient_reverse_tiles_agree(solver):
    # the reverse shape changes only the summation order of the reverse dot
    # products: J is bit-equal (same forward), the gradient agrees to rounding;
    # with 2x2 tiles the 210 workgroups are summed by 64 threads, so the grid
    # total adds three chunks after the first (loaded together, four per pass)
    nx, ny, ns = 30, 28, 3
    q0 = configs.smooth_random2(3, nx, ny, seed=130, amp=0.3)
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, 1.4))
    res = []
    for rtile in (808, 804, 404, 202):
        c = ctx2(3, nx, ny)
        c.set_option(igr.core.OPT_TILE, 808)
        c.set_option(igr.core.OPT_RTILE, rtile)
        c.set_state(q0)
        res.append(c.gradient(3, ns, dt, 1.4, 3e-3, 12, solver, prm=(0.55, 0.5, 0.15, 0.0)))
    for r in res[1:]:
        assert r[0] == res[0][0]
        np.testing.assert_allclose(r[1], res[0][1], rtol=0, atol=1e-11 * np.abs(res[0][1]).max())
        close(r[2], res[0][2], 1e-10)
        close(r[3], res[0][3], 1e-10)


def test_gradient_reverse_groups(orc):
    """The reverse PCG iteration on fewer workgroups than reverse tiles
    (IGR_OPT_RGROUPS), each looping over every n-th tile and storing one
    partial sum per dot product: 210 tiles of 2x2 on 64, 13 and 1 workgroups
    (uneven shares, one workgroup running them all) agree with one workgroup
    per tile to rounding, and the same group count gives the same bits on a
    second call; on 12x10 elements (30 tiles of 2x2 on 7 and 4 workgroups)
    the gradient satisfies the dot-product identity with the oracle's exact
    JVP."""
    nx, ny, ns = 30, 28, 3
    q0 = configs.smooth_random2(3, nx, ny, seed=131, amp=0.3)
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, 1.4))
    prm = (0.55, 0.5, 0.15, 0.0)
    res = {}
    for rg in (0, 64, 13, 1):
        c = ctx2(3, nx, ny)
        c.set_option(igr.core.OPT_RTILE, 202)
        c.set_option(igr.core.OPT_RGROUPS, rg)
        assert c.get_option(igr.core.OPT_RGROUPS) == rg
        c.set_state(q0)
        res[rg] = c.gradient(3, ns, dt, 1.4, 3e-3, 12, "pcg", prm=prm)
        if rg == 13:
            c.set_state(q0)
            again = c.gradient(3, ns, dt, 1.4, 3e-3, 12, "pcg", prm=prm)
            np.testing.assert_array_equal(again[1], res[rg][1])
        c.close()
    for rg in (64, 13, 1):
        assert res[rg][0] == res[0][0]
        np.testing.assert_allclose(res[rg][1], res[0][1], rtol=0, atol=1e-11 * np.abs(res[0][1]).max())
        close(res[rg][2], res[0][2], 1e-10)
        close(res[rg][3], res[0][3], 1e-10)
    nx, ny, ns = 12, 10, 2
    q0 = configs.smooth_random2(3, nx, ny, seed=120, amp=0.3)
    gam, a = 1.4, 4e-3
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, gam))
    p4 = (0.5, 0.5, 0.2, 0.0)
    u = np.random.default_rng(121).standard_normal(q0.shape)
    Jr, dJ, _, _ = orc.Grid2D(3, nx, ny).jvp_objective_run(gam, 0.4, a, 0.3 * a, 15, orc.PCG, ns, dt, 3, q0, u, p4)
    for rg in (7, 4):
        c = ctx2(3, nx, ny)
        c.set_option(igr.core.OPT_RTILE, 202)
        c.set_option(igr.core.OPT_RGROUPS, rg)
        c.set_state(q0)
        J, qb, gb, ab = c.gradient(3, ns, dt, gam, a, 15, "pcg", prm=p4)
        c.close()
        close(J, Jr, 1e-12)
        close(dJ, np.vdot(qb, u) + 0.4 * gb + 0.3 * a * ab, 1e-10)
    with pytest.raises(ValueError):
        ctx2(3, 8, 8).set_option(igr.core.OPT_RGROUPS, -1)


@pytest.mark.parametrize("solver", ["pcg", "chebyshev"])
def test_checkpoint_schedules_bitwise_equal(solver):
    nx, ny, ns = 6, 6, 9
    q0 = configs.smooth_random2(3, nx, ny, seed=70, amp=0.3)
    dt = 0.3 * (1 / nx) / (7 * configs.max_wavespeed2(q0, 1.4))
    res = []
    # slots 2, 3, 5: binomial schedules, with the iterate cache off (keep 0),
    # small (windows of 2 or 4 steps, more windows than free slots) or as
    # large as the memory holds (one window); 10: every step checkpointed
    for slots in (2, 3, 5, 10):
        for keep in (0, 1, 3, -1):
            c = ctx2(3, nx, ny)
            c.set_option(igr.core.OPT_KEEP, keep)
            assert c.get_option(igr.core.OPT_KEEP) == keep
            c.set_state(q0)
            res.append(c.gradient(3, ns, dt, 1.4, 3e-3, 15, solver, prm=(0.55, 0.5, 0.15, 0.0), ckpt_slots=slots))
            c.close()
    with pytest.raises(ValueError):
        ctx2(3, nx, ny).set_option(igr.core.OPT_KEEP, -2)
    for r in res[1:]:
        assert r[0] == res[0][0]
        np.testing.assert_array_equal(r[1], res[0][1])
        assert r[2] == res[0][2] and r[3] == res[0][3]


def test_gradient_c3_32_vs_oracle_and_fd(orc):
    # the C3 gradient check mesh: 32x32 triple deposit, windowed-density objective, PCG-60
    cfg = configs.config2(N=32, seed=2)
    prm = configs.config3_objective()
    ns = 4
    q0, gam, a, dt = cfg["q"], cfg["gamma"], cfg["alpha"], cfg["dt"]
    c = ctx2(3, 32, 32)
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(5, ns, dt, gam, a, 60, "pcg", prm=prm)
    g = orc.Grid2D(3, 32, 32)
    rng = np.random.default_rng(80)
    u = rng.standard_normal(q0.shape) * (q0.std(axis=(1, 2, 3, 4), keepdims=True) + 1e-2)
    Jr, dJ, _, _ = g.jvp_objective_run(gam, 0.1, a, 0.2 * a, 60, orc.PCG, ns, dt, 5, q0, u, prm)
    close(J, Jr, 1e-12)
    close(dJ, np.vdot(qb, u) + 0.1 * gb + 0.2 * a * ab, 1e-9)
    # central FD on the GPU forward map in the gamma and alpha directions
    def Jat(gg, aa):
        c.set_state(q0)
        c.run_steps(ns, dt, gg, aa, 60, "pcg")
        return c.objective(5, gg, prm)
    h = 1e-6
    fd_g = (Jat(gam * (1 + h), a) - Jat(gam * (1 - h), a)) / (2 * h * gam)
    fd_a = (Jat(gam, a * (1 + h)) - Jat(gam, a * (1 - h))) / (2 * h * a)
    close(fd_g, gb, 1e-6)
    close(fd_a, ab, 1e-5)


@pytest.mark.parametrize("alpha", [0.0, 1.5e-3])
@pytest.mark.parametrize("obj", ["l2_density", "weighted_momentum", "kinetic_energy"])
def test_enzyme_2d_case_vs_fd(orc, obj, alpha):
    """The reference's 2D gradient check (IGR_adjoints_2d/tests/run_enzyme_test.jl:14-99):
    Taylor-Green (A = 0.4) on 4x4 elements at p = 2 to T = 0.08 with
    dt = 0.4 h / ((2p + 1) max(|u_x| + c, |u_y| + c)) rounded to a whole step
    count, Chebyshev-60 (1 iteration at alpha = 0); central differences with
    h = 1e-6 at six seeded nodes of every field.  The reference accepts a
    relative error (max |FD - gradient| over max |gradient|) below 1e-3, the
    bar kept here: central differences of this map straddle the switches of
    the LLF wave-speed max (the Taylor-Green state is symmetric, so left and
    right speeds tie at faces), where it has no derivative.  The exact
    derivative of the discrete map is pinned by the dot-product identity
    against the oracle's dual-number JVP: to 1e-10 at alpha = 0, where the
    GPU map is bit-equal to the oracle's and every tie breaks the same way
    (the forward's choice: the first argument of max).  At alpha > 0 the
    Sigma solve agrees to ~1e-13, not bitwise, so a tie of this symmetric
    state can break the other way in one of the two maps, whose one-sided
    derivatives then differ (1e-6 of the gradient here): the identity is
    checked to 1e-10 at q0 perturbed by 1e-9, which breaks the ties."""
    gam, A, T, p, ne = 1.4, 0.4, 0.08, 2, 4
    q0 = configs.taylor_green2(p, ne, ne, A=A, gamma=gam)
    mx, my, r, E = q0
    c_s = np.sqrt(gam * (gam - 1.0) * (E - (mx * mx + my * my) / (2.0 * r)) / r)
    ws = max(np.max(np.abs(mx / r) + c_s), np.max(np.abs(my / r) + c_s))
    dt = 0.4 * (1.0 / ne) / ((2 * p + 1) * ws)
    ns = int(np.ceil(T / dt))
    dt = T / ns
    n_iter = 60 if alpha > 0 else 1
    key = igr.OBJECTIVES[obj]
    c = ctx2(p, ne, ne)
    c.set_state(q0)
    J, qb, _, _ = c.gradient(key, ns, dt, gam, alpha, n_iter, "chebyshev")

    def Jat(q):
        c.set_state(q)
        c.run_steps(ns, dt, gam, alpha, n_iter, "chebyshev")
        return c.objective(key, gam)

    assert abs(Jat(q0) - J) <= 1e-14 * max(abs(J), 1.0)   # weighted momentum: J = 0 by symmetry
    rng = np.random.default_rng(42)
    h = 1e-6
    err = 0.0
    for f in range(4):
        for idx in rng.integers(0, q0[f].size, 6):
            qp, qm = q0.copy(), q0.copy()
            qp[f].flat[idx] += h
            qm[f].flat[idx] -= h
            err = max(err, abs((Jat(qp) - Jat(qm)) / (2 * h) - qb[f].flat[idx]))
    c.close()
    relerr = err / np.max(np.abs(qb))
    print(f"[enzyme 2D] {obj} alpha={alpha:.1e} J={J:.6e} steps={ns} rel={relerr:.2e}", flush=True)
    assert relerr < 1e-3
    rng = np.random.default_rng(43)
    u = rng.standard_normal(q0.shape)
    g = orc.Grid2D(p, ne, ne)
    Jr, dJ, _, _ = g.jvp_objective_run(gam, 0.0, alpha, 0.0, n_iter, orc.CHEBYSHEV, ns, dt, key, q0, u)
    assert abs(J - Jr) <= 1e-12 * max(abs(Jr), 1.0)
    if alpha == 0.0:
        close(dJ, np.vdot(qb, u), 1e-10)
    else:
        q1 = q0 * (1.0 + 1e-9 * rng.standard_normal(q0.shape))
        c = ctx2(p, ne, ne)
        c.set_state(q1)
        J1, qb1, _, _ = c.gradient(key, ns, dt, gam, alpha, n_iter, "chebyshev")
        c.close()
        Jr, dJ, _, _ = g.jvp_objective_run(gam, 0.0, alpha, 0.0, n_iter, orc.CHEBYSHEV, ns, dt, key, q1, u)
        assert abs(J1 - Jr) <= 1e-12 * max(abs(Jr), 1.0)
        close(dJ, np.vdot(qb1, u), 1e-10)


def test_blast_vortex_32_forward_and_gradient(orc):
    """The blast-beside-Oseen-pair workload at 32x32 (3 steps): forward state
    vs the oracle, and the disc-kinetic-energy gradient vs the oracle's JVP."""
    cfg = configs.config5(N=32)
    q0, gam, a, dt = cfg["q"], cfg["gamma"], cfg["alpha"], cfg["dt"]
    c = ctx2(3, 32, 32)
    g = orc.Grid2D(3, 32, 32)
    c.set_state(q0)
    c.run_steps(3, dt, gam, a, 60, "pcg")
    q3 = c.get_state()[0]
    q3r, _ = g.run_steps(gam, a, 60, orc.PCG, 3, dt, q0)
    assert np.max(np.abs(q3 - q3r)) <= 1e-11 * np.max(np.abs(q3r))
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(6, 3, dt, gam, a, 60, "pcg", prm=cfg["disc"])
    u = np.random.default_rng(81).standard_normal(q0.shape) * (q0.std(axis=(1, 2, 3, 4), keepdims=True) + 1e-2)
    Jr, dJ, _, _ = g.jvp_objective_run(gam, 0.3, a, 0.2 * a, 60, orc.PCG, 3, dt, 6, q0, u, cfg["disc"])
    close(J, Jr, 1e-12)
    close(dJ, np.vdot(qb, u) + 0.3 * gb + 0.2 * a * ab, 1e-9)


@pytest.mark.parametrize("p", [2, 4])
def test_gradient_other_degrees(orc, p):
    """The 2D adjoint at p = 2 and p = 4 against the oracle's JVP."""
    nx, ny, ns = 5, 4, 2
    c = ctx2(p, nx, ny)
    g = orc.Grid2D(p, nx, ny)
    q0 = configs.smooth_random2(p, nx, ny, seed=130 + p, amp=0.3)
    gam, a = 1.4, 4e-3
    dt = 0.3 * (1 / nx) / ((2 * p + 1) * configs.max_wavespeed2(q0, gam))
    c.set_state(q0)
    J, qb, gb, ab = c.gradient(1, ns, dt, gam, a, 12, "pcg")
    u = np.random.default_rng(131).standard_normal(q0.shape)
    Jr, dJ, _, _ = g.jvp_objective_run(gam, 0.4, a, 0.3 * a, 12, orc.PCG, ns, dt, 1, q0, u)
    close(J, Jr, 1e-12)
    close(dJ, np.vdot(qb, u) + 0.4 * gb + 0.3 * a * ab, 1e-10)


def test_gradient_across_contexts_and_release_cache():
    """The kept-iterate block outlives its context and serves the next
    gradient (any context); results are bit-equal with a fresh block after
    igr_release_cache."""
    print("[case] gradient pool across contexts", flush=True)
    q0 = configs.smooth_random2(3, 12, 12, seed=4, amp=0.2)
    dt = 0.3 * (1 / 12) / (7 * configs.max_wavespeed2(q0, 1.4))
    outs = []
    for release in (False, False, True):
        if release:
            igr.release_cache()
        c = igr.Context(igr.CartesianMesh2D(12, 12, 1.0, 1.0, igr.DGBasis(3)))
        c.set_state(q0)
        outs.append(c.gradient(0, 6, dt, 1.4, 2e-3, 15, "pcg"))
        c.close()
    for J, qb, gb, ab in outs[1:]:
        assert J == outs[0][0] and gb == outs[0][2] and ab == outs[0][3]
        np.testing.assert_array_equal(qb, outs[0][1])


def _hip_free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_objective_and_gradient_alternating_keep_memory_flat():
    """The objective and VJP entry points (n_iter = 0) reuse the gradient's
    workspace, and a workspace that is replaced frees its buffers: free
    device memory does not fall across repeated objective/gradient calls."""
    print("[case] workspace reuse 24x24", flush=True)
    q0 = configs.smooth_random2(3, 24, 24, seed=91, amp=0.2)
    dt = 0.3 * (1 / 24) / (7 * configs.max_wavespeed2(q0, 1.4))
    c = ctx2(3, 24, 24)
    free = []
    for k in range(4):
        c.set_state(q0)
        c.objective(3, 1.4)
        c.gradient(3, 2, dt, 1.4, 3e-3, 20 + 10 * (k % 2), "pcg", ckpt_slots=2)
        free.append(_hip_free_bytes())
    # one field is 3 MB here; a leaked workspace would be > 100 MB per cycle
    assert max(free[1:]) - min(free[1:]) < 64 << 20, free
