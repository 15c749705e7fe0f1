"""GPU parity of the forward path against the CPU oracle (restated
reference) on the same seeded inputs.  Tolerances: the hyperbolic RHS is
evaluated in the reference's operation order without FMA contraction and
must match bit for bit; SIP/solver results (gather order, reciprocal
multiply) within 1e-12 relative; states within 1e-12 relative after one
step and 1e-9 after a full run."""
import sys

import numpy as np
import pytest

import igr_sensitivities_amd as igr
from igr_sensitivities_amd import configs

pytestmark = pytest.mark.gpu

MESHES_2D = [(8, 8), (6, 5), (1, 1), (3, 7), (32, 32), (9, 4)]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def ctx2(p, nx, ny, lx=1.0, ly=1.0):
    print(f"[case] 2D p={p} {nx}x{ny}", flush=True)
    return igr.Context(igr.CartesianMesh2D(nx, ny, lx, ly, igr.DGBasis(p)))


def ctx1(p, ne, L=1.0):
    print(f"[case] 1D p={p} Ne={ne}", flush=True)
    return igr.Context(igr.PeriodicMesh1D(ne, L, igr.DGBasis(p)))


@pytest.mark.parametrize("nx,ny", MESHES_2D)
def test_apply_sip_2d(orc, nx, ny):
    c = ctx2(3, nx, ny, 1.0, 0.7)
    g = orc.Grid2D(3, nx, ny, 1.0, 0.7)
    rng = np.random.default_rng(nx * 100 + ny)
    rho = 1.0 + 0.5 * rng.random(g.shape)
    x = rng.standard_normal(g.shape)
    for a in (0.0, 3e-3, 0.2):
        assert rel(c.apply_sip(a, x, rho), g.apply_sip(a, x, rho)) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 4])
def test_apply_sip_2d_other_degrees(orc, p):
    c = ctx2(p, 5, 6)
    g = orc.Grid2D(p, 5, 6)
    rng = np.random.default_rng(p)
    rho = 1.0 + 0.5 * rng.random(g.shape)
    x = rng.standard_normal(g.shape)
    assert rel(c.apply_sip(0.01, x, rho), g.apply_sip(0.01, x, rho)) < 1e-13


@pytest.mark.parametrize("nx,ny", MESHES_2D)
def test_hyperbolic_rhs_2d_bitwise(orc, nx, ny):
    c = ctx2(3, nx, ny, 1.0, 1.3)
    g = orc.Grid2D(3, nx, ny, 1.0, 1.3)
    q = configs.smooth_random2(3, nx, ny, seed=nx + ny, amp=0.3)
    S = 0.01 * np.random.default_rng(1).standard_normal(g.shape)
    np.testing.assert_array_equal(c.hyperbolic_rhs(1.4, q, S), g.hyperbolic_rhs(1.4, q, S))


@pytest.mark.parametrize("p,htile", [(1, 0), (2, 0), (3, 404), (3, 804), (3, 202), (4, 0)])
def test_hyperbolic_rhs_2d_bitwise_degrees_and_tiles(orc, p, htile):
    """Bit-equal hyperbolic RHS for every degree and p = 3 tile shape (the
    face items of a tile fill several waves, or loop when they outnumber
    the threads at p = 1), on meshes the tiles divide unevenly and on a
    single element (every face periodic onto itself)."""
    for nx, ny in ((20, 12), (9, 17), (1, 1)):
        c = ctx2(p, nx, ny, 1.0, 0.7)
        if htile:
            c.set_option(igr.core.OPT_HTILE, htile)
        g = orc.Grid2D(p, nx, ny, 1.0, 0.7)
        q = configs.smooth_random2(p, nx, ny, seed=7 * nx + ny + p, amp=0.3)
        S = 0.01 * np.random.default_rng(p).standard_normal(g.shape)
        np.testing.assert_array_equal(c.hyperbolic_rhs(1.4, q, S), g.hyperbolic_rhs(1.4, q, S)), (p, htile, nx, ny)


def test_igr_rhs_and_diag_2d(orc):
    c = ctx2(3, 6, 5)
    g = orc.Grid2D(3, 6, 5)
    q = configs.smooth_random2(3, 6, 5, seed=3, amp=0.3)
    for a in (1e-3, 0.1):
        assert rel(c.igr_rhs(a, q), g.igr_rhs(a, q)) < 1e-13
        assert rel(c.sip_diagonal(a, q[2]), g.sip_diag(a, q[2])) < 1e-13


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
@pytest.mark.parametrize("nx,ny", [(8, 8), (6, 5), (1, 1)])
def test_solve_elliptic_2d(orc, solver, nx, ny):
    c = ctx2(3, nx, ny)
    g = orc.Grid2D(3, nx, ny)
    q = configs.smooth_random2(3, nx, ny, seed=7, amp=0.3)
    S0 = 0.01 * np.random.default_rng(2).standard_normal(g.shape)
    code = igr.SOLVERS[solver]
    for n_iter in (0, 1, 7, 30):
        c.set_state(q, S0)
        c.solve_elliptic(5e-3, n_iter, solver)
        S_gpu = c.get_state()[1]
        S_ref = g.solve_elliptic(5e-3, n_iter, code, q, S0)
        assert rel(S_gpu, S_ref) < 1e-12, (solver, n_iter)


def test_pcg_trace_2d(orc):
    c = ctx2(3, 8, 8)
    g = orc.Grid2D(3, 8, 8)
    q = configs.smooth_random2(3, 8, 8, seed=9, amp=0.3)
    a = 5e-3
    c.set_state(q, np.zeros(g.shape))
    c.solve_elliptic(a, 20, "pcg")
    ta, tb = c.solver_trace(20)
    b = g.igr_rhs(a, q)
    M = g.sip_diag(a, q[2])
    _, ra, rb = g.solver(a, orc.PCG, 20, q[2], b, M, g.field(), trace=True)
    np.testing.assert_allclose(ta, ra, rtol=1e-10)
    np.testing.assert_allclose(tb, rb, rtol=1e-9)


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
def test_euler_rhs_2d(orc, solver):
    c = ctx2(3, 6, 5)
    g = orc.Grid2D(3, 6, 5)
    q = configs.smooth_random2(3, 6, 5, seed=5, amp=0.3)
    c.set_state(q, np.zeros(g.shape))
    dq = c.euler_rhs(1.4, 4e-3, 10, solver)
    dq_ref, S_ref = g.euler_rhs(1.4, 4e-3, 10, igr.SOLVERS[solver], q, g.field())
    assert rel(c.get_state()[1], S_ref) < 1e-12
    assert rel(dq, dq_ref) < 1e-12


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
def test_one_step_and_run_2d(orc, solver):
    c = ctx2(3, 8, 6)
    g = orc.Grid2D(3, 8, 6)
    q0 = configs.taylor_green2(3, 8, 6, A=0.3)
    gam, a = 1.4, 3e-3
    dt = 0.4 * (1 / 8) / (7 * configs.max_wavespeed2(q0, gam))
    code = igr.SOLVERS[solver]
    c.set_state(q0)
    c.run_steps(1, dt, gam, a, 20, solver)
    q1 = c.get_state()[0]
    q1r, _ = g.run_steps(gam, a, 20, code, 1, dt, q0)
    assert rel(q1, q1r) < 1e-12
    c.set_state(q0)
    T = 25 * dt
    ns = c.run_forward(dt, T, gam, a, 20, solver)
    qT, ST = c.get_state()
    qTr, STr, nsr = g.run_forward(gam, a, 20, code, dt, T, q0)
    assert ns == nsr
    assert rel(qT, qTr) < 1e-9
    assert rel(ST, STr) < 1e-9


@pytest.mark.parametrize("n,tile", [(80, None), (30, 202)])
def test_default_tiles_with_more_workgroups_than_threads(orc, n, tile):
    # More partial sums than threads per workgroup, so the grid total adds
    # chunks after the first (one value per thread each).  80x80 with the
    # default 4x4 tiles: 400 workgroups of 256 threads, one further chunk.
    # 30x28 with 2x2 tiles: 210 workgroups of 64 threads, three further
    # chunks, loaded together in one pass of the grid total.
    ny = n if tile is None else 28
    c = ctx2(3, n, ny)
    if tile is None:
        assert c.get_option(igr.core.OPT_TILE) == 404
    else:
        c.set_option(igr.core.OPT_TILE, tile)
    g = orc.Grid2D(3, n, ny)
    q0 = configs.smooth_random2(3, n, ny, seed=140, amp=0.3)
    gam, a = 1.4, configs.alpha_of(1 / n, 3)
    dt = 0.4 * (1 / n) / (7 * configs.max_wavespeed2(q0, gam))
    c.set_state(q0)
    c.run_steps(1, dt, gam, a, 30, "pcg")
    q1, S1 = c.get_state()
    q1r, S1r = g.run_steps(gam, a, 30, orc.PCG, 1, dt, q0)
    assert rel(q1, q1r) < 1e-12
    # Sigma is an inexact 30-iteration PCG iterate after three warm-started
    # solves; on 10^5 nodes the dot products' summation order (a tree here,
    # serial in the oracle) shows in it at about 1e-11.  A partial-sum chunk
    # left out or added twice would move alpha by tens of percent.
    assert rel(S1, S1r) < 1e-10


def test_graph_replay_matches_eager_and_is_deterministic():
    # several replays of the captured step, with eager work and syncs between them,
    # must equal the eager launches bit for bit
    q0 = configs.taylor_green2(3, 16, 16, A=0.3)
    dt = 0.4 * (1 / 16) / (7 * configs.max_wavespeed2(q0, 1.4))
    outs = []
    for graphs in (1, 0, 1):
        c = ctx2(3, 16, 16)
        c.set_option(0, graphs)
        c.set_state(q0)
        for _ in range(3):
            c.run_steps(4, dt, 1.4, 2e-3, 60, "pcg")
            outs.append(c.get_state()[0].copy())
    for k in range(3):
        np.testing.assert_array_equal(outs[k], outs[3 + k])
        np.testing.assert_array_equal(outs[k], outs[6 + k])


@pytest.mark.parametrize("nx,ny", [(24, 24), (20, 12)])
def test_tile_shapes_agree(orc, nx, ny):
    q0 = configs.taylor_green2(3, nx, ny, A=0.3)
    dt = 0.4 * (1 / nx) / (7 * configs.max_wavespeed2(q0, 1.4))
    g = orc.Grid2D(3, nx, ny)
    ref, _ = g.run_steps(1.4, 2e-3, 60, orc.PCG, 2, dt, q0)
    for tile, htile in ((404, 404), (804, 804), (808, 804), (808, 404), (202, 202), (404, 804)):
        c = ctx2(3, nx, ny)
        c.set_option(igr.core.OPT_TILE, tile)
        c.set_option(igr.core.OPT_HTILE, htile)
        assert (c.get_option(igr.core.OPT_TILE), c.get_option(igr.core.OPT_HTILE)) == (tile, htile)
        c.set_state(q0)
        c.run_steps(2, dt, 1.4, 2e-3, 60, "pcg")
        assert rel(c.get_state()[0], ref) < 1e-11, (tile, htile)
    # the PCG x/r update on fewer workgroups than tiles, each running several
    # (uneven counts: 36 or 15 tiles of 4x4 on 7 workgroups)
    for groups in (7, 3):
        c = ctx2(3, nx, ny)
        c.set_option(igr.core.OPT_TILE, 404)
        c.set_option(igr.core.OPT_GROUPS, groups)
        assert c.get_option(igr.core.OPT_GROUPS) == groups
        for solver in ("pcg", "jacobi"):
            c.set_state(q0)
            c.run_steps(2, dt, 1.4, 2e-3, 60, solver)
            want = ref if solver == "pcg" else g.run_steps(1.4, 2e-3, 60, orc.JACOBI, 2, dt, q0)[0]
            assert rel(c.get_state()[0], want) < 1e-11, (groups, solver)
    c = ctx2(3, nx, ny)
    with pytest.raises(ValueError):
        c.set_option(igr.core.OPT_GROUPS, -1)
    with pytest.raises(ValueError):
        c.set_option(igr.core.OPT_HTILE, 808)     # no 8x8 hyperbolic tile
    with pytest.raises(ValueError):
        c.set_option(igr.core.OPT_TILE, 303)


def test_autotune_p4_keeps_results(orc):
    # p = 4: the autotuner times the four SIP-family shapes (2x4, 4x4, 5x5,
    # 8x5) and keeps the fastest; the run after it still matches the oracle
    q0 = configs.taylor_green2(4, 12, 10, A=0.3)
    dt = 0.4 * (1 / 12) / (9 * configs.max_wavespeed2(q0, 1.4))
    c = ctx2(4, 12, 10)
    c.set_state(q0)
    tile, htile = c.autotune()
    print(f"[autotune] p=4 12x10 -> SIP tile {tile}", flush=True)
    assert tile in (204, 404, 505, 805)
    c.set_state(q0)
    c.run_steps(2, dt, 1.4, 2e-3, 60, "pcg")
    ref, _ = orc.Grid2D(4, 12, 10).run_steps(1.4, 2e-3, 60, orc.PCG, 2, dt, q0)
    assert rel(c.get_state()[0], ref) < 1e-11


def test_autotune_keeps_results(orc):
    q0 = configs.taylor_green2(3, 32, 32, A=0.3)
    dt = 0.4 * (1 / 32) / (7 * configs.max_wavespeed2(q0, 1.4))
    c = ctx2(3, 32, 32)
    c.set_state(q0)
    tile, htile = c.autotune()
    print(f"[autotune] 32x32 -> SIP tile {tile}, hyperbolic tile {htile}", flush=True)
    assert tile in (404, 804, 808) and htile in (404, 804)
    c.set_state(q0)
    c.run_steps(2, dt, 1.4, 2e-3, 60, "pcg")
    ref, _ = orc.Grid2D(3, 32, 32).run_steps(1.4, 2e-3, 60, orc.PCG, 2, dt, q0)
    assert rel(c.get_state()[0], ref) < 1e-11


def test_c2_full_size_one_stage(orc):
    # 128x128 p=3 (the measured workload): one compute_igr_euler_rhs! (PCG-60)
    cfg = configs.config2(N=128)
    c = ctx2(3, 128, 128)
    g = orc.Grid2D(3, 128, 128)
    c.set_state(cfg["q"], np.zeros(g.shape))
    dq = c.euler_rhs(cfg["gamma"], cfg["alpha"], 60, "pcg")
    dq_ref, S_ref = g.euler_rhs(cfg["gamma"], cfg["alpha"], 60, orc.PCG, cfg["q"], g.field())
    assert rel(c.get_state()[1], S_ref) < 1e-12
    assert rel(dq, dq_ref) < 1e-12


def test_c2_full_run_properties():
    # full C2 (200 steps): positivity, conservation of mass/energy/momentum to roundoff
    cfg = configs.config2(N=128)
    c = ctx2(3, 128, 128)
    c.set_state(cfg["q"])
    c.run_steps(cfg["n_steps"], cfg["dt"], cfg["gamma"], cfg["alpha"], 60, "pcg")
    q, S = c.get_state()
    assert np.all(np.isfinite(q)) and np.all(np.isfinite(S))
    wq = configs.quad2(3, 128, 128)
    for v in range(4):
        m0, m1 = np.sum(wq * cfg["q"][v]), np.sum(wq * q[v])
        assert abs(m1 - m0) <= 1e-12 * max(1.0, np.sum(wq * np.abs(q[v])))


# ------------------------------------------------------------------ 1D ----
@pytest.mark.parametrize("ne", [1, 2, 7, 128, 300])
def test_apply_sip_and_hyp_1d(orc, ne):
    c = ctx1(3, ne)
    g = orc.Grid1D(3, ne)
    rng = np.random.default_rng(ne)
    rho = 1.0 + 0.5 * rng.random(g.shape)
    x = rng.standard_normal(g.shape)
    assert rel(c.apply_sip(0.02, x, rho), g.apply_sip(0.02, x, rho)) < 1e-13
    q = configs.smooth_random1(3, ne, seed=ne, amp=0.3)
    S = 0.01 * rng.standard_normal(g.shape)
    np.testing.assert_array_equal(c.hyperbolic_rhs(1.4, q, S), g.hyperbolic_rhs(1.4, q, S))


@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
def test_solve_and_run_1d(orc, solver):
    c = ctx1(3, 16)
    g = orc.Grid1D(3, 16)
    code = igr.SOLVERS[solver]
    q0 = configs.smooth_random1(3, 16, seed=3, amp=0.3)
    S0 = 0.01 * np.sin(2 * np.pi * g.coords())
    c.set_state(q0, S0)
    c.solve_elliptic(0.01, 15, solver)
    assert rel(c.get_state()[1], g.solve_elliptic(0.01, 15, code, q0, S0)) < 1e-12
    dt = 0.1 / 16 / (7 * configs.max_wavespeed1(q0, 1.4))
    c.set_state(q0)
    c.run_steps(1, dt, 1.4, 0.01, 15, solver)
    q1r, _ = g.run_steps(1.4, 0.01, 15, code, 1, dt, q0)
    assert rel(c.get_state()[0], q1r) < 1e-12
    c.set_state(q0)
    ns = c.run_forward(dt, 40 * dt, 1.4, 0.01, 15, solver)
    qT, ST = c.get_state()
    qTr, STr, nsr = g.run_forward(1.4, 0.01, 15, code, dt, 40 * dt, q0)
    assert ns == nsr
    assert rel(qT, qTr) < 1e-9 and rel(ST, STr) < 1e-9


@pytest.mark.parametrize("p,ne", [(3, 128), (3, 64), (1, 128), (2, 64), (4, 64)])
@pytest.mark.parametrize("solver", ["pcg", "jacobi", "chebyshev"])
def test_wave1d_run_matches_oracle_and_workgroup_kernel(orc, p, ne, solver):
    """Plain 1D runs of N_e = 64 (and 128 up to p = 3) take the one-wave
    kernel (IGR_OPT_WAVE1D): 12 steps (and a run_forward with its final
    solve) against the oracle and against the one-workgroup kernel, with
    alpha > 0 and alpha = 0.  PCG runs 200 iterations: 20 unconverged PCG
    iterations on these systems move by up to 5e-9 when only the summation
    order of the dot products changes (a dense restatement, summed forwards
    and backwards), and 36 warm-started solves carry that into Sigma at
    1e-5; converged, the same change moves them by <= 1e-13."""
    g = orc.Grid1D(p, ne)
    code = igr.SOLVERS[solver]
    n_iter = 200 if solver == "pcg" else 20
    q0 = configs.smooth_random1(p, ne, seed=40 + p, amp=0.3)
    dt = 0.2 / ne / ((2 * p + 1) * configs.max_wavespeed1(q0, 1.4))
    for alpha in (10.0 * (1.0 / ne) ** 2, 0.0):
        out = {}
        for wave in (1, 0):
            c = ctx1(p, ne)
            c.set_option(igr.core.OPT_WAVE1D, wave)
            assert c.get_option(igr.core.OPT_WAVE1D) == wave
            c.set_state(q0)
            c.run_steps(12, dt, 1.4, alpha, n_iter, solver)
            q, S = c.get_state()
            c.set_state(q0)
            ns = c.run_forward(dt, 7.5 * dt, 1.4, alpha, n_iter, solver)
            qf, Sf = c.get_state()
            out[wave] = (q, S, qf, Sf, ns)
            c.close()
        qr, Sr = g.run_steps(1.4, alpha, n_iter, code, 12, dt, q0)
        qfr, Sfr, nsr = g.run_forward(1.4, alpha, n_iter, code, dt, 7.5 * dt, q0)
        for wave in (1, 0):
            q, S, qf, Sf, ns = out[wave]
            assert ns == nsr
            assert rel(q, qr) < 1e-11 and rel(qf, qfr) < 1e-11, (wave, alpha)
            if alpha > 0:
                assert rel(S, Sr) < 1e-10 and rel(Sf, Sfr) < 1e-10, (wave, alpha)
            else:
                assert np.all(S == 0.0) and np.all(Sf == 0.0)
        assert rel(out[1][0], out[0][0]) < 1e-12 and rel(out[1][2], out[0][2]) < 1e-12


def test_config1_run_matches_oracle(orc):
    # C1 (1D, 128 elements, Jacobi-50), first 300 steps compared with the oracle
    cfg = configs.config1()
    c = ctx1(3, 128)
    g = orc.Grid1D(3, 128)
    c.set_state(cfg["q"])
    c.run_steps(300, cfg["dt"], cfg["gamma"], cfg["alpha"], 50, "jacobi")
    qr, Sr = g.run_steps(cfg["gamma"], cfg["alpha"], 50, orc.JACOBI, 300, cfg["dt"], cfg["q"])
    q, S = c.get_state()
    assert rel(q, qr) < 1e-9 and rel(S, Sr) < 1e-9


# ------------------------------------------------------------- errors -----
def test_errors():
    c = ctx2(3, 4, 4)
    q = configs.smooth_random2(3, 4, 4, seed=1)
    c.set_state(q)
    with pytest.raises(ValueError):
        c.solve_elliptic(0.01, 5, 7)
    with pytest.raises(ValueError):
        c.run_steps(1, 1e-3, 1.4, 0.01, 5, "cg")
    q[3] = -1.0
    with pytest.raises(igr.DomainError):
        c.hyperbolic_rhs(1.4, q, np.zeros(q.shape[1:]))
    c1 = ctx1(3, 8)
    q1 = configs.smooth_random1(3, 8, seed=2)
    q1[2] = -1.0
    with pytest.raises(igr.DomainError):
        c1.hyperbolic_rhs(1.4, q1, np.zeros(q1.shape[1:]))
    # arrays of the wrong size are refused before the C call
    with pytest.raises(ValueError):
        c.set_state(q[:, :2])
    with pytest.raises(ValueError):
        c.apply_sip(0.01, np.zeros(7), np.ones(q.shape[1:]))
    with pytest.raises(ValueError):
        c.objective(3, 1.4, prm=np.zeros(6))


def test_mirror_api_run_forward(orc):
    basis = igr.DGBasis(3)
    mesh = igr.CartesianMesh2D(6, 6, 1.0, 1.0, basis)
    q0 = configs.taylor_green2(3, 6, 6, A=0.2)
    q = q0.copy()
    S = np.zeros(mesh.shape)
    igr.run_forward_(q, S, 2e-3, 0.02, basis, mesh, 1.4, 3e-3, 10, ":pcg")
    g = orc.Grid2D(3, 6, 6)
    qr, Sr, _ = g.run_forward(1.4, 3e-3, 10, orc.PCG, 2e-3, 0.02, q0)
    # q to 1e-10; Sigma (an under-converged PCG-10 result) to 1e-7
    assert rel(q, qr) < 1e-10 and rel(S, Sr) < 1e-7
    with pytest.raises(ValueError):
        igr.run_forward_(q, S, 2e-3, 0.02, basis, mesh, 1.4, 3e-3, 10, ":gmres")


@pytest.mark.parametrize("alpha_on", [False, True])
def test_quasi1d_gpu_2d_equals_gpu_1d(alpha_on):
    """A y-uniform 2D run reproduces the 1D run on the GPU path
    (IGR_adjoints_2d/tests/run_quasi1d_test.jl:27-98: deviation <= 1e-9,
    y-nonuniformity <= 1e-11, PCG-400 when alpha > 0)."""
    print(f"[case] quasi-1D GPU alpha_on={alpha_on}", flush=True)
    gam, L, p, Ne, A, T, CFL = 1.4, 1.0, 3, 12, 0.4, 0.08, 0.5
    Ney, Ly = 4, 0.5
    basis = igr.DGBasis(p)
    c1 = igr.Context(igr.PeriodicMesh1D(Ne, L, basis))
    c2 = igr.Context(igr.CartesianMesh2D(Ne, Ney, L, Ly, basis))
    x1 = configs.coords1(p, Ne, L)
    u1 = A * np.sin(2 * np.pi * x1 / L)
    q1 = np.stack([u1, np.ones_like(u1), 1.0 / (gam - 1) + 0.5 * u1 ** 2])
    x2, _ = configs.coords2(p, Ne, Ney, L, Ly)
    u2 = A * np.sin(2 * np.pi * x2 / L)
    q2 = np.stack([u2, np.zeros_like(u2), np.ones_like(u2), 1.0 / (gam - 1) + 0.5 * u2 ** 2])
    dt = CFL * (L / Ne) / ((2 * p + 1) * configs.max_wavespeed1(q1, gam))
    alpha = 3.0 * (L / Ne) ** 2 if alpha_on else 0.0
    n_iter = 400 if alpha_on else 1
    c1.set_state(q1)
    c2.set_state(q2)
    n1 = c1.run_forward(dt, T, gam, alpha, n_iter, "pcg")
    n2 = c2.run_forward(dt, T, gam, alpha, n_iter, "pcg")
    assert n1 == n2
    f1, S1 = c1.get_state()
    f2, S2 = c2.get_state()
    pairs = [(f2[0], f1[0]), (f2[2], f1[1]), (f2[3], f1[2])] + ([(S2, S1)] if alpha_on else [])
    for a2, a1 in pairs:   # a2 (Ney, Nex, j, i), a1 (Ne, i)
        dev = np.max(np.abs(a2 - a1[None, :, None, :]))
        ynu = np.max(np.abs(a2 - a2[:1, :, :1, :]))
        assert dev < 1e-9 and ynu < 1e-11, (dev, ynu)
    assert np.max(np.abs(f2[1])) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 4])
def test_other_degrees_step_run_2d(orc, p):
    """p = 1, 2, 4 (the reference's 2D heat-map runs use p = 4): one step to
    1e-12 and a short run to 1e-9 against the oracle, every solver."""
    nx, ny = 6, 5
    q0 = configs.smooth_random2(p, nx, ny, seed=40 + p, amp=0.3)
    dt = 0.3 * (1 / nx) / ((2 * p + 1) * configs.max_wavespeed2(q0, 1.4))
    g = orc.Grid2D(p, nx, ny)
    for solver in ("pcg", "jacobi", "chebyshev"):
        print(f"[case] 2D p={p} {solver}", flush=True)
        c = ctx2(p, nx, ny)
        code = igr.SOLVERS[solver]
        c.set_state(q0)
        c.run_steps(1, dt, 1.4, 3e-3, 20, solver)
        q1r, _ = g.run_steps(1.4, 3e-3, 20, code, 1, dt, q0)
        assert rel(c.get_state()[0], q1r) < 1e-12
        c.set_state(q0)
        ns = c.run_forward(dt, 12 * dt, 1.4, 3e-3, 20, solver)
        qT, ST = c.get_state()
        qTr, STr, nsr = g.run_forward(1.4, 3e-3, 20, code, dt, 12 * dt, q0)
        assert ns == nsr and rel(qT, qTr) < 1e-9 and rel(ST, STr) < 1e-9


@pytest.mark.parametrize("p", [1, 2, 4])
def test_other_degrees_run_1d(orc, p):
    """Short 1D runs at p = 1, 2, 4 against the oracle.  The bound is 1e-9 or
    ten times the oracle's own response to a 1e-14 relative change of rho,
    whichever is larger: the 20-iteration PCG at p = 4 is that sensitive
    (a 1e-14 perturbation moves Sigma by ~1e-7), so no summation order can
    match it to 1e-9."""
    ne = 12
    q0 = configs.smooth_random1(p, ne, seed=50 + p, amp=0.3)
    dt = 0.2 / ne / ((2 * p + 1) * configs.max_wavespeed1(q0, 1.4))
    g = orc.Grid1D(p, ne)
    q1 = q0.copy()
    q1[1] *= 1.0 + 1e-14
    for solver in ("pcg", "jacobi", "chebyshev"):
        print(f"[case] 1D p={p} {solver}", flush=True)
        c = ctx1(p, ne)
        code = igr.SOLVERS[solver]
        c.set_state(q0)
        ns = c.run_forward(dt, 15 * dt, 1.4, 5e-3, 20, solver)
        qT, ST = c.get_state()
        qTr, STr, nsr = g.run_forward(1.4, 5e-3, 20, code, dt, 15 * dt, q0)
        qTp, STp, _ = g.run_forward(1.4, 5e-3, 20, code, dt, 15 * dt, q1)
        tq = max(1e-9, 10 * rel(qTp, qTr))
        tS = max(1e-9, 10 * rel(STp, STr))
        assert ns == nsr and rel(qT, qTr) < tq and rel(ST, STr) < tS, (solver, rel(qT, qTr), tq, rel(ST, STr), tS)


def test_p4_tile_shapes_agree(orc):
    """p = 4 (25 nodes per element): every SIP-family tile shape (2x4, 4x4,
    5x5, 8x5: 200 to 1000 threads) against the oracle, on a mesh that none of
    them divides evenly."""
    nx, ny = 11, 9
    q0 = configs.smooth_random2(4, nx, ny, seed=150, amp=0.3)
    dt = 0.3 * (1 / nx) / (9 * configs.max_wavespeed2(q0, 1.4))
    g = orc.Grid2D(4, nx, ny)
    ref, Sref = g.run_steps(1.4, 3e-3, 40, orc.PCG, 2, dt, q0)
    for tile in (204, 404, 505, 805):
        c = ctx2(4, nx, ny)
        c.set_option(igr.core.OPT_TILE, tile)
        assert c.get_option(igr.core.OPT_TILE) == tile
        c.set_state(q0)
        c.run_steps(2, dt, 1.4, 3e-3, 40, "pcg")
        q, S = c.get_state()
        assert rel(q, ref) < 1e-11, tile
        assert rel(S, Sref) < 1e-10, tile
    c = ctx2(4, nx, ny)
    with pytest.raises(ValueError):
        c.set_option(igr.core.OPT_TILE, 808)
    with pytest.raises(ValueError):
        c.set_option(igr.core.OPT_HTILE, 404)
