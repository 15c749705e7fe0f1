"""Pins the CPU oracle (the restated reference) to closed forms, invariants,
dense-matrix solves and finite differences before anything is compared
against it.  Restates the acceptance checks of the reference's own tests
(IGR_adjoints_1d/tests/run_euler_test.jl, IGR_adjoints_2d/tests/
run_euler_test.jl, run_quasi1d_test.jl, run_enzyme_test.jl)."""
import numpy as np
import pytest

from igr_sensitivities_amd import configs


# --------------------------------------------------------------- basis ----
def test_basis_closed_form_p3(orc):
    xi, w, D = orc.basis(3)
    s = 1.0 / np.sqrt(5.0)
    np.testing.assert_allclose(xi, [-1.0, -s, s, 1.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], rtol=0, atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_basis_weights_and_exactness(orc, p):
    # run_euler_test.jl:15-41: sum w = 2 (1e-14); D exact on monomials of degree <= p (1e-12)
    xi, w, D = orc.basis(p)
    assert abs(w.sum() - 2.0) < 1e-14
    for k in range(p + 1):
        f = xi ** k
        df = k * xi ** (k - 1) if k > 0 else np.zeros_like(xi)
        assert np.max(np.abs(D @ f - df)) < 1e-12
    # quadrature exact to degree 2p-1
    for k in range(2 * p):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, xi ** k) - exact) < 1e-13
    # zero row sums (free-stream preservation, basis.jl:44-45)
    assert np.max(np.abs(D.sum(axis=1))) < 1e-13


def test_basis_matches_setup_copy(orc):
    for p in range(1, 5):
        xi, w, _ = orc.basis(p)
        xi2, w2 = configs.lgl(p)
        np.testing.assert_array_equal(xi, xi2)
        np.testing.assert_array_equal(w, w2)


# ---------------------------------------------------------- hyperbolic ----
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_freestream_2d(orc, p):
    # IGR_adjoints_2d/tests/run_euler_test.jl:15-33 (6x5 mesh, Lx != Ly, 1e-11)
    g = orc.Grid2D(p, 6, 5, 1.0, 1.3)
    gam = 1.4
    q = g.state()
    q[0], q[1], q[2] = 1.0, -0.7, 1.0
    q[3] = 1.0 / (gam - 1) + 0.5 * (1.0 + 0.49)
    dq = g.hyperbolic_rhs(gam, q, g.field())
    assert np.max(np.abs(dq)) < 1e-11


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_freestream_1d(orc, p):
    # IGR_adjoints_1d/tests/run_euler_test.jl:46-72 (1e-12)
    g = orc.Grid1D(p, 10)
    gam = 1.4
    q = g.state()
    q[0], q[1], q[2] = 0.8, 1.0, 1.0 / (gam - 1) + 0.5 * 0.64
    assert np.max(np.abs(g.hyperbolic_rhs(gam, q, g.field()))) < 1e-12


def test_conservation_2d(orc):
    # periodic strong-form DG with LLF: sum_nodes w w Jx Jy dq = 0
    g = orc.Grid2D(3, 5, 4, 1.0, 0.8)
    q = configs.smooth_random2(3, 5, 4, seed=4)
    x, y = g.coords()
    q = configs.smooth_random2(3, 5, 4, seed=4)
    S = 0.01 * np.sin(2 * np.pi * x)
    dq = g.hyperbolic_rhs(1.4, q, S)
    wq = g.quad_weights()
    for v in range(4):
        tot = np.sum(wq * dq[v])
        assert abs(tot) < 1e-12 * max(1.0, np.sum(wq * np.abs(dq[v])))


def test_conservation_1d(orc):
    g = orc.Grid1D(3, 9)
    q = configs.smooth_random1(3, 9, seed=5)
    dq = g.hyperbolic_rhs(1.4, q, 0.01 * np.cos(2 * np.pi * g.coords()))
    wq = g.quad_weights()
    for v in range(3):
        assert abs(np.sum(wq * dq[v])) < 1e-12 * max(1.0, np.sum(wq * np.abs(dq[v])))


def test_entropy_wave_order_1d(orc):
    # run_euler_test.jl:77-122: density wave advected one period, alpha = 0, order ~ p+1
    errs = []
    gam, p = 1.4, 3
    for Ne in (8, 16):
        g = orc.Grid1D(p, Ne)
        x = g.coords()
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * x)
        q = g.state()
        q[1] = rho
        q[0] = rho * 1.0
        q[2] = 1.0 / (gam - 1) + 0.5 * rho
        dt = 0.3 * g.dx / ((2 * p + 1) * (1.0 + np.sqrt(gam / 0.8)))
        qT, _, _ = g.run_forward(gam, 0.0, 1, orc.PCG, dt, 1.0, q)
        errs.append(np.sqrt(np.sum(g.quad_weights() * (qT[1] - rho) ** 2)))
    order = np.log2(errs[0] / errs[1])
    assert order > p + 0.5


# ------------------------------------------------------------- elliptic ---
def _dense(apply, shape):
    N = int(np.prod(shape))
    A = np.zeros((N, N))
    for k in range(N):
        e = np.zeros(N)
        e[k] = 1.0
        A[:, k] = apply(e.reshape(shape)).ravel()
    return A


@pytest.mark.parametrize("mesh", [(3, 4, 1.0, 0.7), (2, 2, 1.0, 1.0), (1, 3, 0.5, 1.0)])
def test_sip_symmetric_positive_and_diag(orc, mesh):
    nx, ny, lx, ly = mesh
    g = orc.Grid2D(3, nx, ny, lx, ly)
    rng = np.random.default_rng(7)
    rho = 1.0 + 0.4 * rng.random(g.shape)
    a = 0.013
    A = _dense(lambda e: g.apply_sip(a, e, rho), g.shape)
    assert np.max(np.abs(A - A.T)) < 1e-13 * np.max(np.abs(A))
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    if nx >= 2 and ny >= 2:
        # with a single element across a periodic direction the face terms
        # couple nodes of that same element, so the reference's analytic
        # diagonal (elliptic.jl:237-261) is no longer diag(A)
        np.testing.assert_allclose(np.diag(A), g.sip_diag(a, rho).ravel(), rtol=1e-13, atol=1e-16)


def test_sip_alpha0_is_mass(orc):
    g = orc.Grid2D(3, 3, 2, 1.0, 1.0)
    rng = np.random.default_rng(8)
    rho = 1.0 + rng.random(g.shape)
    x = rng.standard_normal(g.shape)
    y = g.apply_sip(0.0, x, rho)
    np.testing.assert_allclose(y, g.quad_weights() * x / rho, rtol=1e-15)


def test_sip_1d_symmetric_and_diag(orc):
    g = orc.Grid1D(3, 7)
    rng = np.random.default_rng(9)
    rho = 1.0 + 0.4 * rng.random(g.shape)
    a = 0.02
    A = _dense(lambda e: g.apply_sip(a, e, rho), g.shape)
    assert np.max(np.abs(A - A.T)) < 1e-13 * np.max(np.abs(A))
    assert np.linalg.eigvalsh(A).min() > 0
    np.testing.assert_allclose(np.diag(A), g.sip_diag(a, rho).ravel(), rtol=1e-13)


def test_sip_manufactured_convergence(orc):
    # run_euler_test.jl:38-79 (2D): rho = 1, Sigma = sin sin, b = (1 + alpha k^2) M Sigma
    a, p = 0.1, 3
    errs = []
    for N in (4, 8):
        g = orc.Grid2D(p, N, N)
        x, y = g.coords()
        Se = np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
        k2 = 2 * (2 * np.pi) ** 2
        b = g.quad_weights() * (1 + a * k2) * Se
        rho = np.ones(g.shape)
        M = g.sip_diag(a, rho)
        S = g.solver(a, orc.PCG, 800, rho, b, M, g.field())
        errs.append(np.sqrt(np.sum(g.quad_weights() * (S - Se) ** 2)))
    assert errs[1] < errs[0] / 2 ** (p + 0.5)


def test_solvers_against_dense_restatement(orc):
    g = orc.Grid2D(3, 3, 3)
    rng = np.random.default_rng(10)
    rho = 1.0 + 0.3 * rng.random(g.shape)
    a = 0.01
    A = _dense(lambda e: g.apply_sip(a, e, rho), g.shape)
    Md = g.sip_diag(a, rho).ravel()
    b = rng.standard_normal(g.N)
    x0 = rng.standard_normal(g.N)
    # PCG converged == dense solve
    xs = g.solver(a, orc.PCG, 300, rho, b.reshape(g.shape), Md.reshape(g.shape), x0.reshape(g.shape))
    np.testing.assert_allclose(xs.ravel(), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # PCG scalars vs a dense restatement of the same recurrences
    x, ta, tb = g.solver(a, orc.PCG, 12, rho, b.reshape(g.shape), Md.reshape(g.shape), x0.reshape(g.shape),
                         trace=True)
    r = b - A @ x0
    z = r / Md
    pv = z.copy()
    rz = r @ z
    xd = x0.copy()
    for it in range(12):
        Ap = A @ pv
        al = rz / (pv @ Ap + 1e-30)
        xd += al * pv
        r -= al * Ap
        z = r / Md
        rzn = r @ z
        be = rzn / (rz + 1e-30)
        pv = z + be * pv
        rz = rzn
        assert abs(ta[it] - al) <= 1e-10 * abs(al)
        assert abs(tb[it] - be) <= 1e-9 * abs(be) + 1e-300
    np.testing.assert_allclose(x.ravel(), xd, rtol=1e-10, atol=1e-12)
    # Jacobi: x <- x + (2/3)(b - A x)/M
    xj = g.solver(a, orc.JACOBI, 9, rho, b.reshape(g.shape), Md.reshape(g.shape), x0.reshape(g.shape))
    xd = x0.copy()
    for _ in range(9):
        xd = xd + (2.0 / 3.0) * (b - A @ xd) / Md
    np.testing.assert_allclose(xj.ravel(), xd, rtol=1e-12, atol=1e-13)
    # Chebyshev three-term recurrence with lambda_max = 2.1 at p = 3
    xc = g.solver(a, orc.CHEBYSHEV, 7, rho, b.reshape(g.shape), Md.reshape(g.shape), x0.reshape(g.shape))
    lmin, lmax = 1.0, 2.1
    th, de = (lmax + lmin) / 2, (lmax - lmin) / 2
    sg = th / de
    xd = x0.copy()
    r = b - A @ xd
    rc = 1 / sg
    d = None
    for k in range(1, 8):
        z = r / Md
        if k == 1:
            d = z / th
        else:
            rn = 1 / (2 * sg - rc)
            d = rn * rc * d + (2 * rn / de) * z
            rc = rn
        xd = xd + d
        r = b - A @ xd
    np.testing.assert_allclose(xc.ravel(), xd, rtol=1e-12, atol=1e-13)


def test_igr_rhs_closed_form(orc):
    # u = A sin(2 pi x): R = 2 alpha u_x^2 in 1D; 2D: (div u)^2 + tr((Du)^2) = 2 u_x^2 for u = u(x)
    g2 = orc.Grid2D(4, 6, 3)
    x, _ = g2.coords()
    q = g2.state()
    q[2] = 1.0
    q[0] = 0.3 * np.sin(2 * np.pi * x)
    q[3] = 2.5
    a = 0.05
    b = g2.igr_rhs(a, q)
    ux = 0.3 * 2 * np.pi * np.cos(2 * np.pi * x)
    ref = g2.quad_weights() * a * 2 * ux ** 2
    assert np.max(np.abs(b - ref)) < 2e-3 * np.max(np.abs(ref))
    g1 = orc.Grid1D(4, 6)
    x1 = g1.coords()
    q1 = g1.state()
    q1[1] = 1.0
    q1[0] = 0.3 * np.sin(2 * np.pi * x1)
    q1[2] = 2.5
    b1 = g1.igr_rhs(a, q1)
    ref1 = g1.quad_weights() * 2 * a * (0.3 * 2 * np.pi * np.cos(2 * np.pi * x1)) ** 2
    assert np.max(np.abs(b1 - ref1)) < 2e-3 * np.max(np.abs(ref1))


# ------------------------------------------------------------ quasi-1D ----
def _quasi_setup(orc, alpha_on):
    # IGR_adjoints_2d/tests/run_quasi1d_test.jl:27-74
    gam, L, p, Ne, A, T, CFL = 1.4, 1.0, 3, 12, 0.4, 0.08, 0.5
    Ney, Ly = 4, 0.5
    g1 = orc.Grid1D(p, Ne, L)
    g2 = orc.Grid2D(p, Ne, Ney, L, Ly)
    x1 = g1.coords()
    u1 = A * np.sin(2 * np.pi * x1 / L)
    q1 = g1.state()
    q1[0], q1[1], q1[2] = u1, 1.0, 1.0 / (gam - 1) + 0.5 * u1 ** 2
    x2, _ = g2.coords()
    u2 = A * np.sin(2 * np.pi * x2 / L)
    q2 = g2.state()
    q2[0], q2[2], q2[3] = u2, 1.0, 1.0 / (gam - 1) + 0.5 * u2 ** 2
    ws = np.max(np.abs(u1) + np.sqrt(gam * 1.0))
    ws = configs.max_wavespeed1(q1, gam)
    dt = CFL * g1.dx / ((2 * p + 1) * ws)
    n = int(np.ceil(T / dt))
    dt = T / n
    alpha = 3.0 * (L / Ne) ** 2 if alpha_on else 0.0
    return g1, g2, q1, q2, dt, T, gam, alpha


@pytest.mark.parametrize("alpha_on", [False, True])
def test_quasi1d(orc, alpha_on):
    g1, g2, q1, q2, dt, T, gam, alpha = _quasi_setup(orc, alpha_on)
    n_iter = 400 if alpha_on else 1
    f1, S1, _ = g1.run_forward(gam, alpha, n_iter, orc.PCG, dt, T, q1)
    f2, S2, _ = g2.run_forward(gam, alpha, n_iter, orc.PCG, dt, T, q2)
    pairs = [(f2[0], f1[0]), (f2[2], f1[1]), (f2[3], f1[2])] + ([(S2, S1)] if alpha_on else [])
    for a2, a1 in pairs:   # a2 (Ney, Nex, j, i), a1 (Ne, i)
        dev = np.max(np.abs(a2 - a1[None, :, None, :]))
        ynu = np.max(np.abs(a2 - a2[:1, :, :1, :]))
        assert dev < 1e-9 and ynu < 1e-11, (dev, ynu)
    assert np.max(np.abs(f2[1])) < 1e-12


# ------------------------------------------------- JVP machinery vs FD ----
def test_dual_jvp_matches_central_fd_2d(orc):
    g = orc.Grid2D(3, 3, 3)
    q0 = configs.smooth_random2(3, 3, 3, seed=11, amp=0.2)
    rng = np.random.default_rng(12)
    dq = rng.standard_normal(q0.shape)
    gam, al, dt, ns = 1.4, 2e-3, 2e-3, 3
    prm = (0.55, 0.5, 0.15, 0.0)
    for obj in (orc.OBJ_L2_DENSITY, orc.OBJ_WINDOWED_KE, orc.OBJ_KINETIC_ENERGY):
        J, dJ, _, _ = g.jvp_objective_run(gam, 0.3, al, 0.5 * al, 30, orc.PCG, ns, dt, obj, q0, dq, prm)
        h = 1e-6

        def Jat(t):
            qq, _ = g.run_steps(gam + 0.3 * t, al + 0.5 * al * t, 30, orc.PCG, ns, dt, q0 + t * dq)
            return g.objective(obj, gam + 0.3 * t, qq, prm)
        fd = (Jat(h) - Jat(-h)) / (2 * h)
        assert abs(J - Jat(0.0)) <= 1e-14 * abs(J)
        assert abs(dJ - fd) <= 1e-6 * max(abs(fd), 1e-12), (obj, dJ, fd)


def test_dual_jvp_matches_central_fd_1d(orc):
    g = orc.Grid1D(3, 8)
    q0 = configs.smooth_random1(3, 8, seed=13, amp=0.2)
    dq = np.random.default_rng(14).standard_normal(q0.shape)
    gam, al, dt, ns = 1.4, 0.01, 2e-3, 4
    for solver in (orc.PCG, orc.JACOBI, orc.CHEBYSHEV):
        J, dJ, _, _ = g.jvp_objective_run(gam, 0.0, al, 0.0, 10, solver, ns, dt, orc.OBJ_L2_PRESSURE, q0, dq)
        h = 1e-6

        def Jat(t):
            qq, _ = g.run_steps(gam, al, 10, solver, ns, dt, q0 + t * dq)
            return g.objective(orc.OBJ_L2_PRESSURE, gam, qq)
        fd = (Jat(h) - Jat(-h)) / (2 * h)
        assert abs(dJ - fd) <= 1e-6 * abs(fd)


# ----------------------------------------------------------- objectives ---
def test_objective_closed_forms(orc):
    g = orc.Grid2D(3, 4, 3, 1.0, 0.5)
    q = g.state()
    q[2] = 2.0
    q[0] = 0.5
    q[3] = 3.0
    area = 0.5
    assert abs(g.objective(orc.OBJ_L2_DENSITY, 1.4, q) - 4.0 * area) < 1e-14
    P = 0.4 * (3.0 - 0.25 / 4.0)
    assert abs(g.objective(orc.OBJ_L2_PRESSURE, 1.4, q) - P * P * area) < 1e-13
    assert abs(g.objective(orc.OBJ_KINETIC_ENERGY, 1.4, q) - (0.25 / 4.0) / 2.0) < 1e-14
    # weighted momentum of a constant field integrates sin*sin to zero
    assert abs(g.objective(orc.OBJ_WEIGHTED_MOMENTUM, 1.4, q)) < 1e-14


# ------------------------------------------------------------ errors ------
def test_unknown_solver_and_domain_error(orc):
    g = orc.Grid2D(3, 2, 2)
    q = configs.smooth_random2(3, 2, 2, seed=1)
    with pytest.raises(orc.OracleError):
        g.solve_elliptic(0.01, 5, 7, q, g.field())
    q[3] = -1.0
    with pytest.raises(orc.DomainError):
        g.hyperbolic_rhs(1.4, q, g.field())
