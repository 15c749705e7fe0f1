"""Pins the oracle's fixed-iteration solvers and its 2D y-direction flux by
mathematics, not by a second copy of their recurrences.

- damped Jacobi (IGR_adjoints_2d/elliptic.jl:354-361, omega = 2/3 passed at
  :420) is the linear map e -> (I - omega M^-1 A) e on the error;
- Chebyshev-accelerated Jacobi (elliptic.jl:370-397) is the Chebyshev
  semi-iteration of Saad, "Iterative Methods for Sparse Linear Systems",
  2nd ed., Alg. 12.1: after k steps the error is P_k(M^-1 A) e_0 with the
  residual polynomial P_k(t) = T_k((theta - t)/delta) / T_k(theta/delta),
  theta, delta from [lambda_min, lambda_max] = [1, (1.5, 2.0, 2.1, 2.2)[p]]
  (elliptic.jl:375).  T_k is evaluated by numpy's Chebyshev series (Clenshaw)
  or by cosh(k arccosh s), never by the three-term recurrence the oracle uses;
- at alpha = 0 the SIP operator is the lumped mass w_i w_j Jx Jy / rho, its
  analytic diagonal M equals it, and M^-1 A = I: Jacobi then contracts the
  error by exactly (1 - omega)^k = 3^-k and Chebyshev by 1 / T_k(theta/delta);
- the isentropic vortex (IGR_adjoints_2d/tests/run_euler_test.jl:84-119) is
  advected diagonally for one period and converges at order ~ p + 1, which
  pins the y-direction flux and its LLF face terms.

Each check is shown to reject a planted mistake (oracle.fwd_mutation):
omega = 0.5, lambda_max = 2.2 at p = 3, and flux_y with its momentum
components swapped.
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from igr_sensitivities_amd import configs

LMAX = {1: 1.5, 2: 2.0, 3: 2.1, 4: 2.2}


def _cheb_T(k, s):
    """T_k(s) by Clenshaw summation of the series with one unit coefficient."""
    return npcheb.chebval(s, [0.0] * k + [1.0])


def _sigma(p):
    lmax = LMAX[min(p, 4)]
    theta, delta = (lmax + 1.0) / 2.0, (lmax - 1.0) / 2.0
    return theta, delta, theta / delta


def _alpha0_problem(orc, grid, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.5 * rng.random(grid.shape)
    xs = rng.standard_normal(grid.shape)
    x0 = rng.standard_normal(grid.shape)
    b = grid.apply_sip(0.0, xs, rho)
    M = grid.sip_diag(0.0, rho)
    return rho, xs, x0, b, M


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_alpha0_contraction_closed_form(orc, dim, p):
    g = orc.Grid2D(p, 3, 2, 1.0, 0.8) if dim == 2 else orc.Grid1D(p, 5, 1.3)
    rho, xs, x0, b, M = _alpha0_problem(orc, g, 20 + p)
    # M^-1 A = I at alpha = 0 (to rounding)
    np.testing.assert_allclose(g.apply_sip(0.0, x0, rho) / M, x0, rtol=4e-16 * 8, atol=0)
    e0 = xs - x0
    for k in (1, 2, 5, 9):
        xj = g.solver(0.0, orc.JACOBI, k, rho, b, M, x0)
        np.testing.assert_allclose(xs - xj, e0 / 3.0 ** k, rtol=0, atol=1e-13 * np.max(np.abs(e0)))
        theta, delta, s = _sigma(p)
        xc = g.solver(0.0, orc.CHEBYSHEV, k, rho, b, M, x0)
        # T_k(s) for s > 1 in closed form
        tk = np.cosh(k * np.arccosh(s))
        assert abs(tk - _cheb_T(k, s)) <= 1e-12 * tk
        np.testing.assert_allclose(xs - xc, e0 / tk, rtol=0, atol=1e-13 * np.max(np.abs(e0)))


def _dense(apply, shape):
    n = int(np.prod(shape))
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = apply(e.reshape(shape)).ravel()
    return A


@pytest.mark.parametrize("dim,p", [(2, 3), (2, 2), (1, 3)])
def test_residual_polynomials_alpha_positive(orc, dim, p):
    """alpha > 0: the error after k steps is P_k(M^-1 A) e_0, evaluated through
    the eigen-decomposition of the symmetric M^-1/2 A M^-1/2."""
    g = orc.Grid2D(p, 3, 3, 1.0, 0.9) if dim == 2 else orc.Grid1D(p, 6, 1.0)
    rng = np.random.default_rng(31 + p + dim)
    rho = 1.0 + 0.4 * rng.random(g.shape)
    a = 0.012
    A = _dense(lambda e: g.apply_sip(a, e, rho), g.shape)
    Md = g.sip_diag(a, rho).ravel()
    A = 0.5 * (A + A.T)
    ih = 1.0 / np.sqrt(Md)
    lam, V = np.linalg.eigh(ih[:, None] * A * ih[None, :])
    xs = rng.standard_normal(g.N)
    x0 = rng.standard_normal(g.N)
    b = A @ xs
    e0 = xs - x0
    w0 = V.T @ (np.sqrt(Md) * e0)      # error in the eigenbasis of M^-1/2 A M^-1/2

    def predicted(poly):
        return ih * (V @ (poly(lam) * w0))

    theta, delta, s = _sigma(p)
    for k in (1, 3, 8):
        xj = g.solver(a, orc.JACOBI, k, rho, b.reshape(g.shape), Md.reshape(g.shape), x0.reshape(g.shape))
        ej = predicted(lambda t: (1.0 - 2.0 / 3.0 * t) ** k)
        np.testing.assert_allclose((xs - xj.ravel()), ej, rtol=0, atol=1e-11 * np.max(np.abs(e0)))
        xc = g.solver(a, orc.CHEBYSHEV, k, rho, b.reshape(g.shape), Md.reshape(g.shape), x0.reshape(g.shape))
        ec = predicted(lambda t: _cheb_T(k, (theta - t) / delta) / _cheb_T(k, s))
        np.testing.assert_allclose((xs - xc.ravel()), ec, rtol=0, atol=1e-11 * np.max(np.abs(e0)))


def _alpha0_errors(orc, m, solver, p=3, k=6):
    g = orc.Grid2D(p, 3, 2, 1.0, 0.8)
    rho, xs, x0, b, M = _alpha0_problem(orc, g, 23)
    with orc.fwd_mutation(m):
        x = g.solver(0.0, solver, k, rho, b, M, x0)
    e0 = xs - x0
    theta, delta, s = _sigma(p)
    want = e0 / 3.0 ** k if solver == orc.JACOBI else e0 / np.cosh(k * np.arccosh(s))
    return np.max(np.abs((xs - x) - want)) / np.max(np.abs(e0))


def test_solver_pins_reject_planted_mistakes(orc):
    assert _alpha0_errors(orc, 0, orc.JACOBI) < 1e-13
    assert _alpha0_errors(orc, 0, orc.CHEBYSHEV) < 1e-13
    assert _alpha0_errors(orc, 1, orc.JACOBI) > 1e-3           # omega = 0.5
    assert _alpha0_errors(orc, 2, orc.CHEBYSHEV) > 1e-8        # lambda_max = 2.2 at p = 3 (3e-5)
    # the mutations leave the other solver alone
    assert _alpha0_errors(orc, 1, orc.CHEBYSHEV) < 1e-13
    assert _alpha0_errors(orc, 2, orc.JACOBI) < 1e-13


# ------------------------------------------------ 2D vortex convergence ----
def _vortex_l2_error(orc, p, N, mutation=0):
    """run_euler_test.jl:84-119: L = T = 10, beta = 5, (ux0, uy0) = (1, 1),
    CFL 0.4, alpha = 0; the L2 error of rho against the initial state after
    one diagonal period."""
    L, gam = 10.0, 1.4
    g = orc.Grid2D(p, N, N, L, L)
    q0 = configs.isentropic_vortex2(p, N, N, gam, 5.0, 1.0, 1.0, L, L)
    dt = 0.4 * (L / N) / ((2 * p + 1) * configs.max_wavespeed2(q0, gam))
    with orc.fwd_mutation(mutation):
        qf, _, _ = g.run_forward(gam, 0.0, 0, orc.PCG, dt, L, q0)
    return float(np.sqrt(np.sum(g.quad_weights() * (qf[2] - q0[2]) ** 2)))


@pytest.mark.parametrize("p", [2, 3])
def test_isentropic_vortex_order_2d(orc, p):
    e16, e32 = _vortex_l2_error(orc, p, 16), _vortex_l2_error(orc, p, 32)
    order = np.log2(e16 / e32)
    assert order > p + 0.5, (e16, e32, order)


def test_vortex_rejects_swapped_flux_y(orc):
    good = _vortex_l2_error(orc, 2, 16)
    try:
        bad = _vortex_l2_error(orc, 2, 16, mutation=3)
    except orc.DomainError:
        bad = np.inf
    assert bad > 10 * good
