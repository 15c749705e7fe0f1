"""Pins the host restatements of the reference's initial conditions (which
the device builders are compared with) to what mathematics fixes: the Oseen
enthalpy-defect table against the exact integral, and closed-form values."""
import numpy as np
from scipy.special import exp1

from igr_sensitivities_amd import configs


def _oseen_table(gamma=1.4, a=0.05, M_star=0.5, L=1.0):
    """The table oseen_pair_blast builds, rebuilt through its own defect():
    one vortex at the origin, no blast, sampled along a ray."""
    r = np.linspace(0.0, 0.49, 50)
    x, y = r.copy(), np.zeros_like(r)
    q = configs.oseen_pair_blast(x, y, gamma, A=0.0, a_core=a, M_star=M_star,
                                 centers=((0.0, 0.0, 1.0),), blast=(0.5, 0.5), L=L)
    rho = q[2]
    p_vort = rho ** gamma
    h = gamma / (gamma - 1) * p_vort ** ((gamma - 1) / gamma)
    return r, gamma / (gamma - 1) - h      # D(r) recovered from the state


def _defect_exact(r, gamma=1.4, a=0.05, M_star=0.5, L=1.0):
    """(Gamma/2pi)^2 int_r^{2L} (1 - e^{-s^2/a^2})^2 / s^3 ds in closed form:
    with u = s^2/a^2 it is (Gamma/2pi)^2 / (2 a^2) [F(u_r) - F(u_max)],
    F(x) = int_x^inf (1 - e^{-u})^2 / u^2 du = (1 - e^{-x})^2 / x + 2 E1(x) - 2 E1(2x),
    F(0) = 2 ln 2."""
    G = M_star * np.sqrt(gamma) * 2 * np.pi * a / 0.6383
    pref = (G / (2 * np.pi)) ** 2 / (2 * a * a)

    def F(x):
        x = np.asarray(x, dtype=float)
        out = np.full(x.shape, 2 * np.log(2.0))
        nz = x > 0
        xs = x[nz]
        out[nz] = (1 - np.exp(-xs)) ** 2 / xs + 2 * exp1(xs) - 2 * exp1(2 * xs)
        return out
    return pref * (F(r ** 2 / a ** 2) - F(np.array((2 * L) ** 2 / a ** 2)))


def test_oseen_defect_table_matches_exact_integral():
    r, D = _oseen_table()
    exact = _defect_exact(r)
    # trapezoid rule on 4000 points over [0, 2]: h = 5e-4 against the core
    # scale a = 0.05; the error stays below 1e-4 of D(0)
    assert np.max(np.abs(D - exact)) < 1e-4 * exact[0]
    assert abs(D[0] - exact[0]) < 1e-4 * exact[0]
    # radial equilibrium: dD/dr = -v_theta^2 / r (the defect falls off outwards)
    assert np.all(np.diff(D) <= 1e-15)


def test_oseen_pair_mirror_symmetry():
    # the pair (0.45, 0.60, +1), (0.45, 0.40, -1), its 3x3 images and the
    # blast at y = 0.5 map onto themselves under y -> 1 - y (a reflection
    # reverses each vortex's sense): rho, p, u_x even, u_y odd
    rng = np.random.default_rng(0)
    x, y = rng.uniform(0, 1, 200), rng.uniform(0, 1, 200)
    a = configs.oseen_pair_blast(x, y, 1.4)
    b = configs.oseen_pair_blast(x, 1.0 - y, 1.4)
    np.testing.assert_allclose(b[2], a[2], rtol=1e-12)
    np.testing.assert_allclose(b[3], a[3], rtol=1e-12)
    sc = np.max(np.abs(a[:2]))
    np.testing.assert_allclose(b[0], a[0], rtol=0, atol=1e-12 * sc)
    np.testing.assert_allclose(b[1], -a[1], rtol=0, atol=1e-12 * sc)
    assert np.max(np.abs(a[1])) > 0.1 * sc      # u_y is not trivially zero


def test_isentropic_vortex_centre_closed_form():
    # node at the domain centre (even element counts: an element corner)
    p, N, gamma, beta = 3, 4, 1.4, 5.0
    q = configs.isentropic_vortex2(p, N, N, gamma, beta)
    rho_c = (1 - (gamma - 1) * beta ** 2 / (8 * gamma * np.pi ** 2) * np.e) ** (1 / (gamma - 1))
    # element (ex, ey) = (N/2, N/2), node (0, 0) sits at (0.5, 0.5)
    rc = q[2, N // 2, N // 2, 0, 0]
    assert abs(rc - rho_c) < 1e-14
    # isentropic: p = rho^gamma everywhere
    rho, ux, uy = q[2], q[0] / q[2], q[1] / q[2]
    pr = (gamma - 1) * (q[3] - 0.5 * rho * (ux ** 2 + uy ** 2))
    np.testing.assert_allclose(pr, rho ** gamma, rtol=1e-13)


def test_sedov_centre_and_background():
    p, N, A = 3, 4, 7.0
    q = configs.sedov_blast2(p, N, N, A, 1.4, p_bg=0.5, sigma=0.1)
    assert abs(q[3, N // 2, N // 2, 0, 0] * 0.4 - (0.5 + A)) < 1e-12
    assert abs(q[3, 0, 0, 0, 0] * 0.4 - (0.5 + A * np.exp(-0.5 / 0.01))) < 1e-12
