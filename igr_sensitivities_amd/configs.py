"""Seeded synthetic workloads (host-side set-up, numpy FP64).

The reference defines its initial conditions in
IGR_adjoints_2d/initial_conditions.jl and its paper scripts; the workloads
here are the ones this library is measured on:

  C1  1D forward: periodic [0,1], N_e=128, p=3, gamma=1.4, rho=1, u=0,
      pressure 1000 / 0.01 / 100 on [0,0.1) / [0.1,0.9) / [0.9,1] with tanh
      edges of width 0.01, times (1 + 1e-3 eta(x)) with eta a seeded sum of 8
      periodic sine modes (max|eta| = 1), alpha = 10 h^2, CFL 0.1 to
      t = 0.038, Jacobi-50.
  C2  2D forward: [0,1]^2, 128x128, p=3, rho=1, u=v=0, p_bg=1 plus three
      Gaussian energy deposits (sigma=0.02, total E = 1, seeded centres),
      alpha = 10 (h/4)^2, PCG-60, 200 SSP-RK3 steps, CFL 0.4.
  C3  C2 with J = sum w w Jx Jy rho(T) exp(-|x - c_w|^2 / 0.1^2), c_w seeded.
  C4  64 instances, 1D and y-uniform 2D of mixed sizes, shifted C1 data,
      4 tangent directions each, run as one batch.
  C5  2D blast beside the reference's Oseen vortex pair, 256x256, PCG-60,
      400 steps, gradient of the kinetic energy in a disc.

The time step follows the reference scripts' formula
dt = CFL * dx / ((2p+1) * max(|u| + c)) (IGR_adjoints_1d/enzyme_adjoint.jl:272-273,
IGR_adjoints_2d/tests/run_quasi1d_test.jl:70).
"""
import math

import numpy as np


def lgl(p):
    """LGL nodes and weights (host copy for set-up; same Newton iteration as the library)."""
    n = p + 1
    xi = np.empty(n)
    for i in range(n):
        x = -math.cos(math.pi * i / p)
        for _ in range(100):
            P0, P1 = 1.0, x
            for k in range(2, p + 1):
                P0, P1 = P1, ((2.0 * k - 1.0) * x * P1 - (k - 1.0) * P0) / k
            dx = (x * P1 - P0) / (n * P1)
            x -= dx
            if abs(dx) < 1e-17:
                break
        xi[i] = x
    xi[0], xi[-1] = -1.0, 1.0
    for i in range(n // 2):
        h = 0.5 * (xi[n - 1 - i] - xi[i])
        xi[i], xi[n - 1 - i] = -h, h
    if n % 2:
        xi[n // 2] = 0.0
    w = np.empty(n)
    for i in range(n):
        x = xi[i]
        P0, P1 = 1.0, x
        for k in range(2, p + 1):
            P0, P1 = P1, ((2.0 * k - 1.0) * x * P1 - (k - 1.0) * P0) / k
        w[i] = 2.0 / (p * (p + 1.0) * P1 * P1)
    return xi, w


def coords2(p, Nex, Ney, Lx=1.0, Ly=1.0):
    """Node coordinates, shape (Ney, Nex, np, np) (CartesianMesh2D, IGR_adjoints_2d/mesh.jl:48-55)."""
    xi, _ = lgl(p)
    dx, dy = Lx / Nex, Ly / Ney
    ex = np.arange(Nex)[None, :, None, None]
    ey = np.arange(Ney)[:, None, None, None]
    x = ex * dx + (dx / 2.0) * (1.0 + xi[None, None, None, :])
    y = ey * dy + (dy / 2.0) * (1.0 + xi[None, None, :, None])
    shape = (Ney, Nex, p + 1, p + 1)
    return np.broadcast_to(x, shape).copy(), np.broadcast_to(y, shape).copy()


def quad2(p, Nex, Ney, Lx=1.0, Ly=1.0):
    _, w = lgl(p)
    Jx, Jy = Lx / Nex / 2.0, Ly / Ney / 2.0
    wij = w[None, :] * w[:, None]
    return np.broadcast_to(wij * Jx * Jy, (Ney, Nex, p + 1, p + 1)).copy()


def coords1(p, Ne, L=1.0):
    xi, _ = lgl(p)
    dx = L / Ne
    return np.arange(Ne)[:, None] * dx + (dx / 2.0) * (1.0 + xi[None, :])


def max_wavespeed2(q, gamma):
    mx, my, r, E = q
    P = (gamma - 1.0) * (E - (mx * mx + my * my) / (2.0 * r))
    c = np.sqrt(gamma * P / r)
    return float(np.max(np.sqrt(mx * mx + my * my) / r + c))


def max_wavespeed1(q, gamma):
    m, r, E = q
    P = (gamma - 1.0) * (E - m * m / (2.0 * r))
    return float(np.max(np.abs(m / r) + np.sqrt(gamma * P / r)))


def alpha_of(h, p):
    """alpha = 10 (h/(p+1))^2."""
    return 10.0 * (h / (p + 1)) ** 2


def periodic_d2(x, y, xc, yc, Lx=1.0, Ly=1.0):
    dx = np.abs(x - xc)
    dx = np.minimum(dx, Lx - dx)
    dy = np.abs(y - yc)
    dy = np.minimum(dy, Ly - dy)
    return dx * dx + dy * dy


def periodic_box(x, a, b, w):
    """Smoothed indicator of [a, b) on the unit circle (tanh edges of width w)."""
    return sum(0.5 * (np.tanh((x - a - k) / w) - np.tanh((x - b - k) / w)) for k in (-1, 0, 1))


def config1(Ne=128, p=3, seed=1, gamma=1.4, width=0.01):
    """C1 1D blast data with seeded relative noise.

    Pressure 1000 / 0.01 / 100 on [0,0.1) / [0.1,0.9) / [0.9,1], the three
    levels joined by tanh edges of width 0.01 (periodically, across x = 0),
    times (1 + 1e-3 eta).  alpha = 10 h^2 (h = 1/N_e), the reference's own
    1D choice (IGR_adjoints_1d/paper_scripts/generate_gamma_sensitivity_data.jl).
    With sharp steps, or with alpha = 10 (h/4)^2, the unlimited DG scheme
    reaches a negative pressure (Julia DomainError) within the run.
    """
    rng = np.random.default_rng(seed)
    x = coords1(p, Ne)
    k = np.arange(1, 9)
    amp = rng.standard_normal(8)
    ph = rng.uniform(0, 2 * np.pi, 8)
    eta = np.sum(amp[:, None, None] * np.sin(2 * np.pi * k[:, None, None] * x[None] + ph[:, None, None]), axis=0)
    eta /= np.max(np.abs(eta))
    base = 0.01 + (1000.0 - 0.01) * periodic_box(x, 0.0, 0.1, width) + (100.0 - 0.01) * periodic_box(x, 0.9, 1.0, width)
    pr = base * (1.0 + 1e-3 * eta)
    q = np.zeros((3, Ne, p + 1))
    q[1] = 1.0
    q[2] = pr / (gamma - 1.0)
    h = 1.0 / Ne
    lam = max_wavespeed1(q, gamma)
    dt = 0.1 * h / ((2 * p + 1) * lam)
    return dict(q=q, gamma=gamma, alpha=10.0 * h * h, dt=dt, T=0.038, n_iter=50, solver="jacobi",
                p=p, Ne=Ne)


def c1_pressure(x, seed=1, width=0.01, d=False):
    """C1 pressure profile as a periodic function of x (and its x-derivative
    when d=True): 1000 / 0.01 / 100 on [0,0.1) / [0.1,0.9) / [0.9,1] with tanh
    edges of width `width`, times (1 + 1e-3 eta(x)), eta the seeded sum of 8
    sine modes normalised to max|eta| = 1 on a 20001-point grid.  Returns
    (p, base, eta) or, with d=True, (dp/dx, base * eta)."""
    rng = np.random.default_rng(seed)
    k = np.arange(1, 9)
    amp = rng.standard_normal(8)
    ph = rng.uniform(0, 2 * np.pi, 8)
    xf = np.linspace(0.0, 1.0, 20001)
    scale = np.max(np.abs(np.sum(amp[:, None] * np.sin(2 * np.pi * k[:, None] * xf[None] + ph[:, None]), axis=0)))
    arg = 2 * np.pi * k[:, None] * x.ravel()[None] + ph[:, None]
    eta = (np.sum(amp[:, None] * np.sin(arg), axis=0) / scale).reshape(x.shape)
    deta = (np.sum(amp[:, None] * 2 * np.pi * k[:, None] * np.cos(arg), axis=0) / scale).reshape(x.shape)

    def box(a, b):
        return sum(0.5 * (np.tanh((x - a - j) / width) - np.tanh((x - b - j) / width)) for j in (-1, 0, 1))

    def dbox(a, b):
        return sum(0.5 / width * (np.cosh((x - a - j) / width) ** -2 - np.cosh((x - b - j) / width) ** -2)
                   for j in (-1, 0, 1))

    base = 0.01 + (1000.0 - 0.01) * box(0.0, 0.1) + (100.0 - 0.01) * box(0.9, 1.0)
    if d:
        dbase = (1000.0 - 0.01) * dbox(0.0, 0.1) + (100.0 - 0.01) * dbox(0.9, 1.0)
        return dbase * (1.0 + 1e-3 * eta) + base * 1e-3 * deta, base * eta
    return base * (1.0 + 1e-3 * eta), base, eta


def config4(n_inst=64, seed=4, p=3, steps=(3, 6), n_tan=4):
    """C4 batched ensemble: n_inst instances, alternately 1D (N_e in
    {256, ..., 4096}) and 2D (N x N, N in {32, ..., 128}, y-uniform), of the
    C1 data cyclically shifted by a seeded s in [0, 1), gamma ~ U[1.2, 1.67],
    alpha = 10 h^2, CFL 0.1, a seeded number of SSP-RK3 steps in `steps`,
    damped Jacobi-50.  Tangent directions: d/ds (the shift, -dq0/dx), d/d eps
    (the noise amplitude 1e-3) and two seeded smooth random fields."""
    rng = np.random.default_rng(seed)
    runs = []
    for k in range(n_inst):
        dim = 1 if k % 2 == 0 else 2
        N = int(rng.choice([256, 512, 1024, 2048, 4096])) if dim == 1 else int(rng.choice([32, 48, 64, 96, 128]))
        gam = float(rng.uniform(1.2, 1.67))
        sh = float(rng.uniform(0.0, 1.0))
        ns = int(rng.integers(steps[0], steps[1] + 1))
        if dim == 1:
            x = coords1(p, N)
            shape = x.shape
        else:
            x, _ = coords2(p, N, N)
            shape = x.shape
        xs = (x - sh) % 1.0
        pr, _, _ = c1_pressure(xs)
        dp, beta = c1_pressure(xs, d=True)
        nv = dim + 2
        q = np.zeros((nv,) + shape)
        q[nv - 2] = 1.0
        q[nv - 1] = pr / (gam - 1.0)
        s0 = np.zeros((n_tan, nv) + shape)
        s0[0, nv - 1] = -dp / (gam - 1.0)
        if n_tan > 1:
            s0[1, nv - 1] = beta / (gam - 1.0)
        for t in range(2, n_tan):
            for v in range(nv):
                f = np.zeros(shape)
                for m in range(1, 4):
                    f += rng.standard_normal() * np.cos(2 * np.pi * m * x + rng.uniform(0, 2 * np.pi))
                s0[t, v] = 0.1 * f / 3.0
        h = 1.0 / N
        lam = max_wavespeed1(q, gam) if dim == 1 else max_wavespeed2(q, gam)
        dt = 0.1 * h / ((2 * p + 1) * lam)
        runs.append(dict(dim=dim, N=N, q0=q, s0=s0, gamma=gam, alpha=10.0 * h * h, dt=dt, T=(ns - 0.5) * dt,
                         n_steps=ns, shift=sh))
    return dict(runs=runs, n_iter=50, solver="jacobi", p=p)


def config2(N=128, p=3, seed=2, gamma=1.4, n_steps=200, cfl=0.4, sigma=0.02, p_bg=1.0):
    """C2 2D triple energy deposit.

    Background pressure 1, as in the reference's triple-Sedov script
    (IGR_adjoints_2d/paper_scripts/generate_sedov_heatmap_data.jl).  With
    p_bg = 1e-5 the unlimited DG scheme reaches a negative pressure (Julia
    DomainError) after 86 of the 200 steps at 128x128.
    """
    rng = np.random.default_rng(seed)
    x, y = coords2(p, N, N)
    wq = quad2(p, N, N)
    E = np.full(x.shape, p_bg / (gamma - 1.0))
    centres = rng.uniform(0.15, 0.85, size=(3, 2))
    for xc, yc in centres:
        g = np.exp(-periodic_d2(x, y, xc, yc) / sigma ** 2)
        E += (1.0 / 3.0) * g / np.sum(wq * g)
    q = np.zeros((4,) + x.shape)
    q[2] = 1.0
    q[3] = E
    h = 1.0 / N
    lam = max_wavespeed2(q, gamma)
    dt = cfl * h / ((2 * p + 1) * lam)
    return dict(q=q, gamma=gamma, alpha=alpha_of(h, p), dt=dt, n_steps=n_steps, n_iter=60, solver="pcg",
                p=p, N=N, centres=centres)


def config3_objective(seed=3):
    """C3 window centre and width for J = int rho(T) exp(-|x-c|^2/0.1^2)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.7, size=2)
    return (float(c[0]), float(c[1]), 0.1, 0.0)


def oseen_pair_blast(x, y, gamma, A=100.0, sigma=0.03, a_core=0.05, M_star=0.5,
                     centers=((0.45, 0.60, 1.0), (0.45, 0.40, -1.0)), blast=(0.25, 0.50), L=1.0):
    """Counter-rotating pair of Oseen vortices plus a Gaussian over-pressure
    blast, as build_oseen_pair_blast_ic! (IGR_adjoints_2d/initial_conditions.jl:
    152-232) sets it up: circulation from the vortex Mach number (v_max =
    0.6383 Gamma / (2 pi a)), velocity summed over the 3x3 periodic images,
    homentropic radial equilibrium through the tabulated enthalpy defect
    D(r) = (Gamma/2pi)^2 int_r^inf (1 - exp(-s^2/a^2))^2 / s^3 ds (4000 points
    on [0, 2L], trapezoid rule from the outer end, linear interpolation).
    Returns q = (mx, my, rho, E)."""
    Gam = M_star * np.sqrt(gamma) * 2 * np.pi * a_core / 0.6383
    r_max, n_tab = 2.0 * L, 4000
    r_tab = np.linspace(0.0, r_max, n_tab)
    f = np.zeros(n_tab)
    nz = r_tab >= 1e-12
    f[nz] = (1.0 - np.exp(-r_tab[nz] ** 2 / a_core ** 2)) ** 2 / r_tab[nz] ** 3
    seg = 0.5 * (f[:-1] + f[1:]) * np.diff(r_tab)
    D_tab = np.zeros(n_tab)
    D_tab[:-1] = (Gam / (2 * np.pi)) ** 2 * np.cumsum(seg[::-1])[::-1]

    def defect(r):
        t = r / r_max * (n_tab - 1)
        i = np.clip(np.floor(t).astype(int), 0, n_tab - 2)
        w = t - i
        return np.where(r >= r_max, 0.0, (1 - w) * D_tab[i] + w * D_tab[i + 1])

    ux = np.zeros_like(x)
    uy = np.zeros_like(x)
    h = np.full_like(x, gamma / (gamma - 1.0))
    for xc, yc, sgn in centers:
        for sx in (-1, 0, 1):
            for sy in (-1, 0, 1):
                dx = x - (xc + sx * L)
                dy = y - (yc + sy * L)
                r2 = dx * dx + dy * dy
                vr = np.where(r2 < 1e-24, 0.0, sgn * Gam / (2 * np.pi * np.maximum(r2, 1e-300))
                              * (1.0 - np.exp(-r2 / a_core ** 2)))
                ux += -vr * dy
                uy += vr * dx
        dx = x - xc
        dy = y - yc
        dx -= L * np.round(dx / L)
        dy -= L * np.round(dy / L)
        h -= defect(np.sqrt(dx * dx + dy * dy))
    p_vort = ((gamma - 1) * h / gamma) ** (gamma / (gamma - 1))
    rho = p_vort ** (1.0 / gamma)
    pr = p_vort + A * np.exp(-((x - blast[0]) ** 2 + (y - blast[1]) ** 2) / sigma ** 2)
    return np.stack([rho * ux, rho * uy, rho, pr / (gamma - 1) + 0.5 * rho * (ux * ux + uy * uy)])


def config5(N=256, p=3, seed=5, gamma=1.4, n_steps=400, cfl=0.4):
    """C5 blast beside vortices at 256x256: the Gaussian blast of peak 100 and
    sigma 0.03 next to the reference's counter-rotating Oseen pair (core radius
    0.05, vortex Mach number 0.5; oseen_pair_blast), PCG-60, 400 steps, and
    the kinetic energy in a seeded disc (radius 0.15 around the pair) as the
    objective of the gradient with respect to gamma and alpha.

    A single Lamb-Oseen vortex of circulation 1 on rho = p = 1 is not an
    admissible state: its core pressure from radial equilibrium is negative,
    and a net circulation is not periodic.  The pair of the reference's own
    blast-vortex builder keeps zero net circulation and a positive pressure."""
    rng = np.random.default_rng(seed)
    x, y = coords2(p, N, N)
    q = oseen_pair_blast(x, y, gamma)
    h = 1.0 / N
    lam = max_wavespeed2(q, gamma)
    dt = cfl * h / ((2 * p + 1) * lam)
    disc = (float(rng.uniform(0.42, 0.48)), float(rng.uniform(0.47, 0.53)), 0.15, 0.0)
    return dict(q=q, gamma=gamma, alpha=alpha_of(h, p), dt=dt, n_steps=n_steps, n_iter=60, solver="pcg",
                p=p, N=N, disc=disc)


def sedov_heatmap_p4(N=160, p=4, gamma=1.4, cfl=0.4, T=0.10):
    """The p = 4 workload of the reference's triple-Sedov heatmap runs
    (IGR_adjoints_2d/paper_scripts/generate_sedov_heatmap_data.jl:41-96):
    160 x 160 elements, alpha = 3 (L/20)^2, rho = 1, u = 0, p = 1 + three
    Gaussian over-pressures (A = 6, 5, 5.5 at (0.30, 0.34), (0.70, 0.40),
    (0.52, 0.74), sigma = 0.09), CFL 0.4 on max(lambda_x, lambda_y) over
    (2p + 1), n_steps = ceil(T / dt), PCG-60."""
    x, y = coords2(p, N, N)
    pr = np.ones_like(x)
    for xc, yc, A in ((0.30, 0.34, 6.0), (0.70, 0.40, 5.0), (0.52, 0.74, 5.5)):
        pr += A * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / 0.09 ** 2)
    q = np.zeros((4,) + x.shape)
    q[2] = 1.0
    q[3] = pr / (gamma - 1.0)
    dt = cfl * (1.0 / N) / ((2 * p + 1) * max_wavespeed2(q, gamma))
    n_steps = int(math.ceil(T / dt))
    return dict(q=q, gamma=gamma, alpha=3.0 * (1.0 / 20) ** 2, dt=T / n_steps, n_steps=n_steps, T=T, n_iter=60,
                solver="pcg", p=p, N=N)


def taylor_green2(p, Nex, Ney, A=0.1, gamma=1.4, Lx=1.0, Ly=1.0):
    """build_taylor_green_ic! (IGR_adjoints_2d/initial_conditions.jl:105-121)."""
    x, y = coords2(p, Nex, Ney, Lx, Ly)
    cx, cy = 2 * np.pi * x / Lx, 2 * np.pi * y / Ly
    ux = A * np.sin(cx) * np.cos(cy)
    uy = -A * np.cos(cx) * np.sin(cy)
    q = np.zeros((4,) + x.shape)
    q[2] = 1.0
    q[0], q[1] = ux, uy
    q[3] = 1.0 / (gamma - 1.0) + 0.5 * (ux ** 2 + uy ** 2)
    return q


def constant2(p, Nex, Ney, gamma=1.4, rho0=1.0, ux0=1.0, uy0=1.0, p0=1.0, Lx=1.0, Ly=1.0):
    """init_constant_state (IGR_adjoints_2d/initial_conditions.jl:18-25)."""
    shape = (Ney, Nex, p + 1, p + 1)
    q = np.empty((4,) + shape)
    q[0], q[1], q[2] = rho0 * ux0, rho0 * uy0, rho0
    q[3] = p0 / (gamma - 1) + 0.5 * rho0 * (ux0 ** 2 + uy0 ** 2)
    return q


def isentropic_vortex2(p, Nex, Ney, gamma=1.4, beta=5.0, ux0=1.0, uy0=1.0, Lx=1.0, Ly=1.0):
    """init_isentropic_vortex (IGR_adjoints_2d/initial_conditions.jl:42-71)."""
    x, y = coords2(p, Nex, Ney, Lx, Ly)
    dx, dy = x - Lx / 2, y - Ly / 2
    r2 = dx ** 2 + dy ** 2
    ef = np.exp((1 - r2) / 2)
    du = -dy * (beta / (2 * np.pi)) * ef
    dv = dx * (beta / (2 * np.pi)) * ef
    dT = -(gamma - 1) * beta ** 2 / (8 * gamma * np.pi ** 2) * np.exp(1 - r2)
    rho = (1 + dT) ** (1 / (gamma - 1))
    ux, uy = ux0 + du, uy0 + dv
    pr = rho ** gamma
    return np.stack([rho * ux, rho * uy, rho, pr / (gamma - 1) + 0.5 * rho * (ux ** 2 + uy ** 2)])


def sedov_blast2(p, Nex, Ney, A, gamma=1.4, p_bg=1.0, sigma=0.1, Lx=1.0, Ly=1.0):
    """build_sedov_blast_ic! (IGR_adjoints_2d/initial_conditions.jl:136-151)."""
    x, y = coords2(p, Nex, Ney, Lx, Ly)
    r2 = (x - Lx / 2) ** 2 + (y - Ly / 2) ** 2
    q = np.zeros((4,) + x.shape)
    q[2] = 1.0
    q[3] = (p_bg + A * np.exp(-r2 / sigma ** 2)) / (gamma - 1)
    return q


def entropy_wave_init1(p, Ne, gamma=1.4, L=1.0):
    """init_entropy_wave (IGR_adjoints_1d/initial_conditions.jl:24-44)."""
    x = coords1(p, Ne, L)
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x / L)
    return np.stack([rho, rho, 1.0 / (gamma - 1) + 0.5 * rho])


def smooth_random2(p, Nex, Ney, seed, gamma=1.4, amp=0.1):
    """Smooth positive random state (few low modes) for parity tests."""
    rng = np.random.default_rng(seed)
    x, y = coords2(p, Nex, Ney)

    def fld():
        f = np.zeros_like(x)
        for kx in range(3):
            for ky in range(3):
                f += rng.standard_normal() * np.cos(2 * np.pi * (kx * x + ky * y) + rng.uniform(0, 2 * np.pi))
        return f / 9.0

    q = np.zeros((4,) + x.shape)
    q[2] = 1.0 + 0.5 * amp * fld()
    q[0] = amp * fld() * q[2]
    q[1] = amp * fld() * q[2]
    pr = 1.0 + 0.5 * amp * fld()
    q[3] = pr / (gamma - 1.0) + 0.5 * (q[0] ** 2 + q[1] ** 2) / q[2]
    return q


def smooth_random1(p, Ne, seed, gamma=1.4, amp=0.1):
    rng = np.random.default_rng(seed)
    x = coords1(p, Ne)

    def fld():
        f = np.zeros_like(x)
        for k in range(4):
            f += rng.standard_normal() * np.cos(2 * np.pi * k * x + rng.uniform(0, 2 * np.pi))
        return f / 4.0

    q = np.zeros((3,) + x.shape)
    q[1] = 1.0 + 0.5 * amp * fld()
    q[0] = amp * fld() * q[1]
    pr = 1.0 + 0.5 * amp * fld()
    q[2] = pr / (gamma - 1.0) + 0.5 * q[0] ** 2 / q[1]
    return q


# 1D single-mode pulses of IGR_adjoints_1d/initial_conditions.jl:94-204
# (IC_BUILDERS, :218-224): (rho, u, p) = (1, 0, 1) plus A sin(2 pi x / L) on
# the primitives each kind perturbs, E = p/(gamma-1) + rho u^2 / 2.
PULSE_KINDS = {"acoustic": (1, 1, 1), "entropy": (1, None, 0), "density": (1, 0, 0), "pressure": (0, 0, 1),
               "velocity": (0, 1, 0)}


def pulse1(kind, p, Ne, A, gamma=1.4, L=1.0):
    """build_<kind>_pulse_ic! (entropy: build_entropy_wave_ic!, u = 1)."""
    wr, wu, wp = PULSE_KINDS[kind]
    x = coords1(p, Ne, L)
    s = np.sin(2 * np.pi * x / L)
    rho = 1 + A * s if wr else np.ones_like(x)
    u = np.ones_like(x) if wu is None else (A * s if wu else np.zeros_like(x))
    pr = 1 + A * s if wp else np.ones_like(x)
    q = np.zeros((3,) + x.shape)
    q[0] = rho * u
    q[1] = rho
    q[2] = pr / (gamma - 1) + 0.5 * rho * u ** 2
    return q


def pulse1_dA(kind, p, Ne, A, gamma=1.4, L=1.0):
    """d q0 / d A of pulse1 in conservative variables (the chain rule of
    dJ_dA_from_adjoint, IGR_adjoints_1d/tests/run_adjoint_conservative_test.jl:22-47)."""
    wr, wu, wp = PULSE_KINDS[kind]
    x = coords1(p, Ne, L)
    s = np.sin(2 * np.pi * x / L)
    q = pulse1(kind, p, Ne, A, gamma, L)
    rho, u = q[1], q[0] / q[1]
    drho = (wr or 0) * s
    du = (wu or 0) * s
    dp = (wp or 0) * s
    d = np.zeros_like(q)
    d[0] = rho * du + u * drho
    d[1] = drho
    d[2] = dp / (gamma - 1) + 0.5 * (drho * u ** 2 + 2 * rho * u * du)
    return d
