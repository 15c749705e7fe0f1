"""Host-side mirror of the reference's Julia operator API over the C ABI.

The reference (f-t-s/IGR_sensitivities) exposes positional, mutating Julia
functions — run_forward!(q..., Sigma, dt, T, basis, mesh, gamma, alpha,
n_iter, solver) and friends (IGR_adjoints_2d/forward.jl:6,
elliptic.jl:8-10, IGR_adjoints_1d/forward.jl:5, elliptic.jl:7-8).  This
module keeps their names, argument meaning and error behaviour:

* arrays are numpy float64 arrays in the reference layout (2D field shape
  (Ney, Nex, np, np) in C order == Julia (np, np, Nex, Ney); 1D (Ne, np));
* `solver` is "pcg" | "jacobi" | "chebyshev" (Julia Symbols); an unknown
  solver raises ValueError like Julia's ErrorException;
* a negative pressure under sqrt raises DomainError, as Julia's sqrt does;
* the in-place functions (suffix `_`) overwrite their array arguments.

All arithmetic runs in libigr_hip.so on the GPU; nothing here computes.
"""
import ctypes as C

import numpy as np

from . import _lib

SOLVERS = {"pcg": 0, "jacobi": 1, "chebyshev": 2}
OPT_GRAPHS, OPT_TILE, OPT_HTILE, OPT_AUTOTUNE, OPT_RTILE, OPT_GROUPS, OPT_WAVE1D, OPT_KEEP, OPT_RGROUPS = 0, 1, 2, 3, 4, 5, 6, 7, 8


class DomainError(ArithmeticError):
    """Negative pressure under sqrt (Julia DomainError)."""


class IGRError(RuntimeError):
    pass


def _solver_code(solver):
    if isinstance(solver, str):
        key = solver.lstrip(":").lower()
        if key not in SOLVERS:
            raise ValueError(f"Unknown solver: {solver} (use :pcg, :jacobi, or :chebyshev)")
        return SOLVERS[key]
    return int(solver)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _sized(a, count, name):
    """a as contiguous float64 holding exactly `count` values: the C side reads
    and writes that many through the raw pointer, so a wrong size is refused
    here rather than overrunning the array."""
    a = _f64(a)
    if a.size != count:
        raise ValueError(f"{name}: expected {count} values, got {a.size} (shape {a.shape})")
    return a


def _prm4(prm):
    """Objective parameters as the 4 doubles the C side reads (zero-padded)."""
    if prm is None:
        return None
    p = np.ascontiguousarray(prm, dtype=np.float64).ravel()
    if p.size > 4:
        raise ValueError(f"objective parameters: at most 4 values, got {p.size}")
    return np.concatenate([p, np.zeros(4 - p.size)])


class DGBasis:
    """Polynomial degree and LGL node count (DGBasis, IGR_adjoints_2d/basis.jl:7-12)."""

    def __init__(self, p):
        if not 1 <= p <= 4:
            raise ValueError("p must be in 1..4")
        self.p = int(p)
        self.np = self.p + 1


class CartesianMesh2D:
    """Uniform doubly periodic mesh (IGR_adjoints_2d/mesh.jl:21-57)."""

    dim = 2

    def __init__(self, N_ex, N_ey, Lx, Ly, basis):
        self.N_ex, self.N_ey, self.Lx, self.Ly = int(N_ex), int(N_ey), float(Lx), float(Ly)
        self.basis = basis
        self.dx, self.dy = self.Lx / self.N_ex, self.Ly / self.N_ey
        self.Jx, self.Jy = self.dx / 2.0, self.dy / 2.0
        self.shape = (self.N_ey, self.N_ex, basis.np, basis.np)
        self.nv = 4


class PeriodicMesh1D:
    """Uniform periodic 1D mesh (IGR_adjoints_1d/mesh.jl:7-33)."""

    dim = 1

    def __init__(self, N_e, L, basis):
        self.N_e, self.L = int(N_e), float(L)
        self.basis = basis
        self.dx = self.L / self.N_e
        self.J = self.dx / 2.0
        self.shape = (self.N_e, basis.np)
        self.nv = 3


class Context:
    """One device context (mesh + HBM-resident state + HIP stream)."""

    def __init__(self, mesh):
        L = _lib.lib()
        self._L = L
        self.mesh = mesh
        h = C.c_void_p()
        if mesh.dim == 2:
            rc = L.igr_create(C.byref(h), 2, mesh.basis.p, mesh.N_ex, mesh.N_ey, mesh.Lx, mesh.Ly)
        else:
            rc = L.igr_create(C.byref(h), 1, mesh.basis.p, mesh.N_e, 1, mesh.L, 1.0)
        if rc != 0:
            raise IGRError(f"igr_create failed with status {rc}")
        self.h = h
        self.shape = mesh.shape
        self.nv = mesh.nv
        self.N = int(np.prod(mesh.shape))

    def close(self):
        if self.h:
            self._L.igr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc == 0:
            return
        msg = self._L.igr_last_error(self.h)
        msg = msg.decode() if msg else ""
        if rc == _lib.IGR_EDOMAIN:
            raise DomainError(msg)
        if rc == _lib.IGR_EINVAL:
            raise ValueError(msg or "invalid argument")
        raise IGRError(f"status {rc}: {msg}")

    def set_option(self, option, value):
        self.check(self._L.igr_set_option(self.h, int(option), int(value)))

    def get_option(self, option):
        v = C.c_int(0)
        self.check(self._L.igr_get_option(self.h, int(option), C.byref(v)))
        return v.value

    def autotune(self):
        """Time the tile shapes on the current state and keep the fastest (2D, p = 3 and 4);
        returns (SIP tile, hyperbolic tile) as tile_x*100 + tile_y."""
        self.set_option(OPT_AUTOTUNE, 0)
        return self.get_option(OPT_TILE), self.get_option(OPT_HTILE)

    # state transfer
    def set_state(self, q, sigma=None):
        q = _sized(q, self.nv * self.N, "state q")
        s = None if sigma is None else _sized(sigma, self.N, "Sigma")
        self.check(self._L.igr_set_state(self.h, _p(q), _p(s)))

    def get_state(self):
        q = np.empty((self.nv,) + self.shape)
        s = np.empty(self.shape)
        self.check(self._L.igr_get_state(self.h, _p(q), _p(s)))
        return q, s

    # operators on host arrays
    def hyperbolic_rhs(self, gamma, q, sigma):
        q, sigma = _sized(q, self.nv * self.N, "state q"), _sized(sigma, self.N, "Sigma")
        dq = np.empty_like(q)
        self.check(self._L.igr_hyperbolic_rhs(self.h, float(gamma), _p(q), _p(sigma), _p(dq)))
        return dq

    def apply_sip(self, alpha, x, rho):
        x, rho = _sized(x, self.N, "x"), _sized(rho, self.N, "rho")
        y = np.empty_like(x)
        self.check(self._L.igr_apply_sip(self.h, float(alpha), _p(x), _p(rho), _p(y)))
        return y

    def igr_rhs(self, alpha, q):
        q = _sized(q, self.nv * self.N, "state q")
        b = np.empty(self.shape)
        self.check(self._L.igr_igr_rhs(self.h, float(alpha), _p(q), _p(b)))
        return b

    def sip_diagonal(self, alpha, rho):
        rho = _sized(rho, self.N, "rho")
        d = np.empty(self.shape)
        self.check(self._L.igr_sip_diagonal(self.h, float(alpha), _p(rho), _p(d)))
        return d

    # operators on the context state
    def solve_elliptic(self, alpha, n_iter, solver):
        self.check(self._L.igr_solve_elliptic(self.h, float(alpha), int(n_iter), _solver_code(solver)))

    def solver_trace(self, n_iter):
        a, b = np.empty(n_iter), np.empty(n_iter)
        self.check(self._L.igr_solver_trace(self.h, int(n_iter), _p(a), _p(b)))
        return a, b

    def euler_rhs(self, gamma, alpha, n_iter, solver):
        dq = np.empty((self.nv,) + self.shape)
        self.check(self._L.igr_euler_rhs(self.h, float(gamma), float(alpha), int(n_iter), _solver_code(solver),
                                         _p(dq)))
        return dq

    def run_forward(self, dt, T, gamma, alpha, n_iter, solver):
        ns = C.c_int(0)
        self.check(self._L.igr_run_forward(self.h, float(dt), float(T), float(gamma), float(alpha), int(n_iter),
                                           _solver_code(solver), C.byref(ns)))
        return ns.value

    def run_steps(self, n_steps, dt, gamma, alpha, n_iter, solver):
        self.check(self._L.igr_run_steps(self.h, int(n_steps), float(dt), float(gamma), float(alpha), int(n_iter),
                                         _solver_code(solver)))

    # forward tangents
    def run_sensitivity(self, s0, dt, T, gamma, alpha, n_iter, solver="pcg", form=2, sgam=0.0):
        """Advance the context state and tangent states s0 (n_tan, nv, ...) to T; returns (s, varsigma, n_steps)."""
        s0 = _f64(s0)
        if s0.ndim == len(self.shape) + 1:
            s0 = s0[None]
        nt = s0.shape[0]
        s0 = _sized(s0, nt * self.nv * self.N, "tangent states s0")
        so = np.empty_like(s0)
        vo = np.empty((nt,) + self.shape)
        ns = C.c_int(0)
        self.check(self._L.igr_run_sensitivity(self.h, nt, _p(s0), float(dt), float(T), float(gamma), float(alpha),
                                               int(n_iter), _solver_code(solver), int(form), float(sgam), _p(so),
                                               _p(vo), C.byref(ns)))
        return so, vo, ns.value

    # discrete adjoint
    def objective(self, obj, gamma, prm=None):
        J = C.c_double(0)
        p = _prm4(prm)
        self.check(self._L.igr_objective(self.h, int(obj), _p(p), float(gamma), C.byref(J)))
        return J.value

    def build_ic(self, ic, A=0.0, gamma=1.4, prm=None, dA=False):
        """Initial condition built on the device into the context state (the
        reference's IC builders: IGR_adjoints_2d/initial_conditions.jl,
        IGR_adjoints_1d/initial_conditions.jl; names in IC_TYPES).  prm=None
        takes the reference's keyword defaults.  With dA=True returns
        d q0 / d A (conservative variables), the tangent seed of the
        amplitude."""
        p = None if prm is None else np.ascontiguousarray(prm, dtype=np.float64).ravel()
        dq = np.empty((self.nv,) + self.shape) if dA else None
        self.check(self._L.igr_build_ic(self.h, _ic_code(ic), float(A), float(gamma), _p(p),
                                        0 if p is None else int(p.size), _p(dq)))
        return dq

    def gradient(self, obj, n_steps, dt, gamma, alpha, n_iter, solver="pcg", prm=None, ckpt_slots=0):
        """(J, dJ/dq0, dJ/dgamma, dJ/dalpha) of the objective after n_steps fixed steps from the state."""
        J, gb, ab = C.c_double(0), C.c_double(0), C.c_double(0)
        qb = np.empty((self.nv,) + self.shape)
        p = _prm4(prm)
        self.check(self._L.igr_gradient(self.h, int(obj), _p(p), int(n_steps), float(dt), float(gamma),
                                        float(alpha), int(n_iter), _solver_code(solver), int(ckpt_slots),
                                        C.byref(J), _p(qb), C.byref(gb), C.byref(ab)))
        return J.value, qb, gb.value, ab.value

    def hyperbolic_vjp(self, gamma, q, sigma, dqbar):
        nvN = self.nv * self.N
        q, sigma, dqbar = _sized(q, nvN, "state q"), _sized(sigma, self.N, "Sigma"), _sized(dqbar, nvN, "dq-bar")
        qb, sb, gb = np.empty_like(q), np.empty(self.shape), C.c_double(0)
        self.check(self._L.igr_hyperbolic_vjp(self.h, float(gamma), _p(q), _p(sigma), _p(dqbar), _p(qb), _p(sb),
                                              C.byref(gb)))
        return qb, sb, gb.value

    def apply_sip_vjp(self, alpha, x, rho, ybar):
        x, rho, ybar = _sized(x, self.N, "x"), _sized(rho, self.N, "rho"), _sized(ybar, self.N, "y-bar")
        xb, rb, ab = np.empty(self.shape), np.empty(self.shape), C.c_double(0)
        self.check(self._L.igr_apply_sip_vjp(self.h, float(alpha), _p(x), _p(rho), _p(ybar), _p(xb), _p(rb),
                                             C.byref(ab)))
        return xb, rb, ab.value

    def solve_elliptic_vjp(self, alpha, n_iter, solver, sigma0, sigmabar):
        s0, sb = _sized(sigma0, self.N, "sigma0"), _sized(sigmabar, self.N, "Sigma-bar")
        qb, s0b, ab = np.empty((self.nv,) + self.shape), np.empty(self.shape), C.c_double(0)
        self.check(self._L.igr_solve_elliptic_vjp(self.h, float(alpha), int(n_iter), _solver_code(solver), _p(s0),
                                                  _p(sb), _p(qb), _p(s0b), C.byref(ab)))
        return qb, s0b, ab.value

    # continuous adjoint
    def run_forward_store(self, dt, T, gamma, alpha, n_iter, solver="pcg"):
        """run_forward_store from the context state; the snapshots stay on the device. Returns n_steps."""
        ns = C.c_int(0)
        self.check(self._L.igr_run_forward_store(self.h, float(dt), float(T), float(gamma), float(alpha),
                                                 int(n_iter), _solver_code(solver), C.byref(ns)))
        return ns.value

    def snapshot_count(self):
        n = C.c_int(0)
        self.check(self._L.igr_snapshot_count(self.h, C.byref(n)))
        return n.value

    def get_snapshot(self, k):
        """(q, Sigma) of record k (0: after the initial solve, k: after step k)."""
        q, S = np.empty((self.nv,) + self.shape), np.empty(self.shape)
        self.check(self._L.igr_get_snapshot(self.h, int(k), _p(q), _p(S)))
        return q, S

    def set_snapshots(self, records):
        """records: array (n_steps + 1, nv + 1, ...) of (state, Sigma)."""
        r = _f64(records)
        if r.ndim < 1 or r.shape[0] < 1:
            raise ValueError("snapshots: need at least one record")
        r = _sized(r, r.shape[0] * (self.nv + 1) * self.N, "snapshot records")
        self.check(self._L.igr_set_snapshots(self.h, r.shape[0] - 1, _p(r)))

    def run_adjoint_conservative(self, obj, dt, T, gamma, alpha, n_iter, solver="pcg", n_steps=None, prm=None):
        """Adjoint fields at t = 0 (nv, ...) from the stored snapshots."""
        if n_steps is None:
            n_steps = self.snapshot_count()
        a = np.empty((self.nv,) + self.shape)
        p = _prm4(prm)
        self.check(self._L.igr_run_adjoint_conservative(self.h, int(obj), _p(p), int(n_steps), float(dt), float(T),
                                                        float(gamma), float(alpha), int(n_iter),
                                                        _solver_code(solver), _p(a)))
        return a

    def time_kernel(self, kernel, gamma, alpha, reps):
        ms = C.c_double(0)
        self.check(self._L.igr_time_kernel(self.h, int(kernel), float(gamma), float(alpha), int(reps),
                                           C.byref(ms)))
        return ms.value


_CTX = {}


def _ctx(mesh):
    key = id(mesh)
    c = _CTX.get(key)
    if c is None or c.mesh is not mesh:
        c = Context(mesh)
        _CTX[key] = (c)
    return c


def _stack(*fields):
    return np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.float64) for f in fields]))


# ---------------------------------------------------------------- mirror API
def compute_hyperbolic_rhs(q, Sigma, basis, mesh, gamma):
    """compute_hyperbolic_rhs! (IGR_adjoints_2d/hyperbolic.jl:27; 1D :23); returns dq."""
    return _ctx(mesh).hyperbolic_rhs(gamma, q, Sigma)


def apply_sip_(y, x, rho, basis, mesh, alpha):
    """apply_sip!(y, x, rho, basis, mesh, alpha) (IGR_adjoints_2d/elliptic.jl:30)."""
    y[...] = _ctx(mesh).apply_sip(alpha, x, rho)


def compute_igr_rhs_(b, q, basis, mesh, alpha):
    """compute_igr_rhs! (IGR_adjoints_2d/elliptic.jl:159; 1D :126)."""
    b[...] = _ctx(mesh).igr_rhs(alpha, q)


def compute_sip_diagonal(rho, basis, mesh, alpha):
    """compute_sip_diagonal (IGR_adjoints_2d/elliptic.jl:205; 1D :206)."""
    return _ctx(mesh).sip_diagonal(alpha, rho)


def solve_elliptic_(Sigma, q, basis, mesh, alpha, n_iter=10, solver="pcg"):
    """solve_elliptic! (IGR_adjoints_2d/elliptic.jl:413): Sigma is the warm start, overwritten."""
    c = _ctx(mesh)
    code = _solver_code(solver)
    c.set_state(q, Sigma)
    c.solve_elliptic(alpha, n_iter, code)
    Sigma[...] = c.get_state()[1]
    return n_iter


def compute_igr_euler_rhs_(dq, q, Sigma, basis, mesh, gamma, alpha, n_iter=10, solver="pcg"):
    """compute_igr_euler_rhs! (IGR_adjoints_2d/forward.jl:22; 1D :23)."""
    c = _ctx(mesh)
    code = _solver_code(solver)
    c.set_state(q, Sigma)
    dq[...] = c.euler_rhs(gamma, alpha, n_iter, code)
    Sigma[...] = c.get_state()[1]


def run_forward_(q, Sigma, dt, T, basis, mesh, gamma, alpha, n_iter=10, solver="pcg"):
    """run_forward! (IGR_adjoints_2d/forward.jl:41; 1D :50): q and Sigma overwritten."""
    c = _ctx(mesh)
    code = _solver_code(solver)
    c.set_state(q)
    c.run_forward(dt, T, gamma, alpha, n_iter, code)
    qo, so = c.get_state()
    q[...] = qo
    Sigma[...] = so


def run_sensitivity(q0, s0, dt, T, basis, mesh, gamma, alpha, n_iter=10, solver="pcg", elliptic_rhs="form2",
                    sgamma=0.0):
    """run_sensitivity (IGR_adjoints_2d/sensitivity.jl:353; 1D :475): returns
    (q, Sigma, s, varsigma) at T.  elliptic_rhs applies in 1D only; sgamma is
    the gamma source of a d/dgamma tangent (1D as the reference; in 2D the
    same term, which the reference's 2D tangent does not have)."""
    c = _ctx(mesh)
    code = _solver_code(solver)
    form = {"form1": 1, "form2": 2}.get(str(elliptic_rhs).lstrip(":"))
    if form is None:
        raise ValueError(f"Unknown elliptic_rhs: {elliptic_rhs} (use :form1 or :form2)")
    c.set_state(q0)
    s, vs, _ = c.run_sensitivity(s0, dt, T, gamma, alpha, n_iter, code, form, sgamma)
    q, S = c.get_state()
    return q, S, s[0], vs[0]


def run_sensitivity_batch(runs, n_iter=10, solver="pcg", elliptic_rhs="form2"):
    """run_sensitivity for many instances in one batched GPU run.

    runs: list of dicts with keys ctx (a Context whose state is the primal
    initial state), s0 (n_tan, nv, ...) tangent initial states (or None),
    dt, T, gamma, alpha and sgamma.  Each context's state ends at
    (q(T), Sigma(T)); returns [(s, varsigma, n_steps)] per run, each exactly
    what Context.run_sensitivity gives for that run alone.  There is no
    reference counterpart: the reference loops over run_sensitivity calls."""
    code = _solver_code(solver)
    form = {"form1": 1, "form2": 2}.get(str(elliptic_rhs).lstrip(":"))
    if form is None:
        raise ValueError(f"Unknown elliptic_rhs: {elliptic_rhs} (use :form1 or :form2)")
    n = len(runs)
    items = (_lib.BatchItem * n)()
    keep, outs = [], []
    for k, r in enumerate(runs):
        c = r["ctx"]
        s0 = r.get("s0")
        if s0 is None:
            s0 = np.empty((0, c.nv) + c.shape)
        s0 = _f64(s0)
        if s0.ndim == len(c.shape) + 1:
            s0 = s0[None]
        nt = s0.shape[0]
        s0 = _sized(s0, nt * c.nv * c.N, "tangent states s0")
        so = np.empty_like(s0)
        vo = np.empty((nt,) + c.shape)
        keep += [s0, so, vo]
        outs.append((so, vo))
        it = items[k]
        it.ctx = c.h.value if isinstance(c.h, C.c_void_p) else c.h
        it.dt, it.t_end = float(r["dt"]), float(r["T"])
        it.gamma, it.alpha = float(r["gamma"]), float(r["alpha"])
        it.sgam = float(r.get("sgamma", 0.0))
        it.n_tan = nt
        it.s0 = _p(s0) if nt else None
        it.s_out = _p(so) if nt else None
        it.vs_out = _p(vo) if nt else None
    L = _lib.lib()
    rc = L.igr_batch_run_sensitivity(items, n, int(n_iter), code, form)
    if rc != 0:
        bad = next((k for k in range(n) if items[k].status != 0), 0)
        runs[bad]["ctx"].check(items[bad].status or rc)
        raise IGRError(f"batched run failed with status {rc}")
    return [(outs[k][0], outs[k][1], items[k].n_steps) for k in range(n)]


OBJECTIVES = {"l2_density": 0, "l2_pressure": 1, "kinetic_energy": 2, "windowed_kinetic_energy": 3,
              "weighted_momentum": 4, "windowed_density": 5, "disc_kinetic_energy": 6}


def enzyme_adjoint_ic(q0, n_steps, dt, basis, mesh, gamma, alpha, n_iter=10, solver="pcg",
                      objective="l2_density", prm=None, ckpt_slots=0):
    """enzyme_adjoint_ic on the GPU discrete adjoint.

    2D (IGR_adjoints_2d/enzyme_adjoint.jl:270-296): returns (J, dq0) with
    dq0 = dJ/dq0 stacked (mx, my, rho, E).  1D (IGR_adjoints_1d/
    enzyme_adjoint.jl:255-314): returns (J, adj_rho, adj_u, adj_p), the
    conservative adjoints mapped to primitive ones as the reference does
    (:306-311).  Like the reference, n_iter is clamped to >= 1."""
    key = objective.lstrip(":") if isinstance(objective, str) else objective
    if isinstance(key, str):
        if key not in OBJECTIVES:
            raise ValueError(f"Unknown objective: {objective}")
        key = OBJECTIVES[key]
    c = _ctx(mesh)
    code = _solver_code(solver)
    c.set_state(q0)
    J, qb, _, _ = c.gradient(key, n_steps, dt, gamma, alpha, max(int(n_iter), 1), code, prm, ckpt_slots)
    if mesh.dim == 1:
        q0 = np.asarray(q0, dtype=np.float64)
        u0 = q0[0] / q0[1]
        adj_rho = qb[1] + u0 * qb[0] + (u0 ** 2 / 2) * qb[2]
        adj_u = q0[1] * qb[0] + q0[1] * u0 * qb[2]
        adj_p = qb[2] / (gamma - 1)
        return J, adj_rho, adj_u, adj_p
    return J, qb


def run_forward(q0, dt, T, basis, mesh, gamma, alpha, n_iter=10, solver="pcg"):
    """run_forward (IGR_adjoints_2d/forward.jl:84): returns (q, Sigma) at T."""
    q = np.array(q0, dtype=np.float64, copy=True)
    S = np.zeros(mesh.shape)
    run_forward_(q, S, dt, T, basis, mesh, gamma, alpha, n_iter, solver)
    return q, S


def release_cache():
    """igr_release_cache: free the gradient's kept-iterate block that outlives contexts."""
    _lib.lib().igr_release_cache()


class Snapshots:
    """The `snapshots` vector of run_forward_store, kept on the device of the
    context that made it: snaps[k] is (q, Sigma) after step k (k = 0: the
    initial state and its Sigma), copied to the host on access."""

    def __init__(self, ctx, n_steps):
        self.ctx, self.n_steps = ctx, n_steps

    def __len__(self):
        return self.n_steps + 1

    def __getitem__(self, k):
        k = range(len(self))[k]
        return self.ctx.get_snapshot(k)

    def array(self):
        """All records as one array (n_steps + 1, nv + 1, ...)."""
        out = np.empty((len(self), self.ctx.nv + 1) + self.ctx.shape)
        for k in range(len(self)):
            q, S = self[k]
            out[k, :-1], out[k, -1] = q, S
        return out


# IC names: the reference's IC_BUILDERS symbols and init_* functions
IC_TYPES = {"constant": 0, "isentropic_vortex": 1, "taylor_green": 2, "sedov_blast": 3, "oseen_pair_blast": 4,
            "init_entropy_wave": 5, "acoustic_pulse": 6, "entropy_wave": 7, "density_pulse": 8,
            "pressure_pulse": 9, "velocity_pulse": 10}


def _ic_code(ic):
    key = ic.lstrip(":") if isinstance(ic, str) else ic
    if isinstance(key, str):
        if key not in IC_TYPES:
            raise ValueError(f"Unknown IC type: {ic}")
        key = IC_TYPES[key]
    return int(key)


def _objective_code(objective):
    key = objective.lstrip(":") if isinstance(objective, str) else objective
    if isinstance(key, str):
        if key not in OBJECTIVES:
            raise ValueError(f"Unknown objective: {objective}")
        key = OBJECTIVES[key]
    return int(key)


def run_forward_store(q0, dt, T, basis, mesh, gamma, alpha, n_iter=10, solver="pcg"):
    """run_forward_store (IGR_adjoints_2d/forward.jl:111; 1D forward.jl:137):
    returns (snapshots, n_steps); the snapshots stay in HBM (Snapshots)."""
    c = Context(mesh)
    c.set_state(q0)
    n = c.run_forward_store(dt, T, gamma, alpha, n_iter, solver)
    return Snapshots(c, n), n


def build_ic(ic, A, basis, mesh, gamma, prm=None, dA=False):
    """The reference's IC builders (build_<type>_ic!(..., A, mesh, gamma) and
    the init_* functions; names in IC_TYPES) evaluated on the device: returns
    the state stacked as (mx, my, rho, E) / (mu, rho, E), and with dA=True
    also d q0 / d A."""
    c = _ctx(mesh)
    dq = c.build_ic(ic, A, gamma, prm, dA)
    q, _ = c.get_state()
    return (q, dq) if dA else q


def run_adjoint_conservative(snapshots, n_steps, dt, T, basis, mesh, gamma, alpha, n_iter=10,
                             objective="l2_density", solver="pcg", prm=None):
    """run_adjoint_conservative (IGR_adjoints_2d/adjoint_pde.jl:522; 1D
    adjoint_pde.jl:466): the adjoint fields (stacked conservative variables)
    at t = 0.  `snapshots` is what run_forward_store returned, or an array
    (n_steps + 1, nv + 1, ...) of (state, Sigma) records."""
    code = _objective_code(objective)
    if isinstance(snapshots, Snapshots):
        c = snapshots.ctx
    else:
        c = Context(mesh)
        c.set_snapshots(np.asarray(snapshots, dtype=np.float64))
    return c.run_adjoint_conservative(code, dt, T, gamma, alpha, n_iter, solver, n_steps, prm)
