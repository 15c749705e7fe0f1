"""ctypes front end of the CPU oracle — TEST INFRASTRUCTURE ONLY.

The oracle is a plain-loop C++ restatement of the reference Julia packages
(f-t-s/IGR_sensitivities: IGR_adjoints_1d/, IGR_adjoints_2d/); see
igr_oracle.hpp for the file:line each routine follows.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

Arrays are float64 numpy arrays in the reference layout: a 2D field has
shape (Ney, Nex, np, np) in C order (== Julia's column-major
(np, np, Nex, Ney)); a state stacks its conserved variables on axis 0.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

PCG, JACOBI, CHEBYSHEV = 0, 1, 2
OBJ_L2_DENSITY, OBJ_L2_PRESSURE, OBJ_KINETIC_ENERGY = 0, 1, 2
OBJ_WINDOWED_KE, OBJ_WEIGHTED_MOMENTUM, OBJ_WINDOWED_DENSITY, OBJ_DISC_KE = 3, 4, 5, 6
KE_WINDOW = (0.55, 0.50, 0.15)   # IGR_adjoints_2d/enzyme_adjoint.jl:26


def build(force=False):
    """Compile the oracle libraries (make); force=True recompiles them even
    when they are up to date."""
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


_OMP_PATH = os.path.join(_HERE, "_build", "liboracle_omp.so")
_omp_lib = None


def omp_lib():
    """The OpenMP build of the same oracle (identical results; used to time
    the multi-core CPU baseline, never to check results)."""
    global _omp_lib
    if _omp_lib is None:
        if not os.path.exists(_OMP_PATH):
            build()
        _omp_lib = C.CDLL(_OMP_PATH)
    return _omp_lib


def run_steps2_omp(g, gamma, alpha, n_iter, solver, n_steps, dt, q0):
    """Grid2D.run_steps on the OpenMP build: (q, S, threads)."""
    L = omp_lib()
    q = q0.copy()
    S = g.field()
    _chk(L.orc2_run_steps(*g._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _i(n_steps),
                          _d(dt), _p(q), _p(S)))
    return q, S, L.orc_max_threads()


class OracleError(RuntimeError):
    pass


class adj_mutation:
    """Context manager that plants one deliberate mistake in the oracle's
    continuous adjoint (see igr_oracle_adjpde.hpp), for tests that show their
    checks reject it."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        lib().orc_set_adj_mutation(C.c_int(self.m))
        return self

    def __exit__(self, *exc):
        lib().orc_set_adj_mutation(C.c_int(0))
        return False


class fwd_mutation:
    """Context manager that plants one deliberate mistake in the oracle's
    forward map (1: Jacobi omega 0.5; 2: Chebyshev lambda_max 2.2 at p = 3;
    3: flux_y with its momentum components swapped), for tests that show
    their checks reject it."""

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        lib().orc_set_fwd_mutation(C.c_int(self.m))
        return self

    def __exit__(self, *exc):
        lib().orc_set_fwd_mutation(C.c_int(0))
        return False


class DomainError(OracleError):
    """Negative pressure under sqrt (Julia DomainError)."""


def _chk(rc):
    if rc == 1:
        raise DomainError("negative pressure under sqrt")
    if rc != 0:
        raise OracleError(f"oracle returned {rc} (unknown solver/objective)")


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


_d = C.c_double
_i = C.c_int


def basis(p):
    n = p + 1
    xi, w, D = np.zeros(n), np.zeros(n), np.zeros((n, n))
    _chk(lib().orc_basis(_i(p), _p(xi), _p(w), _p(D)))
    return xi, w, D


class Grid2D:
    """Geometry handle; mirrors CartesianMesh2D (IGR_adjoints_2d/mesh.jl:40)."""

    def __init__(self, p, Nex, Ney, Lx=1.0, Ly=1.0):
        self.p, self.Nex, self.Ney, self.Lx, self.Ly = p, Nex, Ney, float(Lx), float(Ly)
        self.np = p + 1
        self.N = self.np * self.np * Nex * Ney
        self.shape = (Ney, Nex, self.np, self.np)
        self.dx, self.dy = self.Lx / Nex, self.Ly / Ney
        self.Jx, self.Jy = self.dx / 2.0, self.dy / 2.0
        self.xi, self.w, self.D = basis(p)

    def _g(self):
        return (_i(self.p), _i(self.Nex), _i(self.Ney), _d(self.Lx), _d(self.Ly))

    def coords(self):
        ex = np.arange(self.Nex)[None, :, None, None]
        ey = np.arange(self.Ney)[:, None, None, None]
        i = self.xi[None, None, None, :]
        j = self.xi[None, None, :, None]
        x = ex * self.dx + self.Jx * (1.0 + i)
        y = ey * self.dy + self.Jy * (1.0 + j)
        return np.broadcast_to(x, self.shape).copy(), np.broadcast_to(y, self.shape).copy()

    def quad_weights(self):
        wij = self.w[None, :] * self.w[:, None]   # [j, i] = w_i * w_j
        return np.broadcast_to(wij * self.Jx * self.Jy, self.shape).copy()

    def field(self):
        return np.zeros(self.shape)

    def state(self):
        return np.zeros((4,) + self.shape)

    # --- restated reference operators ---
    def hyperbolic_rhs(self, gamma, q, S):
        dq = np.zeros_like(q)
        _chk(lib().orc2_hyperbolic_rhs(*self._g(), _d(gamma), _p(q), _p(S), _p(dq)))
        return dq

    def apply_sip(self, alpha, x, rho):
        y = np.zeros_like(x)
        _chk(lib().orc2_apply_sip(*self._g(), _d(alpha), _p(x), _p(rho), _p(y)))
        return y

    def igr_rhs(self, alpha, q):
        b = self.field()
        _chk(lib().orc2_igr_rhs(*self._g(), _d(alpha), _p(q), _p(b)))
        return b

    def sip_diag(self, alpha, rho):
        M = self.field()
        _chk(lib().orc2_sip_diag(*self._g(), _d(alpha), _p(rho), _p(M)))
        return M

    def solver(self, alpha, solver, n_iter, rho, b, M, x0, trace=False):
        x = x0.copy()
        ta, tb = np.zeros(max(n_iter, 1)), np.zeros(max(n_iter, 1))
        _chk(lib().orc2_solver(*self._g(), _d(alpha), _i(solver), _i(n_iter), _p(rho), _p(b), _p(M),
                               _p(x), _p(ta), _p(tb)))
        return (x, ta[:n_iter], tb[:n_iter]) if trace else x

    def solve_elliptic(self, alpha, n_iter, solver, q, S0):
        S = S0.copy()
        _chk(lib().orc2_solve_elliptic(*self._g(), _d(alpha), _i(n_iter), _i(solver), _p(q), _p(S)))
        return S

    def euler_rhs(self, gamma, alpha, n_iter, solver, q, S0):
        S = S0.copy()
        dq = np.zeros_like(q)
        _chk(lib().orc2_euler_rhs(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _p(q), _p(S),
                                  _p(dq)))
        return dq, S

    def run_forward(self, gamma, alpha, n_iter, solver, dt, T, q0):
        q = q0.copy()
        S = self.field()
        ns = _i(0)
        _chk(lib().orc2_run_forward(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _d(dt), _d(T),
                                    _p(q), _p(S), C.byref(ns)))
        return q, S, ns.value

    def run_steps(self, gamma, alpha, n_iter, solver, n_steps, dt, q0, omp=False):
        """omp=True runs the OpenMP build of the same loops (the same bits,
        tests/test_oracle_omp.py), for the large meshes."""
        q = q0.copy()
        S = self.field()
        L = omp_lib() if omp else lib()
        _chk(L.orc2_run_steps(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _i(n_steps),
                              _d(dt), _p(q), _p(S)))
        return q, S

    def objective(self, obj, gamma, q, prm=(0.0, 0.0, 1.0, 0.0)):
        prm = np.ascontiguousarray(prm, dtype=np.float64)
        J = _d(0)
        _chk(lib().orc2_objective(*self._g(), _i(obj), _p(prm), _d(gamma), _p(q), C.byref(J)))
        return J.value

    def run_sensitivity(self, gamma, alpha, n_iter, solver, dt, T, q0, s0, sgam=0.0):
        """sgam: coefficient of the gamma source (d/dgamma tangent); the reference's 2D tangent has none."""
        q, s = q0.copy(), s0.copy()
        S, vs = self.field(), self.field()
        ns = _i(0)
        _chk(lib().orc2_run_sensitivity(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _d(dt),
                                        _d(T), _p(q), _p(S), _p(s), _p(vs), C.byref(ns), _d(sgam)))
        return q, S, s, vs, ns.value

    def sens_elliptic_rhs(self, alpha, q, S, s):
        b = self.field()
        _chk(lib().orc2_sens_elliptic_rhs(*self._g(), _d(alpha), _p(q), _p(S), _p(s), _p(b)))
        return b

    def sens_hyperbolic_rhs(self, gamma, q, S, s, vs, sgam=0.0):
        ds = np.zeros_like(s)
        _chk(lib().orc2_sens_hyperbolic_rhs(*self._g(), _d(gamma), _p(q), _p(S), _p(s), _p(vs), _p(ds),
                                            _d(sgam)))
        return ds

    # --- exact discrete JVPs ---
    def jvp_hyperbolic_rhs(self, gamma, dgamma, q, dq, S, dS):
        o, do = np.zeros_like(q), np.zeros_like(q)
        _chk(lib().orc2_jvp_hyperbolic_rhs(*self._g(), _d(gamma), _d(dgamma), _p(q), _p(dq), _p(S), _p(dS),
                                           _p(o), _p(do)))
        return o, do

    def jvp_apply_sip(self, alpha, dalpha, x, dx, rho, drho):
        y, dy = self.field(), self.field()
        _chk(lib().orc2_jvp_apply_sip(*self._g(), _d(alpha), _d(dalpha), _p(x), _p(dx), _p(rho), _p(drho),
                                      _p(y), _p(dy)))
        return y, dy

    def jvp_igr_rhs(self, alpha, dalpha, q, dq):
        b, db = self.field(), self.field()
        _chk(lib().orc2_jvp_igr_rhs(*self._g(), _d(alpha), _d(dalpha), _p(q), _p(dq), _p(b), _p(db)))
        return b, db

    def jvp_sip_diag(self, alpha, dalpha, rho, drho):
        M, dM = self.field(), self.field()
        _chk(lib().orc2_jvp_sip_diag(*self._g(), _d(alpha), _d(dalpha), _p(rho), _p(drho), _p(M), _p(dM)))
        return M, dM

    def jvp_solver(self, alpha, dalpha, solver, n_iter, rho, drho, b, db, M, dM, x0, dx0):
        x, dx = x0.copy(), dx0.copy()
        _chk(lib().orc2_jvp_solver(*self._g(), _d(alpha), _d(dalpha), _i(solver), _i(n_iter), _p(rho),
                                   _p(drho), _p(b), _p(db), _p(M), _p(dM), _p(x), _p(dx)))
        return x, dx

    def jvp_solve_elliptic(self, alpha, dalpha, n_iter, solver, q, dq, S0, dS0):
        S, dS = S0.copy(), dS0.copy()
        _chk(lib().orc2_jvp_solve_elliptic(*self._g(), _d(alpha), _d(dalpha), _i(n_iter), _i(solver), _p(q),
                                           _p(dq), _p(S), _p(dS)))
        return S, dS

    def jvp_step(self, gamma, dgamma, alpha, dalpha, n_iter, solver, dt, q0, dq0, S0, dS0):
        q, dq, S, dS = q0.copy(), dq0.copy(), S0.copy(), dS0.copy()
        _chk(lib().orc2_jvp_step(*self._g(), _d(gamma), _d(dgamma), _d(alpha), _d(dalpha), _i(n_iter),
                                 _i(solver), _d(dt), _p(q), _p(dq), _p(S), _p(dS)))
        return q, dq, S, dS

    def jvp_objective_run(self, gamma, dgamma, alpha, dalpha, n_iter, solver, n_steps, dt, obj, q0, dq0,
                          prm=(0.0, 0.0, 1.0, 0.0), omp=False):
        q, dq = q0.copy(), dq0.copy()
        S, dS = self.field(), self.field()
        prm = np.ascontiguousarray(prm, dtype=np.float64)
        J, dJ = _d(0), _d(0)
        L = omp_lib() if omp else lib()
        _chk(L.orc2_jvp_objective_run(*self._g(), _d(gamma), _d(dgamma), _d(alpha), _d(dalpha), _i(n_iter),
                                          _i(solver), _i(n_steps), _d(dt), _i(obj), _p(prm), _p(q), _p(dq), _p(S),
                                          _p(dS), C.byref(J), C.byref(dJ)))
        return J.value, dJ.value, q, dq

    def jvp_objective(self, obj, gamma, dgamma, q, dq, prm=(0.0, 0.0, 1.0, 0.0)):
        prm = np.ascontiguousarray(prm, dtype=np.float64)
        J, dJ = _d(0), _d(0)
        _chk(lib().orc2_jvp_objective(*self._g(), _i(obj), _p(prm), _d(gamma), _d(dgamma), _p(q), _p(dq),
                                      C.byref(J), C.byref(dJ)))
        return J.value, dJ.value

    # --- continuous adjoint (IGR_adjoints_2d/adjoint_pde.jl) ---
    def adj_hyp_rhs_local(self, gamma, q, S, a):
        da = np.zeros_like(a)
        _chk(lib().orc2_adj_hyp_rhs_local(*self._g(), _d(gamma), _p(q), _p(S), _p(a), _p(da)))
        return da

    def adj_elliptic_rhs(self, a, q):
        b = self.field()
        _chk(lib().orc2_adj_elliptic_rhs(*self._g(), _p(a), _p(q), _p(b)))
        return b

    def nonlocal_source(self, Pi, q, S):
        Sn = self.state()
        _chk(lib().orc2_nonlocal_source(*self._g(), _p(Pi), _p(q), _p(S), _p(Sn)))
        return Sn

    def adj_rhs_conservative(self, gamma, alpha, n_iter, solver, q, S, a, Pi0):
        """(da, Pi): the full adjoint operator; Pi0 is the warm start of the Pi solve."""
        da, Pi = np.zeros_like(a), np.array(Pi0, dtype=np.float64)
        _chk(lib().orc2_adj_rhs_conservative(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _p(q),
                                             _p(S), _p(a), _p(Pi), _p(da)))
        return da, Pi

    def adj_terminal(self, obj, gamma, qT, prm=KE_WINDOW):
        a = self.state()
        prm = np.ascontiguousarray(prm, dtype=np.float64)
        _chk(lib().orc2_adj_terminal(*self._g(), _i(obj), _p(prm), _d(gamma), _p(qT), _p(a)))
        return a

    def run_forward_store(self, gamma, alpha, n_iter, solver, dt, T, q0):
        """run_forward_store (IGR_adjoints_2d/forward.jl:111): (snaps, n_steps), snaps of shape
        (n_steps + 1, 5, ...) holding (mx, my, rho, E, Sigma) after every step."""
        cap = int(np.ceil(T / dt)) + 1
        snaps = np.zeros((cap, 5) + self.shape)
        ns = C.c_int(0)
        _chk(lib().orc2_run_forward_store(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _d(dt),
                                          _d(T), _p(np.ascontiguousarray(q0)), _i(cap), _p(snaps), C.byref(ns)))
        return snaps[:ns.value + 1], ns.value

    def run_adjoint_conservative(self, snaps, n_steps, dt, T, gamma, alpha, n_iter, solver, obj, prm=KE_WINDOW):
        """run_adjoint_conservative (IGR_adjoints_2d/adjoint_pde.jl:522): adjoint fields at t = 0."""
        a = self.state()
        prm = np.ascontiguousarray(prm, dtype=np.float64)
        _chk(lib().orc2_run_adjoint_conservative(*self._g(), _p(np.ascontiguousarray(snaps)), _i(n_steps), _d(dt),
                                                 _d(T), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _i(obj),
                                                 _p(prm), _p(a)))
        return a


class Grid1D:
    """Geometry handle; mirrors PeriodicMesh1D (IGR_adjoints_1d/mesh.jl:21)."""

    def __init__(self, p, Ne, L=1.0):
        self.p, self.Ne, self.L = p, Ne, float(L)
        self.np = p + 1
        self.N = self.np * Ne
        self.shape = (Ne, self.np)
        self.dx = self.L / Ne
        self.J = self.dx / 2.0
        self.xi, self.w, self.D = basis(p)

    def _g(self):
        return (_i(self.p), _i(self.Ne), _d(self.L))

    def coords(self):
        e = np.arange(self.Ne)[:, None]
        return e * self.dx + self.J * (1.0 + self.xi[None, :])

    def quad_weights(self):
        return np.broadcast_to(self.w[None, :] * self.J, self.shape).copy()

    def field(self):
        return np.zeros(self.shape)

    def state(self):
        return np.zeros((3,) + self.shape)

    def hyperbolic_rhs(self, gamma, q, S):
        dq = np.zeros_like(q)
        _chk(lib().orc1_hyperbolic_rhs(*self._g(), _d(gamma), _p(q), _p(S), _p(dq)))
        return dq

    def apply_sip(self, alpha, x, rho):
        y = np.zeros_like(x)
        _chk(lib().orc1_apply_sip(*self._g(), _d(alpha), _p(x), _p(rho), _p(y)))
        return y

    def igr_rhs(self, alpha, q):
        b = self.field()
        _chk(lib().orc1_igr_rhs(*self._g(), _d(alpha), _p(q), _p(b)))
        return b

    def sip_diag(self, alpha, rho):
        M = self.field()
        _chk(lib().orc1_sip_diag(*self._g(), _d(alpha), _p(rho), _p(M)))
        return M

    def solver(self, alpha, solver, n_iter, rho, b, M, x0, trace=False):
        x = x0.copy()
        ta, tb = np.zeros(max(n_iter, 1)), np.zeros(max(n_iter, 1))
        _chk(lib().orc1_solver(*self._g(), _d(alpha), _i(solver), _i(n_iter), _p(rho), _p(b), _p(M), _p(x),
                               _p(ta), _p(tb)))
        return (x, ta[:n_iter], tb[:n_iter]) if trace else x

    def solve_elliptic(self, alpha, n_iter, solver, q, S0):
        S = S0.copy()
        _chk(lib().orc1_solve_elliptic(*self._g(), _d(alpha), _i(n_iter), _i(solver), _p(q), _p(S)))
        return S

    def euler_rhs(self, gamma, alpha, n_iter, solver, q, S0):
        S = S0.copy()
        dq = np.zeros_like(q)
        _chk(lib().orc1_euler_rhs(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _p(q), _p(S),
                                  _p(dq)))
        return dq, S

    def run_forward(self, gamma, alpha, n_iter, solver, dt, T, q0):
        q = q0.copy()
        S = self.field()
        ns = _i(0)
        _chk(lib().orc1_run_forward(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _d(dt), _d(T),
                                    _p(q), _p(S), C.byref(ns)))
        return q, S, ns.value

    def run_steps(self, gamma, alpha, n_iter, solver, n_steps, dt, q0):
        q = q0.copy()
        S = self.field()
        _chk(lib().orc1_run_steps(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _i(n_steps),
                                  _d(dt), _p(q), _p(S)))
        return q, S

    def objective(self, obj, gamma, q):
        J = _d(0)
        _chk(lib().orc1_objective(*self._g(), _i(obj), _d(gamma), _p(q), C.byref(J)))
        return J.value

    def run_sensitivity(self, gamma, alpha, n_iter, solver, dt, T, q0, s0, form=2, sgam=0.0):
        q, s = q0.copy(), s0.copy()
        S, sS = self.field(), self.field()
        ns = _i(0)
        _chk(lib().orc1_run_sensitivity(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _i(form),
                                        _d(sgam), _d(dt), _d(T), _p(q), _p(S), _p(s), _p(sS), C.byref(ns)))
        return q, S, s, sS, ns.value

    def sens_elliptic_rhs(self, alpha, form, q, S, s):
        b = self.field()
        _chk(lib().orc1_sens_elliptic_rhs(*self._g(), _d(alpha), _i(form), _p(q), _p(S), _p(s), _p(b)))
        return b

    def sens_hyperbolic_rhs(self, gamma, sgam, q, S, s, sS):
        ds = np.zeros_like(s)
        _chk(lib().orc1_sens_hyperbolic_rhs(*self._g(), _d(gamma), _d(sgam), _p(q), _p(S), _p(s), _p(sS),
                                            _p(ds)))
        return ds

    def jvp_hyperbolic_rhs(self, gamma, dgamma, q, dq, S, dS):
        o, do = np.zeros_like(q), np.zeros_like(q)
        _chk(lib().orc1_jvp_hyperbolic_rhs(*self._g(), _d(gamma), _d(dgamma), _p(q), _p(dq), _p(S), _p(dS),
                                           _p(o), _p(do)))
        return o, do

    def jvp_objective_run(self, gamma, dgamma, alpha, dalpha, n_iter, solver, n_steps, dt, obj, q0, dq0):
        q, dq = q0.copy(), dq0.copy()
        S, dS = self.field(), self.field()
        J, dJ = _d(0), _d(0)
        _chk(lib().orc1_jvp_objective_run(*self._g(), _d(gamma), _d(dgamma), _d(alpha), _d(dalpha), _i(n_iter),
                                          _i(solver), _i(n_steps), _d(dt), _i(obj), _p(q), _p(dq), _p(S), _p(dS),
                                          C.byref(J), C.byref(dJ)))
        return J.value, dJ.value, q, dq

    def jvp_run_steps(self, gamma, alpha, n_iter, solver, n_steps, dt, q0, dq0):
        q, dq = q0.copy(), dq0.copy()
        S, dS = self.field(), self.field()
        _chk(lib().orc1_jvp_sensitivity_run(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver),
                                            _i(n_steps), _d(dt), _p(q), _p(dq), _p(S), _p(dS)))
        return q, dq, S, dS

    def jvp_solve_elliptic(self, alpha, dalpha, n_iter, solver, q, dq, S0, dS0):
        S, dS = S0.copy(), dS0.copy()
        _chk(lib().orc1_jvp_solve_elliptic(*self._g(), _d(alpha), _d(dalpha), _i(n_iter), _i(solver), _p(q),
                                           _p(dq), _p(S), _p(dS)))
        return S, dS

    # --- continuous adjoint (IGR_adjoints_1d/adjoint_pde.jl) ---
    def adj_hyp_rhs_local(self, gamma, q, S, a):
        da = np.zeros_like(a)
        _chk(lib().orc1_adj_hyp_rhs_local(*self._g(), _d(gamma), _p(q), _p(S), _p(a), _p(da)))
        return da

    def adj_elliptic_rhs(self, a, q):
        b = self.field()
        _chk(lib().orc1_adj_elliptic_rhs(*self._g(), _p(a), _p(q), _p(b)))
        return b

    def nonlocal_source(self, Pi, q, S):
        Sn = self.state()
        _chk(lib().orc1_nonlocal_source(*self._g(), _p(Pi), _p(q), _p(S), _p(Sn)))
        return Sn

    def adj_rhs_conservative(self, gamma, alpha, n_iter, solver, q, S, a, Pi0):
        da, Pi = np.zeros_like(a), np.array(Pi0, dtype=np.float64)
        _chk(lib().orc1_adj_rhs_conservative(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _p(q),
                                             _p(S), _p(a), _p(Pi), _p(da)))
        return da, Pi

    def adj_terminal(self, obj, gamma, qT):
        a = self.state()
        _chk(lib().orc1_adj_terminal(*self._g(), _i(obj), _d(gamma), _p(qT), _p(a)))
        return a

    def run_forward_store(self, gamma, alpha, n_iter, solver, dt, T, q0):
        """run_forward_store (IGR_adjoints_1d/forward.jl:137): (snaps, n_steps), snaps of shape
        (n_steps + 1, 4, Ne, np) holding (mu, rho, E, Sigma)."""
        cap = int(np.ceil(T / dt)) + 1
        snaps = np.zeros((cap, 4) + self.shape)
        ns = C.c_int(0)
        _chk(lib().orc1_run_forward_store(*self._g(), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _d(dt),
                                          _d(T), _p(np.ascontiguousarray(q0)), _i(cap), _p(snaps), C.byref(ns)))
        return snaps[:ns.value + 1], ns.value

    def run_adjoint_conservative(self, snaps, n_steps, dt, T, gamma, alpha, n_iter, solver, obj):
        a = self.state()
        _chk(lib().orc1_run_adjoint_conservative(*self._g(), _p(np.ascontiguousarray(snaps)), _i(n_steps), _d(dt),
                                                 _d(T), _d(gamma), _d(alpha), _i(n_iter), _i(solver), _i(obj),
                                                 _p(a)))
        return a
