// CPU ORACLE — TEST INFRASTRUCTURE ONLY (see igr_oracle.hpp).
//
// extern "C" wrappers over the templated restatement, for ctypes.  Arrays
// are FP64 in the reference's column-major layout: 2D node (i,j,ex,ey) at
// i + np*(j + np*(ex + Nex*ey)); 1D node (i,e) at i + np*e.  A state array
// holds its conserved variables back to back (2D: mx,my,rho,E; 1D: mu,rho,E).
//
// Return codes: 0 ok, 1 negative pressure under sqrt (Julia DomainError),
// -1 unknown solver/objective code (Julia ErrorException).
#include "igr_oracle.hpp"
#include "igr_oracle_adjpde.hpp"
#ifdef IGR_ORACLE_OMP
#include <omp.h>
#endif

namespace orc {
thread_local int g_domain_error = 0;
thread_local int g_adj_mutation = 0;
thread_local int g_fwd_mutation = 0;
}
using namespace orc;

namespace {
inline int finish(int rc) {
  if (rc) return -1;
  int d = g_domain_error;
  g_domain_error = 0;
  return d ? 1 : 0;
}
template <class T> std::vector<T> to_vec(const double* a, size_t n) { return std::vector<T>(a, a + n); }
std::vector<Dual> to_dual(const double* a, const double* da, size_t n) {
  std::vector<Dual> v(n);
  for (size_t k = 0; k < n; ++k) v[k] = Dual(a[k], da ? da[k] : 0.0);
  return v;
}
void from_dual(const std::vector<Dual>& v, double* a, double* da) {
  for (size_t k = 0; k < v.size(); ++k) { if (a) a[k] = v[k].v; if (da) da[k] = v[k].d; }
}
}  // namespace

#define G2ARGS int p, int Nex, int Ney, double Lx, double Ly
#define MAKE_G2 Basis B = make_basis(p); Mesh2D M(Nex, Ney, Lx, Ly); Grid2 G(B, M); g_domain_error = 0
#define G1ARGS int p, int Ne, double L
#define MAKE_G1 Basis B = make_basis(p); Mesh1D M(Ne, L); Grid1 G(B, M); g_domain_error = 0

extern "C" {

// threads the parallel loops use (1 in the serial build)
int orc_max_threads(void) {
#ifdef IGR_ORACLE_OMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_basis(int p, double* xi, double* w, double* D) {
  if (p < 0 || p + 1 > MAXNP) return -1;
  Basis B = make_basis(p);
  for (int i = 0; i < B.np; ++i) { xi[i] = B.xi[i]; w[i] = B.w[i]; }
  for (int q = 0; q < B.np * B.np; ++q) D[q] = B.D[q];
  return 0;
}

// ------------------------------------------------------------------ 2D ----
int orc2_hyperbolic_rhs(G2ARGS, double g, const double* q, const double* S, double* dq) {
  MAKE_G2; hyperbolic_rhs2<double>(G, g, q, S, dq); return finish(0);
}
int orc2_apply_sip(G2ARGS, double alpha, const double* x, const double* rho, double* y) {
  MAKE_G2; apply_sip2<double>(G, alpha, x, rho, y); return finish(0);
}
int orc2_igr_rhs(G2ARGS, double alpha, const double* q, double* b) {
  MAKE_G2; igr_rhs2<double>(G, alpha, q, b); return finish(0);
}
int orc2_sip_diag(G2ARGS, double alpha, const double* rho, double* Md) {
  MAKE_G2; sip_diag2<double>(G, alpha, rho, Md); return finish(0);
}
// One fixed-iteration solve on given rho, b, M; x is the warm start.
// trace_a / trace_b (n_iter each, may be NULL) receive the PCG alpha/beta.
int orc2_solver(G2ARGS, double alpha, int solver, int n_iter, const double* rho, const double* b,
                const double* Md, double* x, double* trace_a, double* trace_b) {
  MAKE_G2;
  std::vector<double> xv = to_vec<double>(x, G.N), bv = to_vec<double>(b, G.N), Mv = to_vec<double>(Md, G.N);
  auto A = [&](const double* in, double* out) { apply_sip2<double>(G, alpha, in, rho, out); };
  SolverTrace tr;
  int rc = run_solver(solver, A, xv, bv, Mv, n_iter, p, &tr);
  std::copy(xv.begin(), xv.end(), x);
  for (size_t k = 0; k < tr.a.size(); ++k) { if (trace_a) trace_a[k] = tr.a[k]; if (trace_b) trace_b[k] = tr.b[k]; }
  return finish(rc);
}
int orc2_solve_elliptic(G2ARGS, double alpha, int n_iter, int solver, const double* q, double* S) {
  MAKE_G2; return finish(solve_elliptic2<double>(G, alpha, n_iter, solver, q, S));
}
int orc2_euler_rhs(G2ARGS, double g, double alpha, int n_iter, int solver, const double* q, double* S, double* dq) {
  MAKE_G2; return finish(euler_rhs2<double>(G, g, alpha, n_iter, solver, q, S, dq));
}
int orc2_run_forward(G2ARGS, double g, double alpha, int n_iter, int solver, double dt, double T,
                     double* q, double* S, int* n_steps) {
  MAKE_G2; return finish(run_forward2<double>(G, g, alpha, n_iter, solver, dt, T, q, S, n_steps));
}
int orc2_run_steps(G2ARGS, double g, double alpha, int n_iter, int solver, int n_steps, double dt,
                   double* q, double* S) {
  MAKE_G2; return finish(run_steps2<double>(G, g, alpha, n_iter, solver, n_steps, dt, q, S));
}
int orc2_objective(G2ARGS, int obj, const double* prm, double g, const double* q, double* J) {
  MAKE_G2; return finish(objective2<double>(G, obj, prm, g, q, J));
}
int orc2_run_sensitivity(G2ARGS, double g, double alpha, int n_iter, int solver, double dt, double T,
                         double* q, double* S, double* s, double* vs, int* n_steps, double sgam) {
  MAKE_G2; return finish(run_sensitivity2<double>(G, g, alpha, n_iter, solver, dt, T, q, S, s, vs, n_steps, sgam));
}
int orc2_sens_elliptic_rhs(G2ARGS, double alpha, const double* q, const double* S, const double* s, double* b) {
  MAKE_G2; sens_elliptic_rhs2<double>(G, alpha, q, S, s, b); return finish(0);
}
int orc2_sens_hyperbolic_rhs(G2ARGS, double g, const double* q, const double* S, const double* s,
                             const double* vs, double* ds, double sgam) {
  MAKE_G2; sens_hyperbolic_rhs2<double>(G, g, q, S, s, vs, ds, sgam); return finish(0);
}

// ---- exact discrete JVPs (dual numbers) ----
int orc2_jvp_hyperbolic_rhs(G2ARGS, double g, double dg, const double* q, const double* dq_in,
                            const double* S, const double* dS, double* out, double* dout) {
  MAKE_G2;
  auto qd = to_dual(q, dq_in, 4 * (size_t)G.N), Sd = to_dual(S, dS, G.N);
  std::vector<Dual> o(4 * (size_t)G.N);
  hyperbolic_rhs2<Dual>(G, Dual(g, dg), qd.data(), Sd.data(), o.data());
  from_dual(o, out, dout);
  return finish(0);
}
int orc2_jvp_apply_sip(G2ARGS, double alpha, double dalpha, const double* x, const double* dx,
                       const double* rho, const double* drho, double* y, double* dy) {
  MAKE_G2;
  auto xd = to_dual(x, dx, G.N), rd = to_dual(rho, drho, G.N);
  std::vector<Dual> o(G.N);
  apply_sip2<Dual>(G, Dual(alpha, dalpha), xd.data(), rd.data(), o.data());
  from_dual(o, y, dy);
  return finish(0);
}
int orc2_jvp_igr_rhs(G2ARGS, double alpha, double dalpha, const double* q, const double* dq_in,
                     double* b, double* db) {
  MAKE_G2;
  auto qd = to_dual(q, dq_in, 4 * (size_t)G.N);
  std::vector<Dual> o(G.N);
  igr_rhs2<Dual>(G, Dual(alpha, dalpha), qd.data(), o.data());
  from_dual(o, b, db);
  return finish(0);
}
int orc2_jvp_sip_diag(G2ARGS, double alpha, double dalpha, const double* rho, const double* drho,
                      double* Md, double* dMd) {
  MAKE_G2;
  auto rd = to_dual(rho, drho, G.N);
  std::vector<Dual> o(G.N);
  sip_diag2<Dual>(G, Dual(alpha, dalpha), rd.data(), o.data());
  from_dual(o, Md, dMd);
  return finish(0);
}
// Solver JVP on explicit (rho, b, M, x0) inputs.
int orc2_jvp_solver(G2ARGS, double alpha, double dalpha, int solver, int n_iter,
                    const double* rho, const double* drho, const double* b, const double* db,
                    const double* Md, const double* dMd, double* x, double* dx) {
  MAKE_G2;
  auto rd = to_dual(rho, drho, G.N);
  auto bd = to_dual(b, db, G.N), Mdd = to_dual(Md, dMd, G.N), xd = to_dual(x, dx, G.N);
  Dual al(alpha, dalpha);
  auto A = [&](const Dual* in, Dual* out) { apply_sip2<Dual>(G, al, in, rd.data(), out); };
  int rc = run_solver(solver, A, xd, bd, Mdd, n_iter, p, nullptr);
  from_dual(xd, x, dx);
  return finish(rc);
}
int orc2_jvp_solve_elliptic(G2ARGS, double alpha, double dalpha, int n_iter, int solver,
                            const double* q, const double* dq_in, double* S, double* dS) {
  MAKE_G2;
  auto qd = to_dual(q, dq_in, 4 * (size_t)G.N), Sd = to_dual(S, dS, G.N);
  int rc = solve_elliptic2<Dual>(G, Dual(alpha, dalpha), n_iter, solver, qd.data(), Sd.data());
  from_dual(Sd, S, dS);
  return finish(rc);
}
// One SSP-RK3 step (state q and warm start S in/out) with tangents.
int orc2_jvp_step(G2ARGS, double g, double dg, double alpha, double dalpha, int n_iter, int solver,
                  double dt, double* q, double* dq, double* S, double* dS) {
  MAKE_G2;
  const size_t nvN = 4 * (size_t)G.N;
  auto qd = to_dual(q, dq, nvN), Sd = to_dual(S, dS, G.N);
  std::vector<Dual> q1(nvN), q2(nvN), d(nvN);
  Dual gg(g, dg), al(alpha, dalpha);
  auto rhs = [&](const Dual* qs, Dual* o) { return euler_rhs2<Dual>(G, gg, al, n_iter, solver, qs, Sd.data(), o); };
  int rc = ssp_rk3_step(rhs, nvN, qd.data(), dt, q1, q2, d);
  from_dual(qd, q, dq); from_dual(Sd, S, dS);
  return finish(rc);
}
// Objective after n_steps fixed steps (the Enzyme-objective map), with tangent.
int orc2_jvp_objective_run(G2ARGS, double g, double dg, double alpha, double dalpha, int n_iter, int solver,
                           int n_steps, double dt, int obj, const double* prm, double* q, double* dq,
                           double* S, double* dS, double* J, double* dJ) {
  MAKE_G2;
  const size_t nvN = 4 * (size_t)G.N;
  auto qd = to_dual(q, dq, nvN);
  std::vector<Dual> Sd(G.N);
  Dual gg(g, dg), al(alpha, dalpha), Jd;
  int rc = run_steps2<Dual>(G, gg, al, n_iter, solver, n_steps, dt, qd.data(), Sd.data());
  if (!rc) rc = objective2<Dual>(G, obj, prm, gg, qd.data(), &Jd);
  from_dual(qd, q, dq); from_dual(Sd, S, dS);
  *J = Jd.v; *dJ = Jd.d;
  return finish(rc);
}
int orc2_jvp_objective(G2ARGS, int obj, const double* prm, double g, double dg, const double* q,
                       const double* dq, double* J, double* dJ) {
  MAKE_G2;
  auto qd = to_dual(q, dq, 4 * (size_t)G.N);
  Dual Jd;
  int rc = objective2<Dual>(G, obj, prm, Dual(g, dg), qd.data(), &Jd);
  *J = Jd.v; *dJ = Jd.d;
  return finish(rc);
}

// ------------------------------------------------------------------ 1D ----
int orc1_hyperbolic_rhs(G1ARGS, double g, const double* q, const double* S, double* dq) {
  MAKE_G1; hyperbolic_rhs1<double>(G, g, q, S, dq); return finish(0);
}
int orc1_apply_sip(G1ARGS, double alpha, const double* x, const double* rho, double* y) {
  MAKE_G1; apply_sip1<double>(G, alpha, x, rho, y); return finish(0);
}
int orc1_igr_rhs(G1ARGS, double alpha, const double* q, double* b) {
  MAKE_G1; igr_rhs1<double>(G, alpha, q, b); return finish(0);
}
int orc1_sip_diag(G1ARGS, double alpha, const double* rho, double* Md) {
  MAKE_G1; sip_diag1<double>(G, alpha, rho, Md); return finish(0);
}
int orc1_solver(G1ARGS, double alpha, int solver, int n_iter, const double* rho, const double* b,
                const double* Md, double* x, double* trace_a, double* trace_b) {
  MAKE_G1;
  std::vector<double> xv = to_vec<double>(x, G.N), bv = to_vec<double>(b, G.N), Mv = to_vec<double>(Md, G.N);
  auto A = [&](const double* in, double* out) { apply_sip1<double>(G, alpha, in, rho, out); };
  SolverTrace tr;
  int rc = run_solver(solver, A, xv, bv, Mv, n_iter, p, &tr);
  std::copy(xv.begin(), xv.end(), x);
  for (size_t k = 0; k < tr.a.size(); ++k) { if (trace_a) trace_a[k] = tr.a[k]; if (trace_b) trace_b[k] = tr.b[k]; }
  return finish(rc);
}
int orc1_solve_elliptic(G1ARGS, double alpha, int n_iter, int solver, const double* q, double* S) {
  MAKE_G1; return finish(solve_elliptic1<double>(G, alpha, n_iter, solver, q, S));
}
int orc1_euler_rhs(G1ARGS, double g, double alpha, int n_iter, int solver, const double* q, double* S, double* dq) {
  MAKE_G1; return finish(euler_rhs1<double>(G, g, alpha, n_iter, solver, q, S, dq));
}
int orc1_run_forward(G1ARGS, double g, double alpha, int n_iter, int solver, double dt, double T,
                     double* q, double* S, int* n_steps) {
  MAKE_G1; return finish(run_forward1<double>(G, g, alpha, n_iter, solver, dt, T, q, S, n_steps));
}
int orc1_run_steps(G1ARGS, double g, double alpha, int n_iter, int solver, int n_steps, double dt,
                   double* q, double* S) {
  MAKE_G1; return finish(run_steps1<double>(G, g, alpha, n_iter, solver, n_steps, dt, q, S));
}
int orc1_objective(G1ARGS, int obj, double g, const double* q, double* J) {
  MAKE_G1; return finish(objective1<double>(G, obj, g, q, J));
}
int orc1_run_sensitivity(G1ARGS, double g, double alpha, int n_iter, int solver, int form, double sgam,
                         double dt, double T, double* q, double* S, double* s, double* sS, int* n_steps) {
  MAKE_G1;
  return finish(run_sensitivity1<double>(G, g, alpha, n_iter, solver, form, sgam, dt, T, q, S, s, sS, n_steps));
}
int orc1_sens_elliptic_rhs(G1ARGS, double alpha, int form, const double* q, const double* S, const double* s,
                           double* b) {
  MAKE_G1;
  if (form == 1) sens_elliptic_rhs1_form1<double>(G, alpha, q, S, s, b);
  else if (form == 2) sens_elliptic_rhs1_form2<double>(G, alpha, q, S, s, b);
  else return -1;
  return finish(0);
}
int orc1_sens_hyperbolic_rhs(G1ARGS, double g, double sgam, const double* q, const double* S, const double* s,
                             const double* sS, double* ds) {
  MAKE_G1; sens_hyperbolic_rhs1<double>(G, g, q, S, s, sS, sgam, ds); return finish(0);
}
int orc1_jvp_hyperbolic_rhs(G1ARGS, double g, double dg, const double* q, const double* dq_in,
                            const double* S, const double* dS, double* out, double* dout) {
  MAKE_G1;
  auto qd = to_dual(q, dq_in, 3 * (size_t)G.N), Sd = to_dual(S, dS, G.N);
  std::vector<Dual> o(3 * (size_t)G.N);
  hyperbolic_rhs1<Dual>(G, Dual(g, dg), qd.data(), Sd.data(), o.data());
  from_dual(o, out, dout);
  return finish(0);
}
int orc1_jvp_objective_run(G1ARGS, double g, double dg, double alpha, double dalpha, int n_iter, int solver,
                           int n_steps, double dt, int obj, double* q, double* dq, double* S, double* dS,
                           double* J, double* dJ) {
  MAKE_G1;
  const size_t nvN = 3 * (size_t)G.N;
  auto qd = to_dual(q, dq, nvN);
  std::vector<Dual> Sd(G.N);
  Dual gg(g, dg), al(alpha, dalpha), Jd;
  int rc = run_steps1<Dual>(G, gg, al, n_iter, solver, n_steps, dt, qd.data(), Sd.data());
  if (!rc) rc = objective1<Dual>(G, obj, gg, qd.data(), &Jd);
  from_dual(qd, q, dq); from_dual(Sd, S, dS);
  *J = Jd.v; *dJ = Jd.d;
  return finish(rc);
}
int orc1_jvp_sensitivity_run(G1ARGS, double g, double alpha, int n_iter, int solver, int n_steps, double dt,
                             double* q, double* dq, double* S, double* dS) {
  // exact discrete JVP of run_steps1 (no final solve), for comparison with
  // the tangent (continuous-sensitivity) scheme
  MAKE_G1;
  const size_t nvN = 3 * (size_t)G.N;
  auto qd = to_dual(q, dq, nvN);
  std::vector<Dual> Sd(G.N);
  int rc = run_steps1<Dual>(G, Dual(g), Dual(alpha), n_iter, solver, n_steps, dt, qd.data(), Sd.data());
  from_dual(qd, q, dq); from_dual(Sd, S, dS);
  return finish(rc);
}

// ------------------------------------------- continuous adjoint (PDE) ----
// Snapshot arrays hold (n_steps + 1) records of (state, Sigma): 5N doubles in
// 2D, 4N in 1D; `cap` is the number of records the caller allocated.
void orc_set_adj_mutation(int m) { g_adj_mutation = m; }
void orc_set_fwd_mutation(int m) { g_fwd_mutation = m; }

int orc2_adj_hyp_rhs_local(G2ARGS, double g, const double* q, const double* S, const double* a, double* da) {
  MAKE_G2; adj_hyp_rhs_local2(G, g, q, S, a, da); return finish(0);
}
int orc2_adj_elliptic_rhs(G2ARGS, const double* a, const double* q, double* b) {
  MAKE_G2; adj_elliptic_rhs2(G, a, q, b); return finish(0);
}
int orc2_nonlocal_source(G2ARGS, const double* Pi, const double* q, const double* S, double* Sn) {
  MAKE_G2; nonlocal_source2(G, Pi, q, S, Sn); return finish(0);
}
int orc2_adj_rhs_conservative(G2ARGS, double g, double alpha, int n_iter, int solver, const double* q,
                              const double* S, const double* a, double* Pi, double* da) {
  MAKE_G2; return finish(adj_rhs_conservative2(G, g, alpha, n_iter, solver, q, S, a, Pi, da));
}
int orc2_run_forward_store(G2ARGS, double g, double alpha, int n_iter, int solver, double dt, double T,
                           const double* q0, int cap, double* snaps, int* n_steps) {
  MAKE_G2;
  std::vector<double> sv;
  int rc = run_forward_store2(G, g, alpha, n_iter, solver, dt, T, q0, sv, n_steps);
  if (rc) return finish(rc);
  if (sv.size() > (size_t)cap * 5 * (size_t)G.N) return -1;
  std::copy(sv.begin(), sv.end(), snaps);
  return finish(0);
}
int orc2_run_adjoint_conservative(G2ARGS, const double* snaps, int n_steps, double dt, double T, double g,
                                  double alpha, int n_iter, int solver, int obj, const double* prm, double* a) {
  MAKE_G2;
  return finish(run_adjoint_conservative2(G, snaps, n_steps, dt, T, g, alpha, n_iter, solver, obj, prm, a));
}
int orc2_adj_terminal(G2ARGS, int obj, const double* prm, double g, const double* qT, double* a) {
  MAKE_G2; return finish(adj_terminal2(G, obj, prm, g, qT, a));
}
int orc1_adj_hyp_rhs_local(G1ARGS, double g, const double* q, const double* S, const double* a, double* da) {
  MAKE_G1; adj_hyp_rhs_local1(G, g, q, S, a, da); return finish(0);
}
int orc1_adj_elliptic_rhs(G1ARGS, const double* a, const double* q, double* b) {
  MAKE_G1; adj_elliptic_rhs1(G, a, q, b); return finish(0);
}
int orc1_nonlocal_source(G1ARGS, const double* Pi, const double* q, const double* S, double* Sn) {
  MAKE_G1; nonlocal_source1(G, Pi, q, S, Sn); return finish(0);
}
int orc1_adj_rhs_conservative(G1ARGS, double g, double alpha, int n_iter, int solver, const double* q,
                              const double* S, const double* a, double* Pi, double* da) {
  MAKE_G1; return finish(adj_rhs_conservative1(G, g, alpha, n_iter, solver, q, S, a, Pi, da));
}
int orc1_run_forward_store(G1ARGS, double g, double alpha, int n_iter, int solver, double dt, double T,
                           const double* q0, int cap, double* snaps, int* n_steps) {
  MAKE_G1;
  std::vector<double> sv;
  int rc = run_forward_store1(G, g, alpha, n_iter, solver, dt, T, q0, sv, n_steps);
  if (rc) return finish(rc);
  if (sv.size() > (size_t)cap * 4 * (size_t)G.N) return -1;
  std::copy(sv.begin(), sv.end(), snaps);
  return finish(0);
}
int orc1_run_adjoint_conservative(G1ARGS, const double* snaps, int n_steps, double dt, double T, double g,
                                  double alpha, int n_iter, int solver, int obj, double* a) {
  MAKE_G1;
  return finish(run_adjoint_conservative1(G, snaps, n_steps, dt, T, g, alpha, n_iter, solver, obj, a));
}
int orc1_jvp_solve_elliptic(G1ARGS, double alpha, double dalpha, int n_iter, int solver, const double* q,
                            const double* dq, double* S, double* dS) {
  MAKE_G1;
  auto qd = to_dual(q, dq, 3 * (size_t)G.N);
  auto Sd = to_dual(S, dS, G.N);
  int rc = solve_elliptic1<Dual>(G, Dual(alpha, dalpha), n_iter, solver, qd.data(), Sd.data());
  from_dual(Sd, S, dS);
  return finish(rc);
}
int orc1_adj_terminal(G1ARGS, int obj, double g, const double* qT, double* a) {
  MAKE_G1; return finish(adj_terminal1(G, obj, g, qT, a));
}

}  // extern "C"
