// Host driver of the forward tangent (sensitivity) equations: the primal
// state and n_tan tangent states advanced together by SSP-RK3
// (run_sensitivity, IGR_adjoints_2d/sensitivity.jl:353-413;
// IGR_adjoints_1d/sensitivity.jl:384-515).  Every tangent shares the
// stage's primal Sigma, 1/rho and diag(A); each solves its own varsigma
// with the same fixed-iteration solver.
#include <cmath>
#include <vector>
#include <map>
#include <mutex>

#include "igr_ctx.hpp"
#include "host_common.hpp"

namespace igr {

void launch2_hyp(const Ctx& c, const HypArgs& a);
void launch2_sens_rhs(const Ctx& c, const SensArgs& a);
void launch2_sens_hyp(const Ctx& c, const SensArgs& a);
void launch_copy(const Ctx& c, double* dst, const double* src, size_t n);
Geo2 geo2(const Ctx& c);
HypArgs hyp_base(const Ctx& c, double gamma);
void launch1_run(const Ctx& c, const K1Args& a, int block);
int k1_block(const Ctx& c);

// Per-context tangent buffers; the shared map is guarded as the gradient's is.
static std::map<const Ctx*, TanBufs*> g_tan;
static std::mutex g_tan_mu;

int tan_bufs(Ctx& c, int nt, TanBufs*& T) {
  TanBufs*& t = [&]() -> TanBufs*& { std::lock_guard<std::mutex> lk(g_tan_mu); return g_tan[&c]; }();
  if (!t) t = new TanBufs();
  const size_t nvN = c.nv * c.N;
  while (t->nt < nt) {
    double *a = c.alloc(nvN), *b = c.alloc(nvN), *d = c.alloc(nvN), *e = c.alloc(c.N);
    if (!a || !b || !d || !e) return IGR_ENOMEM;
    t->s.push_back(a); t->s1.push_back(b); t->s2.push_back(d); t->vs.push_back(e);
    ++t->nt;
  }
  T = t;
  return IGR_OK;
}

void free_tan_bufs(const Ctx* c) {
  std::lock_guard<std::mutex> lk(g_tan_mu);
  auto it = g_tan.find(c);
  if (it != g_tan.end()) { delete it->second; g_tan.erase(it); }
}

static SensArgs sens_base(const Ctx& c, double gamma, double alpha) {
  SensArgs a;
  a.g = geo2(c); a.B = c.basis; a.gamma = gamma; a.alpha = alpha; a.flag = c.flag;
  a.q = nullptr; a.S = c.S; a.s = nullptr; a.vs = nullptr; a.sn = nullptr; a.out = nullptr;
  a.mode = 0; a.cdt = 0.0; a.sgam = 0.0;
  return a;
}

// tangent varsigma solve on stage qs / tangent ss (solve_sensitivity_elliptic!, sensitivity.jl:301-316)
static void sens_solve2(Ctx& c, const double* qs, const double* ss, double* vs, double gamma, double alpha,
                        int n_iter, int solver) {
  SensArgs r = sens_base(c, gamma, alpha);
  r.q = qs; r.s = ss; r.out = c.tmp;
  launch2_sens_rhs(c, r);
  solve2(c, qs, alpha, n_iter, solver, c.scal, vs, c.tmp);
}

// run_sensitivity (sensitivity.jl:366-413) in three parts, so the batched
// driver can record one step of many instances: Sigma and varsigma zeroed,
// one SSP-RK3 step of the coupled system, the final solves.
void sens2_begin(Ctx& c, TanBufs& T, int nt) {
  launch_fill(c, c.S, c.N, 0.0);
  for (int t = 0; t < nt; ++t) launch_fill(c, T.vs[t], c.N, 0.0);
}

void sens2_step(Ctx& c, TanBufs& T, int nt, double dt, double gamma, double alpha, int n_iter, int solver,
                double sgam) {
  const size_t N = c.N;
  const double cdt[3] = {dt, 1.0 / 4.0 * dt, 2.0 / 3.0 * dt};
  double* qst[3] = {c.q, c.q1, c.q2};
  double* qout[3] = {c.q1, c.q2, c.q};
  for (int st = 0; st < 3; ++st) {
    const double* qs = qst[st];
    if (alpha > 0.0) solve2(c, qs, alpha, n_iter, solver, c.scal);
    else launch_fill(c, c.S, N, 0.0);
    for (int t = 0; t < nt; ++t) {
      const double* ss = st == 0 ? T.s[t] : (st == 1 ? T.s1[t] : T.s2[t]);
      double* so = st == 0 ? T.s1[t] : (st == 1 ? T.s2[t] : T.s[t]);
      if (alpha > 0.0) sens_solve2(c, qs, ss, T.vs[t], gamma, alpha, n_iter, solver);
      else launch_fill(c, T.vs[t], N, 0.0);
      SensArgs h = sens_base(c, gamma, alpha);
      h.q = qs; h.s = ss; h.vs = T.vs[t]; h.sn = T.s[t]; h.out = so; h.mode = st + 1; h.cdt = cdt[st];
      h.sgam = sgam;
      launch2_sens_hyp(c, h);
    }
    HypArgs h = hyp_base(c, gamma);
    h.q = qs; h.mode = st + 1; h.cdt = cdt[st]; h.out = qout[st];
    launch2_hyp(c, h);
  }
}

void sens2_end(Ctx& c, TanBufs& T, int nt, double gamma, double alpha, int n_iter, int solver) {
  if (alpha > 0.0) {
    solve2(c, c.q, alpha, n_iter, solver, c.scal);
    for (int t = 0; t < nt; ++t) sens_solve2(c, c.q, T.s[t], T.vs[t], gamma, alpha, n_iter, solver);
  }
}

// 2D stepping of run_sensitivity: n_steps = ceil(T/dt), dt = T/n_steps (sensitivity.jl:370-371)
int sens2_steps(double dt, double t_end, double& dts) {
  const int n = (int)std::ceil(t_end / dt);
  dts = t_end / n;
  return n;
}

// 1D run_sensitivity as one whole-run argument block: the reference's
// `while t < T - 1e-14; dt = min(Dt, T - t)` stepping (sensitivity.jl:419-420)
// is a run of full steps and at most one shorter last step.
int sens1_args(Ctx& c, TanBufs& T, int n_tan, double dt, double t_end, double gamma, double alpha, int n_iter,
               int solver, int form, double sgam, K1Args& a, int& n_steps) {
  std::vector<std::pair<int, double>> runs;
  double t = 0.0;
  n_steps = 0;
  while (t < t_end - 1e-14) {
    const double d = std::min(dt, t_end - t);
    if (!runs.empty() && runs.back().second == d) ++runs.back().first;
    else runs.push_back({1, d});
    t += d;
    ++n_steps;
  }
  if (runs.size() > 2) return seterr(c, IGR_EINVAL, "1D stepping with more than two step sizes");
  a = k1_base(c, gamma, alpha, n_iter, solver);
  a.op = OP_RUN;
  a.n_steps = runs.empty() ? 0 : runs[0].first;
  a.dt = runs.empty() ? 0.0 : runs[0].second;
  a.n_steps2 = runs.size() > 1 ? runs[1].first : 0;
  a.dt2 = runs.size() > 1 ? runs[1].second : 0.0;
  a.zero_sigma = 1; a.final_solve = 1;
  a.n_tan = n_tan; a.form = form; a.sgam = sgam;
  for (int tt = 0; tt < n_tan; ++tt) {
    a.ts[tt] = T.s[tt]; a.ts1[tt] = T.s1[tt]; a.ts2[tt] = T.s2[tt]; a.tvs[tt] = T.vs[tt];
  }
  return IGR_OK;
}

}  // namespace igr

using namespace igr;

extern "C" {

int igr_run_sensitivity(igr_ctx* c, int n_tan, const double* s0, double dt, double t_end, double gamma,
                        double alpha, int n_iter, int solver, int form, double sgam, double* s_out,
                        double* vs_out, int* n_steps_out) {
  if (!c || n_tan < 1 || n_tan > 8 || !s0 || !(dt > 0.0) || !(t_end > 0.0) || n_iter < 0) return IGR_EINVAL;
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "Unknown solver");
  if (c->dim == 1 && form != 1 && form != 2) return seterr(*c, IGR_EINVAL, "Unknown elliptic_rhs (use form 1 or 2)");
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = ensure_scal(*c, n_iter);
  if (rc) return rc;
  TanBufs* T = nullptr;
  rc = tan_bufs(*c, n_tan, T);
  if (rc) return seterr(*c, rc, "device memory for tangent states");
  const size_t N = c->N, nvN = c->nv * N;
  for (int t = 0; t < n_tan; ++t)
    HIPCHK(*c, hipMemcpyAsync(T->s[t], s0 + t * nvN, nvN * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int n_steps = 0;
  if (c->dim == 2) {
    double dts = 0.0;
    n_steps = sens2_steps(dt, t_end, dts);
    sens2_begin(*c, *T, n_tan);
    for (int k = 0; k < n_steps; ++k) sens2_step(*c, *T, n_tan, dts, gamma, alpha, n_iter, solver, sgam);
    sens2_end(*c, *T, n_tan, gamma, alpha, n_iter, solver);
  } else {
    if (n_tan > K1_MAXTAN) return seterr(*c, IGR_EINVAL, "1D tangent count above 4");
    K1Args a;
    rc = sens1_args(*c, *T, n_tan, dt, t_end, gamma, alpha, n_iter, solver, form, sgam, a, n_steps);
    if (rc) return rc;
    launch1_run(*c, a, k1_block(*c));
  }
  for (int t = 0; t < n_tan; ++t) {
    if (s_out) HIPCHK(*c, hipMemcpyAsync(s_out + t * nvN, T->s[t], nvN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (vs_out) HIPCHK(*c, hipMemcpyAsync(vs_out + t * N, T->vs[t], N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (n_steps_out) *n_steps_out = n_steps;
  return end_call(*c);
}

}  // extern "C"
