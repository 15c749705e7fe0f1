// Host driver of the continuous ("optimise-then-discretise") adjoint and of
// the snapshot-storing forward run it consumes, and their C-ABI entry points:
//   igr_run_forward_store          run_forward_store (IGR_adjoints_2d/forward.jl:111-158,
//                                  IGR_adjoints_1d/forward.jl:137-187)
//   igr_get_snapshot / igr_set_snapshots   the snapshots vector it returns
//   igr_run_adjoint_conservative   run_adjoint_conservative (IGR_adjoints_2d/adjoint_pde.jl:522-611,
//                                  IGR_adjoints_1d/adjoint_pde.jl:466-552)
//
// The snapshots stay in HBM, (nv + 1) N doubles per record: 10 MB per step at
// 128^2 p = 3 and 40 MB at 256^2, so a whole run fits in the 288 GB of one
// MI355X.  2D: one SSP-RK3 step plus its Sigma solve is captured once as a
// hipGraph and replayed; after every replay the state and Sigma are copied
// into their record.  The adjoint step is captured the same way, reading its
// snapshot from a fixed buffer that is refilled from the record before each
// replay.  1D: one workgroup runs the whole forward store (k1_forward.hip,
// OP_STORE) and the whole backward run (k1_adjpde.hip).
#include <cmath>
#include <algorithm>

#include "igr_ctx.hpp"
#include "host_common.hpp"
#include "adjpde_common.hpp"
#include "k2_adjoint_common.hpp"

namespace igr {

void launch2_adj_stage(const Ctx& c, const AdjStageArgs& a);
void launch2_adj_b(const Ctx& c, const AdjBArgs& a);
void launch2_adj_terminal(const Ctx& c, const AdjTermArgs& a);
void launch2_obj(const Ctx& c, const ObjArgs& a, bool grad);
void launch_copy(const Ctx& c, double* dst, const double* src, size_t n);
void launch1_adjpde(const Ctx& c, const K1PdeArgs& a, int block);
Geo2 geo2(const Ctx& c);

static const double KE_WINDOW_PDE[4] = {0.55, 0.50, 0.15, 0.0};   // IGR_adjoints_2d/enzyme_adjoint.jl:26

static size_t rec_doubles(const Ctx& c) { return (size_t)(c.nv + 1) * c.N; }

// room for n_rec snapshot records (contents are not kept on growth)
static int ensure_snaps(Ctx& c, int n_rec) {
  if (c.snaps && c.snaps_cap >= (size_t)n_rec) return IGR_OK;
  HIPCHK(c, hipStreamSynchronize(c.stream));
  if (c.snaps) { (void)hipFree(c.snaps); c.snaps = nullptr; c.snaps_cap = 0; }
  c.snaps_n = -1;
  void* p = nullptr;
  const size_t bytes = (size_t)n_rec * rec_doubles(c) * sizeof(double);
  if (dev_malloc(&p, bytes) != hipSuccess) {
    return seterr(c, IGR_ENOMEM, "cannot allocate " + std::to_string(bytes >> 20) + " MiB of HBM for " +
                                     std::to_string(n_rec) + " snapshots");
  }
  c.snaps = static_cast<double*>(p);
  c.snaps_cap = (size_t)n_rec;
  return IGR_OK;
}

// continuous-adjoint workspace.  2D: A (4N), Pi (N), b (N), current snapshot
// (5N), sums (8).  1D: A, A1, A2 (3N each), Pi (N), b (N).
static int ensure_adjw(Ctx& c) {
  if (c.adjw) return IGR_OK;
  const size_t n = c.dim == 2 ? 11 * c.N + 8 : 11 * c.N;
  c.adjw = c.alloc(n);
  if (!c.adjw) return seterr(c, IGR_ENOMEM, "cannot allocate the continuous-adjoint workspace");
  return IGR_OK;
}

static void copy_record(Ctx& c, int rec) {
  double* d = c.snaps + (size_t)rec * rec_doubles(c);
  launch_copy(c, d, c.q, (size_t)c.nv * c.N);
  launch_copy(c, d + (size_t)c.nv * c.N, c.S, c.N);
}

// ---------------------------------------------------------------- 2D --
// One SSP-RK3 step followed by the Sigma solve of run_forward_store (forward.jl:132-155).
static void store_step2(Ctx& c, double dt, double gamma, double alpha, int n_iter, int solver) {
  step2(c, dt, gamma, alpha, n_iter, solver);
  if (alpha > 0.0) solve2(c, c.q, alpha, n_iter, solver, c.scal);
}

static int replay(Ctx& c, Ctx::GraphKey key, void (*body)(Ctx&, double, double, double, int, int)) {
  auto it = c.graphs.find(key);
  if (it == c.graphs.end()) {
    hipGraph_t g;
    hipGraphExec_t ge;
    HIPCHK(c, hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    body(c, key.dt, key.gamma, key.alpha, key.n_iter, key.solver);
    HIPCHK(c, hipStreamEndCapture(c.stream, &g));
    HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    it = c.graphs.emplace(key, std::make_pair(g, ge)).first;
  }
  HIPCHK(c, hipGraphLaunch(it->second.second, c.stream));
  return IGR_OK;
}

static int forward_store2(Ctx& c, int n_steps, double dt, double gamma, double alpha, int n_iter, int solver) {
  launch_fill(c, c.S, c.N, 0.0);
  if (alpha > 0.0) solve2(c, c.q, alpha, n_iter, solver, c.scal);
  copy_record(c, 0);
  Ctx::GraphKey key{dt, gamma, alpha, n_iter, solver};
  key.kind = 1;
  for (int s = 1; s <= n_steps; ++s) {
    if (c.use_graphs && n_steps >= 2) {
      const int rc = replay(c, key, store_step2);
      if (rc) return rc;
    } else {
      store_step2(c, dt, gamma, alpha, n_iter, solver);
    }
    copy_record(c, s);
  }
  HIPCHK(c, hipGetLastError());
  return IGR_OK;
}

// adjoint workspace views (2D)
struct AdjView2 {
  double *A, *Pi, *b, *cur, *sums;
  explicit AdjView2(Ctx& c) {
    A = c.adjw; Pi = A + 4 * c.N; b = Pi + c.N; cur = b + c.N; sums = cur + 5 * c.N;
  }
};

// One backward SSP-RK3 step of run_adjoint_conservative (adjoint_pde.jl:593-607)
// on the snapshot in the `cur` buffer: a1 = c.q1, a2 = c.q2.
static void adj_step2(Ctx& c, double h, double gamma, double alpha, int n_iter, int solver) {
  AdjView2 v(c);
  const double* qn = v.cur;
  const double* src[3] = {v.A, c.q1, c.q2};
  double* out[3] = {c.q1, c.q2, v.A};
  const double cd[3] = {h, 1.0 / 4.0 * h, 2.0 / 3.0 * h};
  for (int s = 0; s < 3; ++s) {
    if (alpha > 0.0) {
      AdjBArgs b;
      b.g = geo2(c); b.B = c.basis; b.alpha = alpha; b.q = qn; b.A = src[s]; b.b = v.b;
      launch2_adj_b(c, b);
      solve2(c, qn, alpha, n_iter, solver, c.scal, v.Pi, v.b);
    }
    AdjStageArgs a;
    a.g = geo2(c); a.B = c.basis; a.gamma = gamma; a.q = qn; a.S = qn + 4 * c.N; a.A = src[s]; a.an = v.A;
    a.Pi = alpha > 0.0 ? v.Pi : nullptr; a.out = out[s]; a.mode = s + 1; a.cdt = cd[s]; a.flag = c.flag;
    launch2_adj_stage(c, a);
  }
}

static int adjoint2(Ctx& c, int n_steps, int obj, const double* prm, double h, double gamma, double alpha, int n_iter,
                    int solver) {
  AdjView2 v(c);
  const double* qT = c.snaps + (size_t)c.snaps_n * rec_doubles(c);
  const double* pp = prm ? prm : KE_WINDOW_PDE;
  if (obj == 2) {   // kinetic energy / mass: the two sums first
    ObjArgs o;
    o.g = geo2(c); o.B = c.basis; o.obj = 2;
    for (int k = 0; k < 4; ++k) o.prm[k] = pp[k];
    o.Lx = c.Lx; o.Ly = c.Ly; o.gamma = gamma; o.q = qT; o.out = v.sums; o.qbar = nullptr; o.gbar = nullptr;
    o.partials = c.partials; o.counter = c.counter;
    launch2_obj(c, o, false);
  }
  AdjTermArgs t;
  t.g = geo2(c); t.B = c.basis; t.np = c.np; t.obj = obj;
  for (int k = 0; k < 4; ++k) t.prm[k] = pp[k];
  t.Lx = c.Lx; t.Ly = c.Ly; t.gamma = gamma; t.q = qT; t.sums = v.sums; t.a = v.A;
  launch2_adj_terminal(c, t);
  launch_fill(c, v.Pi, c.N, 0.0);
  Ctx::GraphKey key{h, gamma, alpha, n_iter, solver};
  key.kind = 2;
  for (int n = n_steps; n >= 1; --n) {
    launch_copy(c, v.cur, c.snaps + (size_t)n * rec_doubles(c), 5 * c.N);
    if (c.use_graphs && n_steps >= 2) {
      const int rc = replay(c, key, adj_step2);
      if (rc) return rc;
    } else {
      adj_step2(c, h, gamma, alpha, n_iter, solver);
    }
  }
  HIPCHK(c, hipGetLastError());
  return IGR_OK;
}

// ---------------------------------------------------------------- 1D --
static int forward_store1(Ctx& c, int n_steps, double dt, double gamma, double alpha, int n_iter, int solver) {
  K1Args a = k1_base(c, gamma, alpha, n_iter, solver);
  a.op = OP_STORE; a.n_steps = n_steps; a.dt = dt; a.snaps = c.snaps;
  return run1(c, a);
}

static int adjoint1(Ctx& c, int n_steps, int obj, double h, double gamma, double alpha, int n_iter, int solver) {
  K1PdeArgs p;
  p.f = k1_base(c, gamma, alpha, n_iter, solver);
  p.snaps = c.snaps; p.n_steps = n_steps; p.n_last = c.snaps_n; p.h = h; p.obj = obj;
  p.A = c.adjw; p.A1 = p.A + 3 * c.N; p.A2 = p.A1 + 3 * c.N; p.Pi = p.A2 + 3 * c.N; p.bPi = p.Pi + c.N;
  launch1_adjpde(c, p, k1_block(c));
  HIPCHK(c, hipGetLastError());
  return IGR_OK;
}

static bool pde_objective_ok(const Ctx& c, int obj) {
  if (c.dim == 2) return obj >= 0 && obj <= 6;
  return obj == 0 || obj == 1 || obj == 2 || obj == 4;
}

}  // namespace igr

using namespace igr;

extern "C" {

int igr_run_forward_store(igr_ctx* c, double dt, double t_end, double gamma, double alpha, int n_iter, int solver,
                          int* n_steps_out) {
  if (!c) return IGR_EINVAL;
  if (!(dt > 0.0) || !(t_end > 0.0) || n_iter < 0) return seterr(*c, IGR_EINVAL, "bad argument");
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "unknown solver");
  const int n_steps = (int)std::ceil(t_end / dt);
  dt = t_end / n_steps;
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = ensure_scal(*c, n_iter);
  if (rc) return rc;
  rc = ensure_snaps(*c, n_steps + 1);
  if (rc) return rc;
  rc = c->dim == 2 ? forward_store2(*c, n_steps, dt, gamma, alpha, n_iter, solver)
                   : forward_store1(*c, n_steps, dt, gamma, alpha, n_iter, solver);
  if (rc) return rc;
  rc = end_call(*c);
  if (rc) return rc;
  c->snaps_n = n_steps;
  if (n_steps_out) *n_steps_out = n_steps;
  return IGR_OK;
}

int igr_snapshot_count(const igr_ctx* c, int* n_steps) {
  if (!c || !n_steps) return IGR_EINVAL;
  *n_steps = c->snaps_n;
  return IGR_OK;
}

int igr_get_snapshot(igr_ctx* c, int k, double* q, double* sigma) {
  if (!c) return IGR_EINVAL;
  if (k < 0 || k > c->snaps_n) return seterr(*c, IGR_EINVAL, "no such snapshot");
  const double* r = c->snaps + (size_t)k * rec_doubles(*c);
  const size_t nvN = (size_t)c->nv * c->N;
  if (q) HIPCHK(*c, hipMemcpyAsync(q, r, nvN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (sigma) HIPCHK(*c, hipMemcpyAsync(sigma, r + nvN, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(*c, hipStreamSynchronize(c->stream));
  return IGR_OK;
}

int igr_set_snapshots(igr_ctx* c, int n_steps, const double* records) {
  if (!c) return IGR_EINVAL;
  if (n_steps < 0 || !records) return seterr(*c, IGR_EINVAL, "bad argument");
  int rc = ensure_snaps(*c, n_steps + 1);
  if (rc) return rc;
  HIPCHK(*c, hipMemcpyAsync(c->snaps, records, (size_t)(n_steps + 1) * rec_doubles(*c) * sizeof(double),
                            hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipStreamSynchronize(c->stream));
  c->snaps_n = n_steps;
  return IGR_OK;
}

int igr_run_adjoint_conservative(igr_ctx* c, int objective, const double* prm, int n_steps, double dt, double t_end,
                                 double gamma, double alpha, int n_iter, int solver, double* a_out) {
  if (!c) return IGR_EINVAL;
  if (c->snaps_n < 0) return seterr(*c, IGR_EINVAL, "no snapshots: call igr_run_forward_store first");
  if (n_steps < 0 || n_steps > c->snaps_n || !(dt > 0.0) || !(t_end > 0.0) || n_iter < 0)
    return seterr(*c, IGR_EINVAL, "bad argument");
  if (!pde_objective_ok(*c, objective)) return seterr(*c, IGR_EINVAL, "unknown objective");
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "unknown solver");
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = ensure_scal(*c, n_iter);
  if (rc) return rc;
  rc = ensure_adjw(*c);
  if (rc) return rc;
  const double h = std::min(dt, t_end);   // adjoint_pde.jl:585 (1D :525)
  rc = c->dim == 2 ? adjoint2(*c, n_steps, objective, prm, h, gamma, alpha, n_iter, solver)
                   : adjoint1(*c, n_steps, objective, h, gamma, alpha, n_iter, solver);
  if (rc) return rc;
  if (a_out)
    HIPCHK(*c, hipMemcpyAsync(a_out, c->adjw, (size_t)c->nv * c->N * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
  return end_call(*c);
}

}  // extern "C"
