// Host driver of the 2D discrete adjoint: forward recompute of one step with
// every solver iterate stored (PCG, damped Jacobi or Chebyshev), the reverse SSP-RK3 sweep, binomial
// checkpointing of time steps, and the C-ABI gradient / VJP entry points.
//
// The reference differentiates its whole forward map with Enzyme reverse
// mode over a tape kept in host RAM (IGR_adjoints_2d/enzyme_adjoint.jl:
// 270-296).  Here the map is the same (SSP-RK3 with Sigma warm-started from
// the previous stage, zeroed once at t = 0; enzyme_adjoint.jl:50-70) and its
// exact discrete adjoint is assembled from hand-written reverse kernels.
// Memory: one checkpoint (q, Sigma) per slot, plus the iterates of the one
// step being reversed; every other step is recomputed from the nearest
// checkpoint (optimal binomial split).
#include <cmath>
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "igr_ctx.hpp"
#include "host_common.hpp"
#include "k2_adjoint_common.hpp"

namespace igr {

void launch2_sip(const Ctx& c, int mode, const SipArgs& a);
void launch2_hyp(const Ctx& c, const HypArgs& a);
void launch2_hyp_rev(const Ctx& c, const HypRevArgs& a);
void launch2_rev(const Ctx& c, int mode, const RevArgs& a);
void launch2_obj(const Ctx& c, const ObjArgs& a, bool grad);
void launch_sum_blocks(const Ctx& c, const double* blk, int nb, double* out);
void launch_copy(const Ctx& c, double* dst, const double* src, size_t n);
void launch_pcg_final(const Ctx& c, const double* part, double* sc, int nit);
int grid2_blocks(const Ctx& c);
int grid2_rblocks(const Ctx& c);
int grid2_hblocks(const Ctx& c);
Geo2 geo2(const Ctx& c);
SipArgs sip_base(const Ctx& c, double alpha);
HypArgs hyp_base(const Ctx& c, double gamma);
double cheb_lmax(int p);
void launch1_adjoint(const Ctx& c, const K1AdjArgs& a, int block);
void launch1_run(const Ctx& c, const K1Args& a, int block);


// Storage of one stage's solve, written by the recompute and read by the reverse.
struct StageStore {
  // PCG: P = p_k, AP = A p_k (k < n), R = r_k (k <= n).
  // Jacobi: P = x_k, AP = A x_k (k < n).  Chebyshev: P = x_k at k-1 (1 <= k <= n), R = r_k (k < n).
  double *P = nullptr, *AP = nullptr, *R = nullptr;   // n, n, n+1 fields
  double* b = nullptr;                                 // IGR source (Jacobi/Chebyshev)
  double *invR = nullptr, *invM = nullptr;
  double *sc = nullptr, *rs = nullptr;
};

// The stored stage states, Sigmas and solver data of one whole step.
struct StepView {
  StageStore* st;
  double** Q;
  double** Sg;
};
struct StepSet {
  StageStore st[3];
  double* Q[3] = {nullptr, nullptr, nullptr};
  double* Sg[4] = {nullptr, nullptr, nullptr, nullptr};
  StepView view() { return {st, Q, Sg}; }
};

// The block that holds kept whole-step iterates can be most of the HBM, and
// fresh device memory that another allocation just released costs about
// 30 ms per GB to obtain again (measured on MI355X: a 170 GB block, 5 s).  One
// such block per device therefore outlives the gradient workspace that used
// it and is handed to the next one, whichever context it belongs to, until
// igr_release_cache() frees it.
struct SetPool {
  void* p = nullptr;
  size_t bytes = 0;
  bool in_use = false;
};
static std::map<int, SetPool> g_setpool;
static std::mutex g_pool_mu;   // guards g_setpool (contexts may run on several host threads)

static int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}
// bytes of the idle pooled block of this device (counts as free for sizing)
static size_t pool_idle_bytes() {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  auto it = g_setpool.find(current_device());
  return (it != g_setpool.end() && !it->second.in_use) ? it->second.bytes : 0;
}
// a block of at least `bytes`: the pooled one if idle and large enough, else
// a new one (which replaces an idle pooled block, or is private when the pooled
// block is in use by another workspace)
static void* pool_take(size_t bytes, bool& pooled) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  SetPool& sp = g_setpool[current_device()];
  pooled = false;
  void* q = nullptr;
  if (!sp.in_use) {
    if (sp.p && sp.bytes >= bytes) { sp.in_use = true; pooled = true; return sp.p; }
    if (sp.p) { (void)hipDeviceSynchronize(); (void)hipFree(sp.p); sp.p = nullptr; sp.bytes = 0; }
    if (hipMalloc(&q, bytes) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    sp.p = q; sp.bytes = bytes; sp.in_use = true; pooled = true;
    return q;
  }
  if (hipMalloc(&q, bytes) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
  return q;
}
static void pool_give(void* q, bool pooled) {
  if (!q) return;
  if (!pooled) { (void)hipFree(q); return; }
  std::lock_guard<std::mutex> lk(g_pool_mu);
  for (auto& kv : g_setpool)
    if (kv.second.p == q) kv.second.in_use = false;
}
// frees the idle pooled block of device `dev` (caller holds g_pool_mu)
static bool pool_drop_idle(int dev) {
  auto it = g_setpool.find(dev);
  if (it == g_setpool.end() || it->second.in_use || !it->second.p) return false;
  int d = 0;
  (void)hipGetDevice(&d);
  (void)hipSetDevice(dev);
  (void)hipDeviceSynchronize();
  (void)hipFree(it->second.p);
  (void)hipSetDevice(d);
  it->second.p = nullptr;
  it->second.bytes = 0;
  return true;
}

// Workspace of the gradient (allocated per (n_iter) on demand, kept in the context).
struct GradWork {
  int n_iter = -1;
  StageStore st[3];
  double* Q[3] = {nullptr, nullptr, nullptr};   // q^n, q1, q2
  double* Sg[4] = {nullptr, nullptr, nullptr, nullptr};   // Sigma_0..3
  double *QBn = nullptr, *QBa = nullptr, *QBb = nullptr, *QBnext = nullptr;
  double *Xbar = nullptr, *Scarry = nullptr;
  double *RB[2] = {nullptr, nullptr}, *PB[2] = {nullptr, nullptr};
  double *irbar = nullptr, *imbar = nullptr, *bbar = nullptr;
  double *XB[2] = {nullptr, nullptr};        // Chebyshev x-bar ping-pong
  double *gblk = nullptr, *ablk = nullptr;   // per-workgroup accumulators
  double* dsc = nullptr;                     // [0] J, [1] KE, [2] M, [3] gamma-bar, [4] alpha-bar
  std::vector<double*> ck;                   // checkpoint slots (5N each)
  // whole steps kept from the forward sweep (no recompute when reversed):
  // one device block holding `sets`
  std::vector<StepSet> sets;
  void* set_block = nullptr;
  size_t set_block_bytes = 0;
  bool set_block_pooled = false;
  // every other buffer of this workspace, freed with it (not with the context)
  std::vector<void*> bufs;
  double* alloc(size_t n) {
    void* p = nullptr;
    if (dev_malloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) return nullptr;
    bufs.push_back(p);
    return static_cast<double*>(p);
  }
  StepView view() { return {st, Q, Sg}; }
  // the owner synchronizes the context stream before deleting a workspace
  ~GradWork() {
    pool_give(set_block, set_block_pooled);
    for (void* p : bufs) (void)hipFree(p);
  }
};

// Doubles of one StepSet for n solver iterations on N nodes, as alloc_sets
// carves it: every piece rounded up to 32 doubles (256 B).
static size_t carve_round(size_t cnt) { return (cnt + 31) / 32 * 32; }
static size_t step_set_doubles(size_t N, int n) {
  const size_t n1 = (size_t)std::max(n, 1);
  const size_t per_stage = 2 * carve_round(n1 * N) + carve_round((size_t)(n + 1) * N) + 3 * carve_round(N) +
                           carve_round(4 * n1 + 1) + carve_round(5 * n1 + 1);
  return 3 * per_stage + 3 * carve_round(4 * N) + 4 * carve_round(N);
}

// Carve K step sets out of one allocation (kept until the context goes).
static int alloc_sets(Ctx& c, GradWork& W, int K) {
  if ((int)W.sets.size() >= K) return IGR_OK;
  const size_t N = c.N;
  const int n = W.n_iter;
  const size_t per = step_set_doubles(N, n);
  if (W.set_block) { pool_give(W.set_block, W.set_block_pooled); W.set_block = nullptr; W.sets.clear(); }
  const size_t bytes = per * sizeof(double) * (size_t)K;
  W.set_block = pool_take(bytes, W.set_block_pooled);
  if (!W.set_block) return IGR_ENOMEM;
  W.set_block_bytes = bytes;
  double* p = (double*)W.set_block;
  double* const end = p + per * (size_t)K;
  auto take = [&](size_t cnt) { double* r = p; p += carve_round(cnt); return r; };
  W.sets.resize(K);
  for (int k = 0; k < K; ++k) {
    StepSet& S = W.sets[k];
    for (int st = 0; st < 3; ++st) {
      StageStore& t = S.st[st];
      t.P = take((size_t)std::max(n, 1) * N); t.AP = take((size_t)std::max(n, 1) * N); t.R = take((size_t)(n + 1) * N);
      t.invR = take(N); t.invM = take(N); t.b = take(N);
      t.sc = take(4 * (size_t)std::max(n, 1) + 1); t.rs = take(5 * (size_t)std::max(n, 1) + 1);
    }
    for (int q = 0; q < 3; ++q) S.Q[q] = take(4 * N);
    for (int q = 0; q < 4; ++q) S.Sg[q] = take(N);
  }
  if (p != end) {   // the carve must fill exactly what step_set_doubles sized
    pool_give(W.set_block, W.set_block_pooled);
    W.set_block = nullptr;
    W.sets.clear();
    return IGR_EINVAL;
  }
  return IGR_OK;
}

// Per-context gradient workspaces.  The maps are shared by all contexts, so
// their lookups and erasures hold g_work_mu (contexts may be used from
// several host threads); the slot a lookup returns belongs to one context.
static std::map<const Ctx*, GradWork*> g_work;
static std::mutex g_work_mu;

// A workspace sized for n solver iterations serves every call with n or fewer
// (the per-iteration stores and scalar records are indexed with the call's
// own n_iter), so the objective and VJP entry points (n_iter = 0) reuse the
// gradient's.  A call with more iterations replaces it, and the old one's
// buffers are freed once the context stream has drained.
static int alloc_work(Ctx& c, int n_iter, int slots, GradWork*& W) {
  GradWork*& w = [&]() -> GradWork*& { std::lock_guard<std::mutex> lk(g_work_mu); return g_work[&c]; }();
  const size_t N = c.N;
  const int n = std::max(n_iter, 0);
  if (!w || w->n_iter < n) {
    if (w) {
      (void)hipStreamSynchronize(c.stream);
      delete w;
      w = nullptr;
    }
    GradWork* nw = new GradWork();
    nw->n_iter = n;
    auto fail = [&]() { delete nw; return IGR_ENOMEM; };
    for (int s = 0; s < 3; ++s) {
      StageStore& st = nw->st[s];
      st.P = nw->alloc(std::max<size_t>((size_t)n * N, 1));
      st.AP = nw->alloc(std::max<size_t>((size_t)n * N, 1));
      st.R = nw->alloc((size_t)(n + 1) * N);
      st.invR = nw->alloc(N); st.invM = nw->alloc(N); st.b = nw->alloc(N);
      st.sc = nw->alloc(4 * (size_t)std::max(n, 1) + 1);
      st.rs = nw->alloc(5 * (size_t)std::max(n, 1) + 1);
      if (!st.P || !st.AP || !st.R || !st.invR || !st.invM || !st.b || !st.sc || !st.rs) return fail();
    }
    for (int s = 0; s < 3; ++s) if (!(nw->Q[s] = nw->alloc(4 * N))) return fail();
    for (int s = 0; s < 4; ++s) if (!(nw->Sg[s] = nw->alloc(N))) return fail();
    double** one[] = {&nw->Xbar, &nw->Scarry, &nw->RB[0], &nw->RB[1], &nw->PB[0], &nw->PB[1], &nw->irbar,
                      &nw->imbar, &nw->bbar, &nw->XB[0], &nw->XB[1]};
    for (auto p : one) if (!(*p = nw->alloc(N))) return fail();
    double** four[] = {&nw->QBn, &nw->QBa, &nw->QBb, &nw->QBnext};
    for (auto p : four) if (!(*p = nw->alloc(4 * N))) return fail();
    const size_t nb = c.partials_n;
    if (!(nw->gblk = nw->alloc(nb)) || !(nw->ablk = nw->alloc(nb)) || !(nw->dsc = nw->alloc(16))) return fail();
    w = nw;
  }
  while ((int)w->ck.size() < slots) {
    double* p = w->alloc(5 * N);
    if (!p) return IGR_ENOMEM;
    w->ck.push_back(p);
  }
  W = w;
  return IGR_OK;
}

// 1D workspace: the stored step, the adjoint vectors and checkpoint slots (4N each).
struct GradWork1 {
  int n_iter = -1;
  K1AdjArgs base;
  std::vector<double*> ck;
  std::vector<void*> bufs;   // every buffer above, freed with the workspace
  double* alloc(size_t n) {
    void* p = nullptr;
    if (dev_malloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) return nullptr;
    bufs.push_back(p);
    return static_cast<double*>(p);
  }
  ~GradWork1() {
    for (void* p : bufs) (void)hipFree(p);
  }
};
static std::map<const Ctx*, GradWork1*> g_work1;

static int alloc_work1(Ctx& c, int n_iter, int slots, GradWork1*& W) {
  GradWork1*& w = [&]() -> GradWork1*& { std::lock_guard<std::mutex> lk(g_work_mu); return g_work1[&c]; }();
  const size_t N = c.N;
  const int n = std::max(n_iter, 0);
  if (!w || w->n_iter < n) {   // as alloc_work: a larger workspace serves smaller n_iter
    if (w) {
      (void)hipStreamSynchronize(c.stream);
      delete w;
      w = nullptr;
    }
    GradWork1* nw = new GradWork1();
    nw->n_iter = n;
    auto fail = [&]() { delete nw; return IGR_ENOMEM; };
    K1AdjArgs& a = nw->base;
    for (int s = 0; s < 3; ++s) {
      a.Q[s] = nw->alloc(3 * N); a.invR[s] = nw->alloc(N); a.invM[s] = nw->alloc(N); a.b[s] = nw->alloc(N);
      a.P[s] = nw->alloc(std::max<size_t>((size_t)n * N, 1));
      a.AP[s] = nw->alloc(std::max<size_t>((size_t)n * N, 1));
      a.R[s] = nw->alloc((size_t)(n + 1) * N); a.sc[s] = nw->alloc(4 * (size_t)std::max(n, 1) + 1);
      if (!a.Q[s] || !a.invR[s] || !a.invM[s] || !a.b[s] || !a.P[s] || !a.AP[s] || !a.R[s] || !a.sc[s])
        return fail();
    }
    for (int s = 0; s < 4; ++s) if (!(a.Sg[s] = nw->alloc(N))) return fail();
    double** three[] = {&a.QBnext, &a.QBn, &a.QBa, &a.QBb, &a.Ab};
    for (auto p : three) if (!(*p = nw->alloc(3 * N))) return fail();
    double** one[] = {&a.Xbar, &a.Scarry, &a.w, &a.y, &a.rb, &a.pb, &a.xb, &a.irbar, &a.imbar, &a.bbar};
    for (auto p : one) if (!(*p = nw->alloc(N))) return fail();
    if (!(a.dsc = nw->alloc(16))) return fail();
    w = nw;
  }
  while ((int)w->ck.size() < slots) {
    double* p = w->alloc(4 * N);
    if (!p) return IGR_ENOMEM;
    w->ck.push_back(p);
  }
  W = w;
  return IGR_OK;
}

// Chebyshev recurrence coefficients of iterations k = 1..n, as the forward
// solve forms them (chebyshev_fixed!, IGR_adjoints_2d/elliptic.jl:370-397).
static void cheb_coeffs(int p, int n, double& theta, std::vector<double>& c1, std::vector<double>& c2) {
  const double lmin = 1.0, lmax = cheb_lmax(p);
  theta = (lmax + lmin) / 2.0;
  const double delta = (lmax - lmin) / 2.0;
  const double sigma = theta / std::max(delta, 1e-10);
  double rc = 1.0 / sigma;
  c1.assign(n + 1, 0.0);
  c2.assign(n + 1, 0.0);
  for (int k = 2; k <= n; ++k) {
    const double rn = 1.0 / (2.0 * sigma - rc);
    c1[k] = rn * rc; c2[k] = 2.0 * rn / delta; rc = rn;
  }
}

// PCG solve on stage state qs with warm start/result x, keeping every iterate.
static void solve2_store_pcg(Ctx& c, const double* qs, double* x, double alpha, int n_iter, StageStore& st) {
  const size_t N = c.N;
  SipArgs a = sip_base(c, alpha);
  a.mx = qs; a.my = qs + N; a.rho = qs + 2 * N; a.x = x;
  a.invR_out = st.invR; a.invM_out = st.invM; a.sc = st.sc; a.n_iter = n_iter; a.rout = st.R;
  double* R0 = c.partials;
  double* R1 = c.partials + c.pstride;
  a.part_out = R0;
  launch2_sip(c, M_INIT_PCG, a);
  for (int it = 0; it < n_iter; ++it) {
    SipArgs k = sip_base(c, alpha);
    k.sc = st.sc; k.n_iter = n_iter; k.it = it; k.first = (it == 0);
    k.r = st.R + (size_t)it * N; k.invM = st.invM; k.invR = st.invR;
    k.pold = st.P + (size_t)(it > 0 ? it - 1 : 0) * N; k.pnew = st.P + (size_t)it * N; k.y = st.AP + (size_t)it * N;
    k.xout = x; k.part_in = R0; k.part_out = R1;
    launch2_sip(c, M_PCG_A, k);
    SipArgs m = sip_base(c, alpha);
    m.sc = st.sc; m.n_iter = n_iter; m.it = it;
    m.xout = x; m.pold = st.P + (size_t)it * N; m.y = st.AP + (size_t)it * N;
    m.r = st.R + (size_t)it * N; m.rout = st.R + (size_t)(it + 1) * N; m.invM = st.invM;
    m.part_in = R1; m.part_out = R0;
    launch2_sip(c, it + 1 < n_iter ? M_PCG_B : M_PCG_BX, m);
  }
  launch_pcg_final(c, R0, st.sc, n_iter);
}

// Damped Jacobi (jacobi_fixed!, elliptic.jl:354-361) keeping x_k and A x_k.
static void solve2_store_jac(Ctx& c, const double* qs, double* x, double alpha, int n_iter, StageStore& st) {
  const size_t N = c.N;
  SipArgs a = sip_base(c, alpha);
  a.mx = qs; a.my = qs + N; a.rho = qs + 2 * N; a.x = x;
  a.invR_out = st.invR; a.invM_out = st.invM; a.b_out = st.b;
  launch2_sip(c, M_INIT_JAC, a);
  if (n_iter > 0) launch_copy(c, st.P, x, N);
  for (int k = 0; k < n_iter; ++k) {
    SipArgs m = sip_base(c, alpha);
    m.x = st.P + (size_t)k * N; m.xout = (k + 1 < n_iter) ? st.P + (size_t)(k + 1) * N : x;
    m.y = st.AP + (size_t)k * N; m.b = st.b; m.invM = st.invM; m.invR = st.invR; m.omega = 2.0 / 3.0;
    launch2_sip(c, M_JACOBI, m);
  }
}

// Chebyshev-accelerated Jacobi (chebyshev_fixed!, elliptic.jl:370-397) keeping x_k and r_k.
static void solve2_store_cheb(Ctx& c, const double* qs, double* x, double alpha, int n_iter, StageStore& st) {
  const size_t N = c.N;
  SipArgs a = sip_base(c, alpha);
  a.mx = qs; a.my = qs + N; a.rho = qs + 2 * N; a.x = x;
  a.invR_out = st.invR; a.invM_out = st.invM; a.b_out = st.b; a.rout = st.R;
  launch2_sip(c, M_INIT_CHEB, a);
  double theta;
  std::vector<double> c1, c2;
  cheb_coeffs(c.p, n_iter, theta, c1, c2);
  double* Dd[2] = {c.dA, c.dB};
  for (int k = 1; k <= n_iter; ++k) {
    SipArgs m = sip_base(c, alpha);
    m.first = (k == 1); m.theta = theta; m.c1 = c1[k]; m.c2 = c2[k];
    m.x = (k == 1) ? x : st.P + (size_t)(k - 2) * N; m.d = Dd[(k - 1) % 2]; m.r = st.R + (size_t)(k - 1) * N;
    m.xout = st.P + (size_t)(k - 1) * N; m.dout = Dd[k % 2]; m.rout = st.R + (size_t)k * N;
    m.b = st.b; m.invM = st.invM; m.invR = st.invR;
    launch2_sip(c, M_CHEB, m);
  }
  if (n_iter > 0) launch_copy(c, x, st.P + (size_t)(n_iter - 1) * N, N);
}

static void solve2_store(Ctx& c, const double* qs, double* x, double alpha, int n_iter, int solver,
                         StageStore& st) {
  if (solver == IGR_SOLVER_JACOBI) solve2_store_jac(c, qs, x, alpha, n_iter, st);
  else if (solver == IGR_SOLVER_CHEBYSHEV) solve2_store_cheb(c, qs, x, alpha, n_iter, st);
  else solve2_store_pcg(c, qs, x, alpha, n_iter, st);
}

// One forward step from (q^n in W.Q[0], Sigma_0 in W.Sg[0]) storing what the
// reverse needs; q^{n+1} is written to out (4N), Sigma_3 stays in W.Sg[3].
static void step_store(Ctx& c, StepView W, double dt, double gamma, double alpha, int n_iter, int solver,
                       double* out) {
  const size_t N = c.N;
  const double cdt[3] = {dt, 1.0 / 4.0 * dt, 2.0 / 3.0 * dt};
  for (int s = 0; s < 3; ++s) {
    const double* qs = W.Q[s];
    if (alpha > 0.0) {
      launch_copy(c, W.Sg[s + 1], W.Sg[s], N);
      solve2_store(c, qs, W.Sg[s + 1], alpha, n_iter, solver, W.st[s]);
    } else {
      launch_fill(c, W.Sg[s + 1], N, 0.0);
    }
    HypArgs h = hyp_base(c, gamma);
    h.q = qs; h.S = W.Sg[s + 1]; h.qn = W.Q[0]; h.mode = s + 1; h.cdt = cdt[s];
    h.out = (s < 2) ? W.Q[s + 1] : out;
    launch2_hyp(c, h);
  }
}

// Reverse of one fixed-iteration PCG solve whose iterates are in st.
// xbar: adjoint of the solve output; x0: warm start; qs: stage state.
// Adds the state adjoint into qbar, writes the warm-start adjoint to x0bar.
static void solve2_rev_pcg(Ctx& c, GradWork& W, const double* qs, const double* x0, const double* xbar,
                           double alpha, int n, StageStore& st, double* qbar, double* x0bar) {
  const size_t N = c.N;
  launch_fill(c, W.irbar, N, 0.0);
  launch_fill(c, W.imbar, N, 0.0);
  RevArgs a;
  a.g = geo2(c); a.B = c.basis; a.alpha = alpha; a.n_iter = n;
  a.invR = st.invR; a.invM = st.invM; a.P = st.P; a.AP = st.AP; a.R = st.R; a.sc = st.sc; a.rs = st.rs;
  a.xbar = xbar; a.irbar = W.irbar; a.imbar = W.imbar; a.abar_blk = W.ablk;
  a.partials = c.partials; a.counter = c.counter;
  double* R0 = c.partials;
  double* R1 = c.partials + c.pstride;
  if (n > 0) {
    a.part_out = R0;
    launch2_rev(c, RV_PCG_START, a);
    for (int k = n - 1; k >= 0; --k) {
      const int m = n - 1 - k;
      RevArgs r = a;
      r.k = k; r.first = (k == n - 1);
      r.part_in = (m % 2) ? R1 : R0; r.part_out = (m % 2) ? R0 : R1;
      r.rbar_in = W.RB[(m + 1) % 2]; r.rbar_out = W.RB[m % 2];
      r.pbar_in = W.PB[(m + 1) % 2]; r.pbar_out = W.PB[m % 2];
      launch2_rev(c, RV_PCG_ITER, r);
    }
  }
  RevArgs f = a;
  f.rbar_in = W.RB[(n + 1) % 2]; f.pbar_in = W.PB[(n + 1) % 2];
  if (n > 0) { f.rbar_in = W.RB[(n - 1) % 2]; f.pbar_in = W.PB[(n - 1) % 2]; }
  f.x0 = x0; f.q = qs; f.qbar = qbar; f.x0bar = x0bar;
  launch2_rev(c, RV_PCG_FIN, f);
}

// Reverse of damped Jacobi: x-bar_k = x-bar_{k+1} - A (omega x-bar_{k+1} / M), plus
// the adjoints of 1/M, b and 1/rho (and alpha) of every iteration.
static void solve2_rev_jac(Ctx& c, GradWork& W, const double* qs, const double* x0, const double* xbar,
                           double alpha, int n, StageStore& st, double* qbar, double* x0bar) {
  const size_t N = c.N;
  launch_fill(c, W.irbar, N, 0.0);
  launch_fill(c, W.imbar, N, 0.0);
  launch_fill(c, W.bbar, N, 0.0);
  RevArgs a;
  a.g = geo2(c); a.B = c.basis; a.alpha = alpha; a.n_iter = n;
  a.invR = st.invR; a.invM = st.invM; a.X = st.P; a.AX = st.AP; a.b = st.b;
  a.irbar = W.irbar; a.imbar = W.imbar; a.bbar = W.bbar; a.abar_blk = W.ablk; a.omega = 2.0 / 3.0;
  const double* cur = xbar;
  for (int k = n - 1; k >= 0; --k) {
    const int m = n - 1 - k;
    RevArgs r = a;
    r.k = k; r.xbar = cur; r.xbar_out = W.RB[m % 2];
    launch2_rev(c, RV_JAC_ITER, r);
    cur = W.RB[m % 2];
  }
  RevArgs f = a;
  f.xbar = cur; f.x0 = x0; f.q = qs; f.qbar = qbar; f.x0bar = x0bar;
  launch2_rev(c, RV_JAC_FIN, f);
}

// Reverse of Chebyshev: per iteration k = n..1 the adjoints of x_k, d_k, r_{k-1};
// the final kernel reverses r_0 = b - A x_0 and the set-up.
static void solve2_rev_cheb(Ctx& c, GradWork& W, const double* qs, const double* x0, const double* xbar,
                            double alpha, int n, StageStore& st, double* qbar, double* x0bar) {
  const size_t N = c.N;
  launch_fill(c, W.irbar, N, 0.0);
  launch_fill(c, W.imbar, N, 0.0);
  launch_fill(c, W.bbar, N, 0.0);
  double theta;
  std::vector<double> c1, c2;
  cheb_coeffs(c.p, n, theta, c1, c2);
  RevArgs a;
  a.g = geo2(c); a.B = c.basis; a.alpha = alpha; a.n_iter = n;
  a.invR = st.invR; a.invM = st.invM; a.X = st.P; a.R = st.R; a.b = st.b; a.theta = theta;
  a.irbar = W.irbar; a.imbar = W.imbar; a.bbar = W.bbar; a.abar_blk = W.ablk;
  const double* cur = xbar;
  for (int k = n; k >= 1; --k) {
    const int m = n - k;
    RevArgs r = a;
    r.k = k; r.first = (k == n); r.cc1 = c1[k]; r.cc2 = c2[k];
    r.xbar = cur; r.xbar_out = W.XB[m % 2];
    r.pbar_in = W.PB[(m + 1) % 2]; r.pbar_out = W.PB[m % 2];
    r.rbar_in = W.RB[(m + 1) % 2]; r.rbar_out = W.RB[m % 2];
    launch2_rev(c, RV_CHEB_ITER, r);
    cur = W.XB[m % 2];
  }
  RevArgs f = a;
  f.xbar = cur; f.rbar_in = (n > 0) ? W.RB[(n - 1) % 2] : nullptr;
  f.x0 = x0; f.q = qs; f.qbar = qbar; f.x0bar = x0bar;
  launch2_rev(c, RV_CHEB_FIN, f);
}

static void solve2_rev(Ctx& c, GradWork& W, const double* qs, const double* x0, const double* xbar, double alpha,
                       int n, int solver, StageStore& st, double* qbar, double* x0bar) {
  if (solver == IGR_SOLVER_JACOBI) solve2_rev_jac(c, W, qs, x0, xbar, alpha, n, st, qbar, x0bar);
  else if (solver == IGR_SOLVER_CHEBYSHEV) solve2_rev_cheb(c, W, qs, x0, xbar, alpha, n, st, qbar, x0bar);
  else solve2_rev_pcg(c, W, qs, x0, xbar, alpha, n, st, qbar, x0bar);
}

// Reverse of one step stored by step_store.  W.QBnext holds the adjoint of
// q^{n+1}; sbar_next the adjoint of Sigma_3 (null if none).  On return W.QBn
// holds the adjoint of q^n and W.Scarry the adjoint of Sigma_0.
static void step_rev(Ctx& c, GradWork& W, StepView V, double dt, double gamma, double alpha, int n_iter, int solver,
                     bool has_sbar) {
  const size_t N = c.N;
  launch_fill(c, W.QBn, 4 * N, 0.0);
  const bool solve = alpha > 0.0;
  struct Stage { double* qbar_out; double* qbar_stage; int acc; double c_dt, c_prev, c_n; };
  const Stage sg[3] = {
      {W.QBa, W.QBn, 1, dt, 1.0, 0.0},
      {W.QBb, W.QBa, 0, 1.0 / 4.0 * dt, 1.0 / 4.0, 3.0 / 4.0},
      {W.QBnext, W.QBb, 0, 2.0 / 3.0 * dt, 2.0 / 3.0, 1.0 / 3.0},
  };
  bool sbar_valid = has_sbar && solve;
  for (int s = 2; s >= 0; --s) {
    HypRevArgs h;
    h.g = geo2(c); h.B = c.basis; h.gamma = gamma;
    h.q = V.Q[s]; h.S = V.Sg[s + 1];
    h.qbar_out = sg[s].qbar_out; h.qbar_stage = sg[s].qbar_stage; h.accumulate = sg[s].acc;
    h.qbar_n = (s > 0) ? W.QBn : nullptr;
    h.sbar_in = sbar_valid ? W.Scarry : nullptr; h.sbar_out = W.Xbar;
    h.c_dt = sg[s].c_dt; h.c_prev = sg[s].c_prev; h.c_n = sg[s].c_n;
    h.gbar_blk = W.gblk;
    launch2_hyp_rev(c, h);
    if (solve) {
      solve2_rev(c, W, V.Q[s], V.Sg[s], W.Xbar, alpha, n_iter, solver, V.st[s], sg[s].qbar_stage, W.Scarry);
      sbar_valid = true;
    }
  }
  if (!solve) launch_fill(c, W.Scarry, N, 0.0);
}

// ---- binomial checkpointing -------------------------------------------------
// cost(n, s): forward steps (without storage) spent reversing n steps whose
// first state is checkpointed, with s free slots; split(n, s): steps to
// advance before placing the next checkpoint.  With an iterate cache of K
// whole steps (2D), a segment of at most L = K + 1 steps is a window: one
// forward pass from its first state stores the iterates of its first K steps
// in the cache, its last step is reversed as a base step, then the cached
// ones, with no recomputation; every step is stored once either way, so a
// window costs no forward steps.  A segment with no free slot advances to
// its last window, reverses it, and does the same for the rest from its own
// first state.  L = 1 is the plain binomial (Revolve) schedule.
struct Schedule {
  std::vector<std::vector<long long>> cost;
  std::vector<std::vector<int>> split;
  bool store_all = false;
  int win_len = 1;
  void build(int n, int s, int L) {
    if (s >= n - 1) { store_all = true; return; }   // every step gets a slot: advance one step at a time
    win_len = std::max(L, 1);
    cost.assign(n + 1, std::vector<long long>(s + 1, 0));
    split.assign(n + 1, std::vector<int>(s + 1, 1));
    for (int m = 1; m <= n; ++m)
      for (int k = 0; k <= s; ++k) {
        if (m <= win_len) { cost[m][k] = 0; continue; }
        if (k == 0) { cost[m][k] = (m - win_len) + cost[m - win_len][0]; continue; }
        long long best = -1; int bj = 1;
        for (int j = 1; j < m; ++j) {
          const long long v = j + cost[m - j][k - 1] + cost[j][k];
          if (best < 0 || v < best) { best = v; bj = j; }
        }
        cost[m][k] = best; split[m][k] = bj;
      }
  }
  int split_of(int n, int s) const {
    if (store_all) return 1;
    s = std::min(s, (int)split[n].size() - 1);
    return split[n][s];
  }
};

struct GradRun {
  Ctx& c;
  GradWork* W = nullptr;     // 2D
  GradWork1* W1 = nullptr;   // 1D
  double dt, gamma, alpha;
  int n_iter, solver;
  Schedule sch;
  int obj;
  double prm[4];
  bool first_base = true;
  long long fwd_steps = 0;
  int n_total = 0;
  int kfirst = 1 << 30;   // 2D: steps kfirst .. n_total-2 keep their iterates from the forward sweep

  bool kept(int m) const { return c.dim == 2 && m >= kfirst && m < n_total - 1; }
  double* slot(int k) const { return c.dim == 1 ? W1->ck[k] : W->ck[k]; }
  void save(int k) {
    launch_copy(c, slot(k), c.q, c.nv * c.N);
    launch_copy(c, slot(k) + c.nv * c.N, c.S, c.N);
  }
  void restore(int k) {
    launch_copy(c, c.q, slot(k), c.nv * c.N);
    launch_copy(c, c.S, slot(k) + c.nv * c.N, c.N);
  }
  // advance from step index `at` by `steps` steps (the state is in c.q, c.S)
  void advance(int at, int steps) {
    if (steps <= 0) return;
    if (c.dim == 2) {
      for (int m = at; m < at + steps; ++m) {
        if (kept(m)) {
          StepSet& S = W->sets[m - kfirst];
          launch_copy(c, S.Q[0], c.q, 4 * c.N);
          launch_copy(c, S.Sg[0], c.S, c.N);
          step_store(c, S.view(), dt, gamma, alpha, n_iter, solver, c.q);
          launch_copy(c, c.S, S.Sg[3], c.N);
        } else {
          step2(c, dt, gamma, alpha, n_iter, solver);
        }
      }
      fwd_steps += steps;
      return;
    }
    if (c.dim == 1) {
      K1Args a = k1_base(c, gamma, alpha, n_iter, solver);
      a.op = OP_RUN; a.n_steps = steps; a.dt = dt; a.zero_sigma = 0; a.final_solve = 0;
      launch1_run(c, a, k1_block(c));
    } else {
      for (int s = 0; s < steps; ++s) step2(c, dt, gamma, alpha, n_iter, solver);
    }
    fwd_steps += steps;
  }
  // reverse step m: from its kept iterates, or recomputed (with storage) from
  // its start state in (c.q, c.S)
  void base_step(int m) {
    if (kept(m)) {
      step_rev(c, *W, W->sets[m - kfirst].view(), dt, gamma, alpha, n_iter, solver, !first_base);
      first_base = false;
      std::swap(W->QBn, W->QBnext);
      return;
    }
    ++fwd_steps;
    if (c.dim == 1) {
      K1AdjArgs& b = W1->base;
      launch_copy(c, b.Q[0], c.q, 3 * c.N);
      launch_copy(c, b.Sg[0], c.S, c.N);
      K1AdjArgs a = b;
      a.f = k1_base(c, gamma, alpha, n_iter, solver);
      a.f.q = b.Q[0];
      a.dt = dt; a.obj = obj; a.out = c.q;
      a.op = OPA_STORE;
      launch1_adjoint(c, a, k1_block(c));
      if (first_base) {
        a.op = OPA_OBJ; a.f.q = c.q;
        launch1_adjoint(c, a, k1_block(c));
        a.f.q = b.Q[0];
      }
      a.op = OPA_REV; a.has_sbar = first_base ? 0 : 1;
      launch1_adjoint(c, a, k1_block(c));
      first_base = false;
      std::swap(b.QBn, b.QBnext);
      return;
    }
    launch_copy(c, W->Q[0], c.q, 4 * c.N);
    launch_copy(c, W->Sg[0], c.S, c.N);
    step_store(c, W->view(), dt, gamma, alpha, n_iter, solver, c.q);
    if (first_base) {
      // q(T) is now in c.q: the objective and its gradient seed the sweep
      objective_seed();
      first_base = false;
      step_rev(c, *W, W->view(), dt, gamma, alpha, n_iter, solver, false);
    } else {
      step_rev(c, *W, W->view(), dt, gamma, alpha, n_iter, solver, true);
    }
    std::swap(W->QBn, W->QBnext);
  }
  // reverse the window [a, b) (b - a <= cache + 1) from its first state, in
  // (c.q, c.S) already or restored from slot k: steps a .. b-2 run once,
  // storing their iterates in the cache, step b-1 is a base step, then the
  // cached steps are reversed
  void window(int a, int b, int k, bool restore_first) {
    if (restore_first) restore(k);
    for (int m = a; m < b - 1; ++m) {
      StepSet& S = W->sets[m - a];
      launch_copy(c, S.Q[0], c.q, 4 * c.N);
      launch_copy(c, S.Sg[0], c.S, c.N);
      step_store(c, S.view(), dt, gamma, alpha, n_iter, solver, c.q);
      launch_copy(c, c.S, S.Sg[3], c.N);
    }
    base_step(b - 1);
    for (int m = b - 2; m >= a; --m) {
      step_rev(c, *W, W->sets[m - a].view(), dt, gamma, alpha, n_iter, solver, !first_base);
      first_base = false;
      std::swap(W->QBn, W->QBnext);
    }
  }
  // reverse steps [a, b): the state at a is in `slot`; the slots after it are free
  void reverse_range(int a, int b, int k, int free_slots) {
    const int n = b - a;
    if (n == 1) {
      if (!kept(a)) restore(k);
      base_step(a);
      return;
    }
    const int L = sch.win_len;
    if (!sch.store_all && L > 1 && n <= L) {
      window(a, b, k, true);
      return;
    }
    const int fs = std::min(free_slots, n - 1);
    if (fs == 0) {
      if (!sch.store_all && L > 1) {
        restore(k);
        advance(a, n - L);
        window(b - L, b, k, false);
        reverse_range(a, b - L, k, 0);
        return;
      }
      for (int m = b - 1; m >= a; --m) { restore(k); advance(a, m - a); base_step(m); }
      return;
    }
    const int j = sch.split_of(n, fs);
    restore(k);
    advance(a, j);
    save(k + 1);
    reverse_range(a + j, b, k + 1, free_slots - 1);
    reverse_range(a, a + j, k, free_slots);
  }
  void objective_seed();
};

void GradRun::objective_seed() {
  ObjArgs o;
  o.g = geo2(c); o.B = c.basis; o.obj = obj;
  for (int k = 0; k < 4; ++k) o.prm[k] = prm[k];
  o.Lx = c.Lx; o.Ly = c.Ly; o.gamma = gamma; o.q = c.q;
  o.out = W->dsc; o.qbar = W->QBnext; o.gbar = W->dsc + 3;
  o.partials = c.partials; o.counter = c.counter;
  launch2_obj(c, o, false);
  launch2_obj(c, o, true);
}

}  // namespace igr

using namespace igr;

static const double KE_WINDOW_DEFAULT[4] = {0.55, 0.50, 0.15, 0.0};   // enzyme_adjoint.jl:26

extern "C" {

int igr_objective(igr_ctx* c, int objective, const double* prm, double gamma, double* J) {
  if (!c || !J || c->dim != 2) return IGR_EINVAL;
  if (objective < 0 || objective > 6) return seterr(*c, IGR_EINVAL, "unknown objective");
  int rc = begin_call(*c);
  if (rc) return rc;
  GradWork* W = nullptr;
  rc = alloc_work(*c, 0, 0, W);
  if (rc) return rc;
  ObjArgs o;
  o.g = geo2(*c); o.B = c->basis; o.obj = objective;
  const double* pp = prm ? prm : KE_WINDOW_DEFAULT;
  for (int k = 0; k < 4; ++k) o.prm[k] = pp[k];
  o.Lx = c->Lx; o.Ly = c->Ly; o.gamma = gamma; o.q = c->q; o.out = W->dsc; o.qbar = W->QBnext;
  o.gbar = W->dsc + 3; o.partials = c->partials; o.counter = c->counter;
  launch2_obj(*c, o, false);
  HIPCHK(*c, hipMemcpyAsync(J, W->dsc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return end_call(*c);
}

int igr_gradient(igr_ctx* c, int objective, const double* prm, int n_steps, double dt, double gamma, double alpha,
                 int n_iter, int solver, int ckpt_slots, double* J, double* qbar0, double* gammabar,
                 double* alphabar) {
  if (!c || n_steps < 1 || n_iter < 0) return c ? seterr(*c, IGR_EINVAL, "bad argument") : IGR_EINVAL;
  if (objective < 0 || objective > 6) return seterr(*c, IGR_EINVAL, "unknown objective");
  if (c->dim == 1 && !(objective == 0 || objective == 1 || objective == 2 || objective == 4))
    return seterr(*c, IGR_EINVAL, "unknown 1D objective (l2_density, l2_pressure, kinetic_energy, weighted_momentum)");
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "unknown solver");
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = ensure_scal(*c, n_iter);
  if (rc) return rc;
  const size_t N = c->N, nv = c->nv;
  int slots = ckpt_slots;
  if (slots <= 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(*c, hipMemGetInfo(&fr, &tot));
    const size_t per = (nv + 1) * N * sizeof(double);
    slots = (int)std::min<size_t>((size_t)n_steps, std::max<size_t>(2, (fr / 2) / per));
  }
  slots = std::max(2, std::min(slots, n_steps + 1));
  GradRun run{*c};
  run.dt = dt; run.gamma = gamma; run.alpha = alpha; run.n_iter = n_iter; run.solver = solver;
  double* dsc = nullptr;
  double* qb = nullptr;
  if (c->dim == 1) {
    rc = alloc_work1(*c, n_iter, slots, run.W1);
    if (rc) return seterr(*c, rc, "device memory for the gradient workspace");
    dsc = run.W1->base.dsc;
  } else {
    rc = alloc_work(*c, n_iter, slots, run.W);
    if (rc) return seterr(*c, rc, "device memory for the gradient workspace");
    launch_fill(*c, run.W->gblk, c->partials_n, 0.0);
    launch_fill(*c, run.W->ablk, c->partials_n, 0.0);
    dsc = run.W->dsc;
  }
  launch_fill(*c, dsc, 16, 0.0);
  // state at t = 0: q0 from the context, Sigma = 0 (enzyme_adjoint.jl:51)
  launch_fill(*c, c->S, N, 0.0);
  run.obj = objective;
  run.n_total = n_steps;
  const double* pp = prm ? prm : KE_WINDOW_DEFAULT;
  for (int k = 0; k < 4; ++k) run.prm[k] = pp[k];
  // The iterate cache: whole steps whose solver iterates are kept in HBM, as
  // many as 85% of the free memory holds (capped by IGR_OPT_KEEP).  With
  // every step checkpointed it keeps the last steps from the forward sweep;
  // otherwise it sets the schedule's window length.
  const bool store_all = slots - 1 >= n_steps - 1;
  int K = 0;
  if (c->dim == 2 && n_steps >= 2 && c->keep_steps != 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(*c, hipMemGetInfo(&fr, &tot));
    const size_t per = step_set_doubles(N, n_iter) * sizeof(double);
    const size_t budget = (size_t)(0.85 * (double)(fr + run.W->set_block_bytes + pool_idle_bytes()));
    K = (int)std::min<size_t>((size_t)(n_steps - 1), budget / per);
    if (c->keep_steps > 0) K = std::min(K, c->keep_steps);
    // the block grows only when it would at least double: growing it frees
    // and obtains it again (about 30 ms per GB of a just-released block),
    // while a few more kept steps save a few step recomputes (the free
    // memory seen here can differ by a set between calls)
    const int K_have = (int)run.W->sets.size();
    if (K_have > 0 && K > K_have && K < 2 * K_have) K = K_have;
    if (K < 1 || alloc_sets(*c, *run.W, K) != IGR_OK) K = 0;
  }
  run.sch.build(n_steps, slots - 1, store_all ? 1 : K + 1);
  if (store_all && K >= 1) run.kfirst = n_steps - 1 - K;
  run.save(0);
  run.reverse_range(0, n_steps, 0, slots - 1);
  // the last reversed step left dJ/dq0 in QBnext
  if (c->dim == 1) {
    qb = run.W1->base.QBnext;
  } else {
    const int nb = grid2_rblocks(*c);
    launch_sum_blocks(*c, run.W->gblk, grid2_hblocks(*c), dsc + 3);
    launch_sum_blocks(*c, run.W->ablk, nb, dsc + 4);
    qb = run.W->QBnext;
  }
  double h[5];
  HIPCHK(*c, hipMemcpyAsync(h, dsc, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (qbar0) HIPCHK(*c, hipMemcpyAsync(qbar0, qb, nv * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  rc = end_call(*c);
  if (rc) return rc;
  if (J) *J = h[0];
  if (gammabar) *gammabar = h[3];
  if (alphabar) *alphabar = h[4];
  return IGR_OK;
}

// VJP of compute_hyperbolic_rhs!: given dq-bar, returns q-bar, Sigma-bar, gamma-bar.
int igr_hyperbolic_vjp(igr_ctx* c, double gamma, const double* q, const double* sigma, const double* dqbar,
                       double* qbar, double* sigmabar, double* gammabar) {
  if (!c || c->dim != 2 || !q || !sigma || !dqbar) return IGR_EINVAL;
  int rc = begin_call(*c);
  if (rc) return rc;
  GradWork* W = nullptr;
  rc = alloc_work(*c, 0, 0, W);
  if (rc) return rc;
  const size_t N = c->N;
  HIPCHK(*c, hipMemcpyAsync(W->Q[0], q, 4 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipMemcpyAsync(W->Sg[0], sigma, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipMemcpyAsync(W->QBnext, dqbar, 4 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  launch_fill(*c, W->gblk, c->partials_n, 0.0);
  launch_fill(*c, W->dsc, 16, 0.0);
  HypRevArgs h;
  h.g = geo2(*c); h.B = c->basis; h.gamma = gamma; h.q = W->Q[0]; h.S = W->Sg[0];
  h.qbar_out = W->QBnext; h.qbar_stage = W->QBn; h.accumulate = 0; h.qbar_n = nullptr;
  h.sbar_in = nullptr; h.sbar_out = W->Xbar; h.c_dt = 1.0; h.c_prev = 0.0; h.c_n = 0.0; h.gbar_blk = W->gblk;
  launch2_hyp_rev(*c, h);
  launch_sum_blocks(*c, W->gblk, grid2_hblocks(*c), W->dsc + 3);
  if (qbar) HIPCHK(*c, hipMemcpyAsync(qbar, W->QBn, 4 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (sigmabar) HIPCHK(*c, hipMemcpyAsync(sigmabar, W->Xbar, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  double gb = 0.0;
  HIPCHK(*c, hipMemcpyAsync(&gb, W->dsc + 3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  rc = end_call(*c);
  if (gammabar) *gammabar = gb;
  return rc;
}

// VJP of apply_sip!: y = A(rho, alpha) x; given y-bar returns x-bar = A y-bar, rho-bar, alpha-bar.
int igr_apply_sip_vjp(igr_ctx* c, double alpha, const double* x, const double* rho, const double* ybar,
                      double* xbar, double* rhobar, double* alphabar) {
  if (!c || c->dim != 2 || !x || !rho || !ybar) return IGR_EINVAL;
  int rc = begin_call(*c);
  if (rc) return rc;
  GradWork* W = nullptr;
  rc = alloc_work(*c, 0, 0, W);
  if (rc) return rc;
  const size_t N = c->N;
  HIPCHK(*c, hipMemcpyAsync(W->RB[0], x, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipMemcpyAsync(W->RB[1], ybar, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipMemcpyAsync(W->Q[0], rho, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // 1/rho via the init kernel path is not needed: a reciprocal pass
  SipArgs ia = sip_base(*c, alpha);
  ia.mx = W->Q[1]; ia.my = W->Q[1]; ia.rho = W->Q[0]; ia.x = W->Q[1];
  launch_fill(*c, W->Q[1], N, 0.0);
  ia.invR_out = W->irbar; ia.invM_out = W->imbar; ia.b_out = W->PB[0];
  launch2_sip(*c, M_INIT_JAC, ia);
  launch_fill(*c, W->ablk, c->partials_n, 0.0);
  launch_fill(*c, W->dsc, 16, 0.0);
  RevArgs a;
  a.g = geo2(*c); a.B = c->basis; a.alpha = alpha;
  a.invR = W->irbar; a.v = W->RB[0]; a.wbar = W->RB[1]; a.y = W->PB[1]; a.irbar = W->Xbar; a.abar_blk = W->ablk;
  launch2_rev(*c, RV_VJP, a);
  if (xbar) HIPCHK(*c, hipMemcpyAsync(xbar, W->PB[1], N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (rhobar) HIPCHK(*c, hipMemcpyAsync(rhobar, W->Xbar, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  rc = end_call(*c);
  if (rc) return rc;
  // alpha-bar: <y-bar, K x> = (<y-bar, A x> - mass)/alpha, summed on the host from the block partials
  if (alphabar) {
    std::vector<double> blk(grid2_rblocks(*c));
    HIPCHK(*c, hipMemcpy(blk.data(), W->ablk, blk.size() * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0.0;
    for (double v : blk) s += v;
    *alphabar = s;
  }
  return IGR_OK;
}

// VJP of one solve_elliptic! (any solver) on the context state q with warm start
// sigma0: given sigma-bar (adjoint of the result) returns q-bar, the warm-start
// adjoint and alpha-bar.
int igr_solve_elliptic_vjp(igr_ctx* c, double alpha, int n_iter, int solver, const double* sigma0,
                           const double* sigmabar, double* qbar, double* sigma0bar, double* alphabar) {
  if (!c || c->dim != 2 || !sigma0 || !sigmabar || n_iter < 0) return IGR_EINVAL;
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "unknown solver");
  int rc = begin_call(*c);
  if (rc) return rc;
  GradWork* W = nullptr;
  rc = alloc_work(*c, n_iter, 0, W);
  if (rc) return rc;
  const size_t N = c->N;
  HIPCHK(*c, hipMemcpyAsync(W->Sg[0], sigma0, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipMemcpyAsync(W->Xbar, sigmabar, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  launch_copy(*c, W->Q[0], c->q, 4 * N);
  launch_copy(*c, W->Sg[1], W->Sg[0], N);
  solve2_store(*c, W->Q[0], W->Sg[1], alpha, n_iter, solver, W->st[0]);
  launch_fill(*c, W->QBn, 4 * N, 0.0);
  launch_fill(*c, W->ablk, c->partials_n, 0.0);
  launch_fill(*c, W->dsc, 16, 0.0);
  solve2_rev(*c, *W, W->Q[0], W->Sg[0], W->Xbar, alpha, n_iter, solver, W->st[0], W->QBn, W->Scarry);
  launch_sum_blocks(*c, W->ablk, grid2_rblocks(*c), W->dsc + 4);
  if (qbar) HIPCHK(*c, hipMemcpyAsync(qbar, W->QBn, 4 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (sigma0bar) HIPCHK(*c, hipMemcpyAsync(sigma0bar, W->Scarry, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  double ab = 0.0;
  HIPCHK(*c, hipMemcpyAsync(&ab, W->dsc + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  rc = end_call(*c);
  if (alphabar) *alphabar = ab;
  return rc;
}

}  // extern "C"

extern "C" int igr_release_cache(void) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  for (auto& kv : g_setpool) pool_drop_idle(kv.first);
  return IGR_OK;
}

namespace igr {
hipError_t dev_malloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes) == hipSuccess) return hipSuccess;
  (void)hipGetLastError();
  bool dropped = false;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    dropped = pool_drop_idle(current_device());
  }
  if (dropped && hipMalloc(p, bytes) == hipSuccess) return hipSuccess;
  (void)hipGetLastError();
  return hipErrorOutOfMemory;
}

void free_grad_work(const Ctx* c) {
  std::lock_guard<std::mutex> lk(g_work_mu);
  auto it = g_work.find(c);
  if (it != g_work.end()) { delete it->second; g_work.erase(it); }
  auto it1 = g_work1.find(c);
  if (it1 != g_work1.end()) { delete it1->second; g_work1.erase(it1); }
}
}  // namespace igr
