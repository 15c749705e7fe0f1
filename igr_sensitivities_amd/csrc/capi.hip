// C ABI (include/igr.h): context management and host orchestration of the
// 1D and 2D forward paths.  Every entry point returns an IGR_* status and
// keeps a message for igr_last_error().
#include <cmath>
#include <cstring>
#include <algorithm>
#include <map>
#include <tuple>

#include "igr_ctx.hpp"
#include "k1_common.hpp"
#include "k2_common.hpp"
#include "host_common.hpp"

namespace igr {

// launchers defined in the kernel translation units
void launch2_sip(const Ctx& c, int mode, const SipArgs& a);
void launch2_hyp(const Ctx& c, const HypArgs& a);
int grid2_blocks(const Ctx& c);
void launch_fill(const Ctx& c, double* x, size_t n, double v);
void launch_copy(const Ctx& c, double* dst, const double* src, size_t n);
void launch1_run(const Ctx& c, const K1Args& a, int block);
void set_kernel_timer(hipEvent_t t0, hipEvent_t t1);
void launch_pcg_final(const Ctx& c, const double* part, double* sc, int nit);
void free_grad_work(const Ctx* c);
void free_tan_bufs(const Ctx* c);

double* Ctx::alloc(size_t n) {
  void* p = nullptr;
  if (dev_malloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) return nullptr;
  owned.push_back(p);
  return static_cast<double*>(p);
}

// ---------------------------------------------------------------- basis --
// Legendre-Gauss-Lobatto nodes/weights (Newton on (1-x^2)P_p'), barycentric
// weights and the differentiation matrix with diagonal = -(row sum of the
// off-diagonal), as DGBasis builds them (IGR_adjoints_2d/basis.jl:19-68).
void make_basis(int p, double* xi, double* w, double* Dm) {
  const int n = p + 1;
  for (int i = 0; i < n; ++i) {
    double x = -std::cos(M_PI * i / p);
    for (int it = 0; it < 100; ++it) {
      double P0 = 1.0, P1 = x;
      for (int k = 2; k <= p; ++k) { double P2 = ((2.0 * k - 1.0) * x * P1 - (k - 1.0) * P0) / k; P0 = P1; P1 = P2; }
      const double dx = (x * P1 - P0) / (n * P1);
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    xi[i] = x;
  }
  xi[0] = -1.0; xi[n - 1] = 1.0;
  for (int i = 0; i < n / 2; ++i) { const double h = 0.5 * (xi[n - 1 - i] - xi[i]); xi[i] = -h; xi[n - 1 - i] = h; }
  if (n % 2 == 1) xi[n / 2] = 0.0;
  for (int i = 0; i < n; ++i) {
    double x = xi[i], P0 = 1.0, P1 = x;
    for (int k = 2; k <= p; ++k) { double P2 = ((2.0 * k - 1.0) * x * P1 - (k - 1.0) * P0) / k; P0 = P1; P1 = P2; }
    w[i] = 2.0 / (p * (p + 1.0) * P1 * P1);
  }
  double lam[MAXNP];
  for (int j = 0; j < n; ++j) {
    lam[j] = 1.0;
    for (int i = 0; i < n; ++i) if (i != j) lam[j] *= (xi[j] - xi[i]);
    lam[j] = 1.0 / lam[j];
  }
  for (int q = 0; q < n * n; ++q) Dm[q] = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      if (i == j) continue;
      Dm[i * n + j] = (lam[j] / lam[i]) / (xi[i] - xi[j]);
      Dm[i * n + i] -= Dm[i * n + j];
    }
}

Geo2 geo2(const Ctx& c) {
  Geo2 g;
  g.Nex = c.Nex; g.Ney = c.Ney;
  g.dx = c.Lx / c.Nex; g.dy = c.Ly / c.Ney;
  g.Jx = g.dx / 2.0; g.Jy = g.dy / 2.0;
  g.invJx = 1.0 / g.Jx; g.invJy = 1.0 / g.Jy;
  g.eta = (double)((c.p + 1) * (c.p + 1));
  return g;
}

double cheb_lmax(int p) {
  static const double tab[4] = {1.5, 2.0, 2.1, 2.2};   // elliptic.jl:375
  return tab[std::min(p, 4) - 1];
}

int ensure_scal(Ctx& c, int n_iter) {
  const size_t need = 4 * (size_t)std::max(n_iter, 1) + 1;
  if (c.scal_n >= need) return IGR_OK;
  c.scal = c.alloc(need);
  if (!c.scal) return IGR_ENOMEM;
  c.scal_n = need;
  return IGR_OK;
}

// ------------------------------------------------------------- 2D solve --
SipArgs sip_base(const Ctx& c, double alpha) {
  SipArgs a;
  a.g = geo2(c); a.B = c.basis; a.alpha = alpha;
  a.partials = c.partials; a.counter = c.counter; a.flag = c.flag;
  return a;
}

// Sigma solve on stage state qs (device, 4N); warm start and result c.S.
void solve2(Ctx& c, const double* qs, double alpha, int n_iter, int solver, double* sc, double* x,
            const double* b_in) {
  solve2_fields(c, qs, qs + c.N, qs + 2 * c.N, alpha, n_iter, solver, sc, x, b_in);
}

// The same solve on separate momentum and density fields (device, N each).
void solve2_fields(Ctx& c, const double* mx, const double* my, const double* rho, double alpha, int n_iter,
                   int solver, double* sc, double* x, const double* b_in) {
  const size_t N = c.N;
  if (!x) x = c.S;
  SipArgs a = sip_base(c, alpha);
  a.mx = mx; a.my = my; a.rho = rho; a.x = x; a.b_in = b_in;
  a.invR_out = c.invR; a.invM_out = c.invM; a.sc = sc; a.n_iter = n_iter;
  if (solver == IGR_SOLVER_PCG) {
    double* R0 = c.partials;
    double* R1 = c.partials + c.pstride;
    a.rout = c.r; a.part_out = R0;
    launch2_sip(c, M_INIT_PCG, a);
    double* P[2] = {c.pA, c.pB};
    for (int it = 0; it < n_iter; ++it) {
      SipArgs k = sip_base(c, alpha);
      k.sc = sc; k.n_iter = n_iter; k.it = it; k.first = (it == 0);
      k.r = c.r; k.invM = c.invM; k.invR = c.invR;
      k.pold = P[(it + 1) % 2]; k.pnew = P[it % 2]; k.y = c.Ap; k.xout = x;
      k.part_in = R0; k.part_out = R1;
      launch2_sip(c, M_PCG_A, k);
      SipArgs m = sip_base(c, alpha);
      m.sc = sc; m.n_iter = n_iter; m.it = it;
      m.xout = x; m.pold = P[it % 2]; m.y = c.Ap; m.rout = c.r; m.invM = c.invM;
      m.part_in = R1; m.part_out = R0;
      launch2_sip(c, it + 1 < n_iter ? M_PCG_B : M_PCG_BX, m);
    }
    launch_pcg_final(c, R0, sc, n_iter);
  } else if (solver == IGR_SOLVER_JACOBI) {
    a.b_out = c.b;
    launch2_sip(c, M_INIT_JAC, a);
    double* X[2] = {x, c.xA};
    for (int it = 0; it < n_iter; ++it) {
      SipArgs k = sip_base(c, alpha);
      k.x = X[it % 2]; k.xout = X[(it + 1) % 2]; k.b = c.b; k.invM = c.invM; k.invR = c.invR;
      k.omega = 2.0 / 3.0;
      launch2_sip(c, M_JACOBI, k);
    }
    if (n_iter % 2) launch_copy(c, x, c.xA, N);
  } else {
    a.rout = c.r; a.b_out = c.b;
    launch2_sip(c, M_INIT_CHEB, a);
    const double lmin = 1.0, lmax = cheb_lmax(c.p);
    const double theta = (lmax + lmin) / 2.0, delta = (lmax - lmin) / 2.0;
    const double sigma = theta / std::max(delta, 1e-10);
    double rc = 1.0 / sigma;
    double* X[2] = {x, c.xA};
    double* Dd[2] = {c.dA, c.dB};
    double* R[2] = {c.r, c.r2};
    for (int k = 1; k <= n_iter; ++k) {
      SipArgs m = sip_base(c, alpha);
      m.first = (k == 1); m.theta = theta;
      if (k > 1) { const double rn = 1.0 / (2.0 * sigma - rc); m.c1 = rn * rc; m.c2 = 2.0 * rn / delta; rc = rn; }
      m.x = X[(k - 1) % 2]; m.d = Dd[(k - 1) % 2]; m.r = R[(k - 1) % 2];
      m.xout = X[k % 2]; m.dout = Dd[k % 2]; m.rout = R[k % 2];
      m.b = c.b; m.invM = c.invM; m.invR = c.invR;
      launch2_sip(c, M_CHEB, m);
    }
    if (n_iter % 2) launch_copy(c, x, c.xA, N);
  }
}

HypArgs hyp_base(const Ctx& c, double gamma) {
  HypArgs h;
  h.g = geo2(c); h.B = c.basis; h.gamma = gamma; h.flag = c.flag;
  h.q = nullptr; h.S = c.S; h.qn = c.q; h.out = nullptr; h.mode = 0; h.cdt = 0.0;
  return h;
}

// One stage: Sigma solve (or Sigma = 0), then hyperbolic RHS + stage update.
void stage2(Ctx& c, const double* qs, double gamma, double alpha, int n_iter, int solver, int mode, double cdt,
            double* out) {
  if (alpha > 0.0) solve2(c, qs, alpha, n_iter, solver, c.scal);
  else launch_fill(c, c.S, c.N, 0.0);
  HypArgs h = hyp_base(c, gamma);
  h.q = qs; h.mode = mode; h.cdt = cdt; h.out = out;
  launch2_hyp(c, h);
}

// One SSP-RK3 step (IGR_adjoints_2d/forward.jl:51-69).
void step2(Ctx& c, double dt, double gamma, double alpha, int n_iter, int solver) {
  stage2(c, c.q, gamma, alpha, n_iter, solver, 1, dt, c.q1);
  stage2(c, c.q1, gamma, alpha, n_iter, solver, 2, 1.0 / 4.0 * dt, c.q2);
  stage2(c, c.q2, gamma, alpha, n_iter, solver, 3, 2.0 / 3.0 * dt, c.q);
}

// ------------------------------------------------------------------- 1D --
K1Args k1_base(Ctx& c, double gamma, double alpha, int n_iter, int solver) {
  K1Args a;
  a.Ne = c.Nex; a.dx = c.Lx / c.Nex; a.J = a.dx / 2.0; a.invJ = 1.0 / a.J;
  a.eta = (double)((c.p + 1) * (c.p + 1));
  a.B = c.basis; a.gamma = gamma; a.alpha = alpha; a.n_iter = n_iter; a.solver = solver;
  a.cheb_lmax = cheb_lmax(c.p);
  a.q = c.q; a.q1 = c.q1; a.q2 = c.q2; a.S = c.S; a.dq = c.dq;
  a.invR = c.invR; a.invM = c.invM; a.b = c.b; a.r = c.r; a.p = c.pA; a.z = c.pB; a.Ap = c.Ap;
  a.u = c.xA; a.g = c.dA; a.F = c.k1F; a.fstar = c.k1face; a.fjump = c.k1face + 3 * c.Nex;
  a.favg = c.k1face + 4 * c.Nex; a.fsig = c.k1face + 5 * c.Nex; a.flag = c.flag;
  a.bs = c.dB;
  return a;
}
int k1_block(const Ctx& c) {  // threads of the 1D whole-instance workgroup
  const size_t N = c.N;
  return N >= 1024 ? 1024 : (int)(((N + 63) / 64) * 64);
}
// ---- tile shapes (p = 3) -------------------------------------------------
bool tile_ok(int tx, int ty, bool hyp) {
  return (tx == 4 && ty == 4) || (tx == 8 && ty == 4) || (tx == 2 && ty == 2) || (!hyp && tx == 8 && ty == 8);
}
// SIP-family tiles at p = 4
bool tile_ok5(int tx, int ty) {
  return (tx == 2 && ty == 4) || (tx == 4 && ty == 4) || (tx == 5 && ty == 5) || (tx == 8 && ty == 5);
}
void set_tiles(Ctx& c, int tx, int ty, int htx, int hty) {
  if (tx == c.tile_x && ty == c.tile_y && htx == c.htile_x && hty == c.htile_y) return;
  (void)hipStreamSynchronize(c.stream);
  destroy_graphs(c);   // captured steps hold launches of the old shapes
  c.tile_x = tx; c.tile_y = ty; c.htile_x = htx; c.htile_y = hty;
}

// Default shapes by mesh size (elements E = Nex * Ney), from timing every
// shape on MI355X at 16^2 .. 1024^2: up to ~64^2 elements the launches are
// latency-bound and 4x4 tiles are fastest; from ~96^2 8x4 tiles win for
// every family (fewer partial sums for every workgroup to read than 4x4,
// and two or more 8-wave workgroups per CU hide each other's barriers where
// one 16-wave 8x8 workgroup cannot).  The shape of a given mesh is fixed, so
// repeated runs give bit-identical results; the autotuner may choose other
// forward shapes.
void default_tiles(Ctx& c) {
  if (c.np == 5) {
    // p = 4: the SIP family's tile by mesh size (the other families keep
    // 2x4).  A PCG-60 step of the triple-Sedov data on MI355X, 4x4 / 5x5 /
    // 8x5 tiles (x/r update on max(256, tiles / 16) workgroups beyond 512
    // tiles): 64^2 1.96 / 2.08 / 2.31 ms; 96^2 2.96 / 2.69 / 2.45; 128^2
    // 3.64 / 3.74 / 3.70; 160^2 (the reference's p = 4 heatmap runs) 5.03 /
    // 5.04 / 4.91; 200^2 7.21 / 7.20 / 6.33; 256^2 10.45 / 10.90 / 10.55.
    const long long E = (long long)c.Nex * c.Ney;
    if (E >= 96 * 96) { c.tile_x = 8; c.tile_y = 5; }
    else if (E >= 32 * 32) { c.tile_x = 4; c.tile_y = 4; }
    else { c.tile_x = 2; c.tile_y = 4; }
    const int nt = ((c.Nex + c.tile_x - 1) / c.tile_x) * ((c.Ney + c.tile_y - 1) / c.tile_y);
    c.bgroups = nt > 512 ? std::max(256, (nt / 16 + 7) / 8 * 8) : 0;
    return;
  }
  if (c.np != 4) return;
  // p = 3: 4x4 tiles below 96^2 elements, 8x4 from there for every family
  // (8x8 SIP tiles, two 16-wave workgroups per CU once PCG_A fits 64
  // registers, ran PCG_A 24.4 us at 256^2 against 17.7 for 8x4; 1024^2 step
  // 103.5 against 88.5 ms).
  const long long E = (long long)c.Nex * c.Ney;
  if (E < 96 * 96) { c.tile_x = 4; c.tile_y = 4; c.htile_x = 4; c.htile_y = 4; }
  else { c.tile_x = 8; c.tile_y = 4; c.htile_x = 8; c.htile_y = 4; }
  c.rtile_x = c.tile_x;
  c.rtile_y = c.tile_y;
  // The PCG x/r update streams 7 fields per node and no halo: beyond 512
  // tiles it runs on max(512, tiles / 16) workgroups that each loop over
  // their tiles with the next tile's loads in flight, and the next PCG_A
  // reads that many partial sums instead of one per tile (measured on
  // MI355X with 8x4 tiles: 256^2 step 5.57 ms on 512 workgroups against
  // 6.82 on 256 and 5.95 on 1024; 512^2 17.3 ms on 256 or 512; 1024^2
  // 88.5 ms on 1024 or 2048; no change up to 128^2).
  const int nt = ((c.Nex + c.tile_x - 1) / c.tile_x) * ((c.Ney + c.tile_y - 1) / c.tile_y);
  c.bgroups = nt > 512 ? std::max(512, (nt / 16 + 7) / 8 * 8) : 0;
  // The reverse PCG iteration of the gradient keeps two 8-wave workgroups
  // per CU (128 VGPRs): beyond 512 reverse tiles it runs on 512 workgroups
  // that loop over their tiles, and the next iteration reduces 512 partial
  // sums per dot product instead of one per tile (256^2, 20-step gradient:
  // 0.277 s on 512 workgroups against 0.301 on 256, 0.312 on 1024 and 0.365
  // with one per tile).
  c.rgroups = nt > 512 ? 512 : 0;
}

// The autotuner: on the current state, time a 30-iteration PCG solve for
// every SIP-family shape and 20 hyperbolic stage launches for every
// hyperbolic-family shape (HIP events, best of 3), and keep the fastest.
// Results change only in the order of the solver's partial sums.
int autotune(Ctx& c) {
  // p = 3: three SIP-family and two hyperbolic-family shapes; p = 4: the
  // four SIP-family shapes (its hyperbolic family has one)
  const bool p4 = c.np == 5;
  const int sip3[3][2] = {{4, 4}, {8, 4}, {8, 8}}, sip4[4][2] = {{2, 4}, {4, 4}, {5, 5}, {8, 5}};
  const int (*sip)[2] = p4 ? sip4 : sip3;
  const int nsip = p4 ? 4 : 3, nhyp = p4 ? 1 : 2;
  const int hyp[2][2] = {{4, 4}, {8, 4}};
  const double h = c.Lx / c.Nex;
  const double alpha = 10.0 * (h / 4.0) * (h / 4.0), gamma = 1.4;
  const int n_it = 30;
  int rc = ensure_scal(c, n_it);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  auto timed = [&](auto&& body) {
    float best = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {
      (void)hipEventRecord(e0, c.stream);
      body();
      (void)hipEventRecord(e1, c.stream);
      (void)hipEventSynchronize(e1);
      float t = 0.f;
      (void)hipEventElapsedTime(&t, e0, e1);
      if (rep > 0) best = std::min(best, t);    // the first run loads code objects
    }
    return best;
  };
  int bt = 0, bh = 0;
  float tb = 1e30f, th = 1e30f;
  const int htx0 = c.htile_x, hty0 = c.htile_y;
  for (int k = 0; k < nsip; ++k) {
    set_tiles(c, sip[k][0], sip[k][1], htx0, hty0);
    const float t = timed([&]() {
      launch_copy(c, c.xA, c.S, c.N);
      solve2(c, c.q, alpha, n_it, IGR_SOLVER_PCG, c.scal, c.xA);
    });
    if (t < tb) { tb = t; bt = k; }
  }
  for (int k = 0; k < nhyp && !p4; ++k) {
    set_tiles(c, sip[bt][0], sip[bt][1], hyp[k][0], hyp[k][1]);
    HypArgs a = hyp_base(c, gamma);
    a.q = c.q; a.mode = 2; a.cdt = 1e-6; a.out = c.q2; a.qn = c.q;
    const float t = timed([&]() { for (int r = 0; r < 20; ++r) launch2_hyp(c, a); });
    if (t < th) { th = t; bh = k; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (p4) set_tiles(c, sip[bt][0], sip[bt][1], htx0, hty0);
  else set_tiles(c, sip[bt][0], sip[bt][1], hyp[bh][0], hyp[bh][1]);
  return IGR_OK;
}

int run1(Ctx& c, const K1Args& a) {
  launch1_run(c, a, k1_block(c));
  HIPCHK(c, hipGetLastError());
  return IGR_OK;
}

}  // namespace igr

using namespace igr;

// =================================================================== ABI ==
extern "C" {

int igr_create(igr_ctx** out, int dim, int p, int nex, int ney, double lx, double ly) {
  if (!out) return IGR_EINVAL;
  *out = nullptr;
  if ((dim != 1 && dim != 2) || p < 1 || p > 4 || nex < 1 || (dim == 2 && ney < 1) || !(lx > 0.0) ||
      (dim == 2 && !(ly > 0.0)))
    return IGR_EINVAL;
  igr_ctx* c = new igr_ctx();
  c->dim = dim; c->p = p; c->np = p + 1; c->Nex = nex; c->Ney = dim == 2 ? ney : 1;
  c->Lx = lx; c->Ly = dim == 2 ? ly : 1.0;
  c->nv = dim == 2 ? 4 : 3;
  c->N = dim == 2 ? (size_t)c->np * c->np * nex * ney : (size_t)c->np * nex;
  double xi[MAXNP], w[MAXNP], Dm[MAXNP * MAXNP];
  make_basis(p, xi, w, Dm);
  for (int i = 0; i < c->np; ++i) { c->basis.xi[i] = xi[i]; c->basis.w[i] = w[i]; }
  for (int k = 0; k < c->np * c->np; ++k) c->basis.D[k] = Dm[k];
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return IGR_EHIP; }
  const size_t N = c->N, nvN = c->nv * N;
  bool ok = true;
  auto A = [&](double*& ptr, size_t n) { ptr = c->alloc(n); ok = ok && ptr; };
  A(c->q, nvN); A(c->S, N); A(c->q1, nvN); A(c->q2, nvN); A(c->dq, nvN);
  A(c->invR, N); A(c->invM, N); A(c->b, N); A(c->r, N); A(c->r2, N);
  A(c->pA, N); A(c->pB, N); A(c->Ap, N); A(c->xA, N); A(c->dA, N); A(c->dB, N); A(c->tmp, N);
  A(c->dscalar, 64);
  if (dim == 1) { A(c->k1F, 3 * N); A(c->k1face, 6 * (size_t)nex); A(c->k1args, sizeof(K1Args) / sizeof(double) + 1); }
  // two regions of 4 partial sums per workgroup, sized for the smallest tile
  // (2x2 elements) so any tile choice fits
  default_tiles(*c);
  c->pstride = 4 * (size_t)(((nex + 1) / 2) * ((c->Ney + 1) / 2)) + 64;
  c->partials_n = 2 * c->pstride;
  A(c->partials, c->partials_n);
  void* pp = nullptr;
  if (ok && dev_malloc(&pp, 64) == hipSuccess) { c->owned.push_back(pp); c->counter = (unsigned*)pp; } else ok = false;
  if (ok && dev_malloc(&pp, 64) == hipSuccess) { c->owned.push_back(pp); c->flag = (int*)pp; } else ok = false;
  if (ok) ok = ensure_scal(*c, 64) == IGR_OK;
  // on the context's own stream and waited for: the stream does not wait for
  // the null stream, so a null-stream memset could still be pending (behind
  // other work queued there) when the first kernel reads Sigma or the flags
  if (ok) ok = hipMemsetAsync(c->counter, 0, 64, c->stream) == hipSuccess &&
               hipMemsetAsync(c->flag, 0, 64, c->stream) == hipSuccess &&
               hipMemsetAsync(c->S, 0, N * sizeof(double), c->stream) == hipSuccess &&
               hipMemsetAsync(c->q, 0, nvN * sizeof(double), c->stream) == hipSuccess &&
               hipStreamSynchronize(c->stream) == hipSuccess;
  if (!ok) { igr_destroy(c); return IGR_ENOMEM; }
  *out = c;
  return IGR_OK;
}

int igr_destroy(igr_ctx* c) {
  if (!c) return IGR_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_graphs(*c);
  free_grad_work(c);
  free_tan_bufs(c);
  if (c->snaps) (void)hipFree(c->snaps);
  for (void* p : c->owned) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return IGR_OK;
}

const char* igr_last_error(const igr_ctx* c) { return c ? c->err.c_str() : "null context"; }

int igr_sizes(const igr_ctx* c, long long* n_nodes, int* n_vars) {
  if (!c) return IGR_EINVAL;
  if (n_nodes) *n_nodes = (long long)c->N;
  if (n_vars) *n_vars = c->nv;
  return IGR_OK;
}

int igr_set_option(igr_ctx* c, int option, int value) {
  if (!c) return IGR_EINVAL;
  if (option == IGR_OPT_GRAPHS) { c->use_graphs = value ? 1 : 0; return IGR_OK; }
  if (option == IGR_OPT_WAVE1D) { c->k1_wave = value ? 1 : 0; return IGR_OK; }
  if (option == IGR_OPT_TILE || option == IGR_OPT_HTILE) {
    const int tx = value / 100, ty = value % 100;
    const bool h = option == IGR_OPT_HTILE;
    if (c->np == 5) {
      if (h) return seterr(*c, IGR_EINVAL, "the hyperbolic family has one tile shape at p = 4");
      if (!tile_ok5(tx, ty)) return seterr(*c, IGR_EINVAL, "unsupported tile shape");
      set_tiles(*c, tx, ty, c->htile_x, c->htile_y);
      return IGR_OK;
    }
    if (!tile_ok(tx, ty, h)) return seterr(*c, IGR_EINVAL, "unsupported tile shape");
    if (c->np == 4) set_tiles(*c, h ? c->tile_x : tx, h ? c->tile_y : ty, h ? tx : c->htile_x, h ? ty : c->htile_y);
    return IGR_OK;
  }
  if (option == IGR_OPT_RTILE) {
    // the reverse kernels are launched directly (no captured step holds them)
    const int tx = value / 100, ty = value % 100;
    if (!tile_ok(tx, ty, false)) return seterr(*c, IGR_EINVAL, "unsupported tile shape");
    if (c->np == 4) { c->rtile_x = tx; c->rtile_y = ty; }
    return IGR_OK;
  }
  if (option == IGR_OPT_GROUPS) {
    if (value < 0) return seterr(*c, IGR_EINVAL, "workgroup count must be >= 0");
    if (value != c->bgroups) {
      (void)hipStreamSynchronize(c->stream);
      destroy_graphs(*c);   // captured steps hold launches of the old grid
      c->bgroups = value;
    }
    return IGR_OK;
  }
  if (option == IGR_OPT_RGROUPS) {
    // the reverse kernels are launched directly (no captured step holds them)
    if (value < 0) return seterr(*c, IGR_EINVAL, "workgroup count must be >= 0");
    c->rgroups = value;
    return IGR_OK;
  }
  if (option == IGR_OPT_KEEP) {
    if (value < -1) return seterr(*c, IGR_EINVAL, "kept steps must be >= -1");
    c->keep_steps = value;
    return IGR_OK;
  }
  if (option == IGR_OPT_AUTOTUNE) {
    if (c->dim != 2 || (c->np != 4 && c->np != 5)) return IGR_OK;   // one tile shape exists for the other cases
    int rc = begin_call(*c);
    if (rc) return rc;
    rc = autotune(*c);
    if (rc) return rc;
    return end_call(*c);
  }
  return seterr(*c, IGR_EINVAL, "unknown option");
}

int igr_get_option(const igr_ctx* c, int option, int* value) {
  if (!c || !value) return IGR_EINVAL;
  if (option == IGR_OPT_GRAPHS) *value = c->use_graphs;
  else if (option == IGR_OPT_TILE) *value = 100 * c->tile_x + c->tile_y;
  else if (option == IGR_OPT_HTILE) *value = 100 * c->htile_x + c->htile_y;
  else if (option == IGR_OPT_RTILE) *value = 100 * c->rtile_x + c->rtile_y;
  else if (option == IGR_OPT_GROUPS) *value = c->bgroups;
  else if (option == IGR_OPT_WAVE1D) *value = c->k1_wave;
  else if (option == IGR_OPT_KEEP) *value = c->keep_steps;
  else if (option == IGR_OPT_RGROUPS) *value = c->rgroups;
  else return IGR_EINVAL;
  return IGR_OK;
}

int igr_set_state(igr_ctx* c, const double* q, const double* sigma) {
  if (!c || !q) return IGR_EINVAL;
  HIPCHK(*c, hipMemcpyAsync(c->q, q, c->nv * c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (sigma) HIPCHK(*c, hipMemcpyAsync(c->S, sigma, c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipStreamSynchronize(c->stream));
  return IGR_OK;
}

int igr_get_state(igr_ctx* c, double* q, double* sigma) {
  if (!c) return IGR_EINVAL;
  if (q) HIPCHK(*c, hipMemcpyAsync(q, c->q, c->nv * c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (sigma) HIPCHK(*c, hipMemcpyAsync(sigma, c->S, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(*c, hipStreamSynchronize(c->stream));
  return IGR_OK;
}

int igr_hyperbolic_rhs(igr_ctx* c, double gamma, const double* q, const double* sigma, double* dq) {
  if (!c || !q || !sigma || !dq) return IGR_EINVAL;
  const size_t N = c->N, nvN = c->nv * N;
  int rc = begin_call(*c);
  if (rc) return rc;
  HIPCHK(*c, hipMemcpyAsync(c->q1, q, nvN * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(*c, hipMemcpyAsync(c->tmp, sigma, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (c->dim == 2) {
    HypArgs h = hyp_base(*c, gamma);
    h.q = c->q1; h.S = c->tmp; h.mode = 0; h.out = c->dq;
    launch2_hyp(*c, h);
  } else {
    K1Args a = k1_base(*c, gamma, 0.0, 0, 0);
    a.op = OP_HYP; a.q = c->q1; a.S = c->tmp;
    rc = run1(*c, a);
    if (rc) return rc;
  }
  HIPCHK(*c, hipMemcpyAsync(dq, c->dq, nvN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return end_call(*c);
}

int igr_apply_sip(igr_ctx* c, double alpha, const double* x, const double* rho, double* y) {
  if (!c || !x || !rho || !y) return IGR_EINVAL;
  const size_t N = c->N;
  int rc = begin_call(*c);
  if (rc) return rc;
  HIPCHK(*c, hipMemcpyAsync(c->xA, x, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (c->dim == 2) {
    HIPCHK(*c, hipMemcpyAsync(c->tmp, rho, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SipArgs a = sip_base(*c, alpha);
    a.x = c->xA; a.rho = c->tmp; a.y = c->Ap;
    launch2_sip(*c, M_APPLY, a);
  } else {
    // OP_APPLY reads rho from the second field of q1
    HIPCHK(*c, hipMemcpyAsync(c->q1 + N, rho, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    K1Args a = k1_base(*c, 0.0, alpha, 0, 0);
    a.op = OP_APPLY; a.q = c->q1; a.x_in = c->xA; a.y_out = c->Ap; a.u = c->tmp;
    rc = run1(*c, a);
    if (rc) return rc;
  }
  HIPCHK(*c, hipMemcpyAsync(y, c->Ap, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return end_call(*c);
}

// b and diag(A) on their own: the fused solver-init kernel computes both,
// so run it with a zero warm start and read them back.
static int init_fields(igr_ctx* c, double alpha, const double* q_host, const double* rho_host) {
  const size_t N = c->N;
  if (q_host) HIPCHK(*c, hipMemcpyAsync(c->q1, q_host, c->nv * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  else {
    launch_fill(*c, c->q1, c->nv * N, 0.0);
    HIPCHK(*c, hipMemcpyAsync(c->q1 + (c->dim == 2 ? 2 : 1) * N, rho_host, N * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
  }
  launch_fill(*c, c->tmp, N, 0.0);
  if (c->dim == 2) {
    SipArgs a = sip_base(*c, alpha);
    a.mx = c->q1; a.my = c->q1 + N; a.rho = c->q1 + 2 * N; a.x = c->tmp;
    a.invR_out = c->invR; a.invM_out = c->invM; a.b_out = c->b;
    launch2_sip(*c, M_INIT_JAC, a);
  } else {
    K1Args a = k1_base(*c, 0.0, alpha, 0, IGR_SOLVER_JACOBI);
    a.op = OP_SOLVE; a.q = c->q1; a.S = c->tmp;
    return run1(*c, a);
  }
  return IGR_OK;
}

int igr_igr_rhs(igr_ctx* c, double alpha, const double* q, double* b) {
  if (!c || !q || !b) return IGR_EINVAL;
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = init_fields(c, alpha, q, nullptr);
  if (rc) return rc;
  HIPCHK(*c, hipMemcpyAsync(b, c->b, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return end_call(*c);
}

int igr_sip_diagonal(igr_ctx* c, double alpha, const double* rho, double* diag) {
  if (!c || !rho || !diag) return IGR_EINVAL;
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = init_fields(c, alpha, nullptr, rho);
  if (rc) return rc;
  std::vector<double> im(c->N);
  HIPCHK(*c, hipMemcpyAsync(im.data(), c->invM, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  rc = end_call(*c);
  if (rc) return rc;
  for (size_t n = 0; n < c->N; ++n) diag[n] = 1.0 / im[n];
  return IGR_OK;
}

int igr_solve_elliptic(igr_ctx* c, double alpha, int n_iter, int solver) {
  if (!c || n_iter < 0 || solver < 0 || solver > 2) return c ? seterr(*c, IGR_EINVAL, "unknown solver") : IGR_EINVAL;
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = ensure_scal(*c, n_iter);
  if (rc) return rc;
  if (c->dim == 2) {
    solve2(*c, c->q, alpha, n_iter, solver, c->scal);
  } else {
    K1Args a = k1_base(*c, 0.0, alpha, n_iter, solver);
    a.op = OP_SOLVE; a.trace = c->scal;
    rc = run1(*c, a);
    if (rc) return rc;
  }
  return end_call(*c);
}

int igr_solver_trace(igr_ctx* c, int n_iter, double* a, double* beta) {
  if (!c || n_iter < 0 || (size_t)(4 * n_iter + 1) > c->scal_n) return IGR_EINVAL;
  std::vector<double> h(4 * (size_t)n_iter + 1);
  HIPCHK(*c, hipMemcpy(h.data(), c->scal, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < n_iter; ++k) {
    if (c->dim == 2) { if (a) a[k] = h[2 * n_iter + 1 + k]; if (beta) beta[k] = h[3 * n_iter + 1 + k]; }
    else { if (a) a[k] = h[k]; if (beta) beta[k] = h[n_iter + k]; }
  }
  return IGR_OK;
}

int igr_euler_rhs(igr_ctx* c, double gamma, double alpha, int n_iter, int solver, double* dq) {
  if (!c || n_iter < 0) return IGR_EINVAL;
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "unknown solver");
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = ensure_scal(*c, n_iter);
  if (rc) return rc;
  if (c->dim == 2) {
    if (alpha > 0.0) solve2(*c, c->q, alpha, n_iter, solver, c->scal);
    else launch_fill(*c, c->S, c->N, 0.0);
    HypArgs h = hyp_base(*c, gamma);
    h.q = c->q; h.mode = 0; h.out = c->dq;
    launch2_hyp(*c, h);
  } else {
    K1Args a = k1_base(*c, gamma, alpha, n_iter, solver);
    a.op = OP_RHS;
    rc = run1(*c, a);
    if (rc) return rc;
  }
  if (dq) HIPCHK(*c, hipMemcpyAsync(dq, c->dq, c->nv * c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return end_call(*c);
}

int igr_run_steps(igr_ctx* c, int n_steps, double dt, double gamma, double alpha, int n_iter, int solver) {
  if (!c || n_steps < 0 || n_iter < 0) return IGR_EINVAL;
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "unknown solver");
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = run_steps_impl(*c, n_steps, dt, gamma, alpha, n_iter, solver, 1, 0);
  if (rc) return rc;
  return end_call(*c);
}

int igr_run_forward(igr_ctx* c, double dt, double t_end, double gamma, double alpha, int n_iter, int solver,
                    int* n_steps_out) {
  if (!c || !(dt > 0.0) || !(t_end > 0.0) || n_iter < 0) return IGR_EINVAL;
  if (solver < 0 || solver > 2) return seterr(*c, IGR_EINVAL, "unknown solver");
  const int n_steps = (int)std::ceil(t_end / dt);
  dt = t_end / n_steps;
  if (n_steps_out) *n_steps_out = n_steps;
  int rc = begin_call(*c);
  if (rc) return rc;
  rc = run_steps_impl(*c, n_steps, dt, gamma, alpha, n_iter, solver, 1, 1);
  if (rc) return rc;
  return end_call(*c);
}

int igr_time_kernel(igr_ctx* c, int kernel, double gamma, double alpha, int reps, double* ms) {
  if (!c || !ms || reps < 1 || reps > 4096 || kernel < 0 || kernel > 5 || c->dim != 2) return IGR_EINVAL;
  int rc = begin_call(*c);
  if (rc) return rc;
  const size_t N = c->N;
  // events around a batch of back-to-back launches: the mean in-stream time
  // per launch, gaps included
  auto batch_ms = [&](auto&& launch) -> int {
    hipEvent_t e0, e1;
    HIPCHK(*c, hipEventCreate(&e0));
    HIPCHK(*c, hipEventCreate(&e1));
    for (int w = 0; w < 3; ++w) launch();
    HIPCHK(*c, hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) launch();
    HIPCHK(*c, hipEventRecord(e1, c->stream));
    HIPCHK(*c, hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(*c, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    return IGR_OK;
  };
  if (kernel == 2 || kernel == 5) {
    // the hyperbolic RHS + RK3 stage-2 update, launched back to back on the current state
    HypArgs h = hyp_base(*c, gamma);
    h.q = c->q; h.mode = 2; h.cdt = 1e-6; h.out = c->q2; h.qn = c->q1;
    launch_copy(*c, c->q1, c->q, c->nv * N);
    if (kernel == 5) {
      rc = batch_ms([&]() { launch2_hyp(*c, h); });
      if (rc) return rc;
      return end_call(*c);
    }
    std::vector<hipEvent_t> ev(2 * (size_t)reps);
    for (auto& e : ev) HIPCHK(*c, hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
    for (int w = 0; w < 3; ++w) launch2_hyp(*c, h);
    for (int r = 0; r < reps; ++r) {
      set_kernel_timer(ev[2 * r], ev[2 * r + 1]);
      launch2_hyp(*c, h);
    }
    HIPCHK(*c, hipEventSynchronize(ev.back()));
    double tot = 0.0;
    for (int r = 0; r < reps; ++r) {
      float t = 0.f;
      HIPCHK(*c, hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]));
      tot += t;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms = tot / reps;
    return end_call(*c);
  }
  // PCG_A / PCG_B in the context they run in: a real PCG solve of `reps`
  // iterations on the current state (into scratch vectors), every launch of
  // the timed kernel stamped with start/end events by the kernel launch.
  rc = ensure_scal(*c, reps);
  if (rc) return rc;
  launch_copy(*c, c->xA, c->S, N);
  double* R0 = c->partials;
  double* R1 = c->partials + c->pstride;
  SipArgs a = sip_base(*c, alpha);
  a.mx = c->q; a.my = c->q + N; a.rho = c->q + 2 * N; a.x = c->xA;
  a.invR_out = c->invR; a.invM_out = c->invM; a.rout = c->r; a.sc = c->scal; a.n_iter = reps; a.part_out = R0;
  launch2_sip(*c, M_INIT_PCG, a);
  double* pbuf[2] = {c->pA, c->pB};
  if (kernel == 3 || kernel == 4) {
    // PCG_A (3) or PCG_B (4) of iteration 1 launched `reps` times back to
    // back, events around the whole batch only.  PCG_A's inputs (r, 1/M,
    // p_0, 1/rho, the pAp partials) do not change, so every launch does the
    // same work; PCG_B updates x and r in place with the same fields and
    // scalars each time.  The mean is the in-stream duration plus the gap
    // between launches.
    const int nit = std::max(reps, 2);
    rc = ensure_scal(*c, nit);
    if (rc) return rc;
    for (int it = 0; it < 2; ++it) {
      SipArgs k = sip_base(*c, alpha);
      k.sc = c->scal; k.n_iter = nit; k.it = it; k.first = (it == 0);
      k.r = c->r; k.invM = c->invM; k.invR = c->invR; k.pold = pbuf[(it + 1) % 2]; k.pnew = pbuf[it % 2];
      k.y = c->Ap; k.xout = c->xA; k.part_in = R0; k.part_out = R1;
      if (it == 1 && kernel == 3) {
        rc = batch_ms([&]() { launch2_sip(*c, M_PCG_A, k); });
        if (rc) return rc;
        return end_call(*c);
      }
      launch2_sip(*c, M_PCG_A, k);
      SipArgs m = sip_base(*c, alpha);
      m.sc = c->scal; m.n_iter = nit; m.it = it;
      m.xout = c->xA; m.pold = pbuf[it % 2]; m.y = c->Ap; m.rout = c->r; m.r = nullptr; m.invM = c->invM;
      m.part_in = R1; m.part_out = R0;
      if (it == 1 && kernel == 4) {
        // it reads PCG_A's partial sums (R1) and writes R0, which it does not read
        rc = batch_ms([&]() { launch2_sip(*c, M_PCG_B, m); });
        if (rc) return rc;
        return end_call(*c);
      }
      launch2_sip(*c, it + 1 < nit ? M_PCG_B : M_PCG_BX, m);
    }
  }
  std::vector<hipEvent_t> ev(2 * (size_t)reps);
  for (auto& e : ev) HIPCHK(*c, hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
  for (int it = 0; it < reps; ++it) {
    SipArgs k = sip_base(*c, alpha);
    k.sc = c->scal; k.n_iter = reps; k.it = it; k.first = (it == 0);
    k.r = c->r; k.invM = c->invM; k.invR = c->invR; k.pold = pbuf[(it + 1) % 2]; k.pnew = pbuf[it % 2]; k.y = c->Ap;
    k.xout = c->xA; k.part_in = R0; k.part_out = R1;
    SipArgs m = sip_base(*c, alpha);
    m.sc = c->scal; m.n_iter = reps; m.it = it;
    m.xout = c->xA; m.pold = pbuf[it % 2]; m.y = c->Ap; m.rout = c->r; m.r = nullptr; m.invM = c->invM;
    m.part_in = R1; m.part_out = R0;
    if (kernel == 0) set_kernel_timer(ev[2 * it], ev[2 * it + 1]);
    launch2_sip(*c, M_PCG_A, k);
    if (kernel == 1) set_kernel_timer(ev[2 * it], ev[2 * it + 1]);
    launch2_sip(*c, it + 1 < reps ? M_PCG_B : M_PCG_BX, m);
  }
  HIPCHK(*c, hipEventSynchronize(ev.back()));
  double tot = 0.0;
  for (int it = 0; it < reps; ++it) {
    float t = 0.f;
    HIPCHK(*c, hipEventElapsedTime(&t, ev[2 * it], ev[2 * it + 1]));
    tot += t;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  *ms = tot / reps;
  return end_call(*c);
}

}  // extern "C"
