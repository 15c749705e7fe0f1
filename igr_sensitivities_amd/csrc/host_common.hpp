// Host-side helpers for the C ABI: status handling, call prologue/epilogue
// (domain flag), and the SSP-RK3 driver with hipGraph capture of a step.
#pragma once
#include <cmath>
#include <string>
#include <vector>
#include "igr_ctx.hpp"
#include "k1_common.hpp"
#include "k2_common.hpp"

namespace igr {

inline int seterr(Ctx& c, int code, const std::string& msg) {
  c.err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                          \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ::igr::seterr((ctx), IGR_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline int begin_call(Ctx& c) {
  c.err.clear();
  HIPCHK(c, hipMemsetAsync(c.flag, 0, sizeof(int), c.stream));
  return IGR_OK;
}

inline int end_call(Ctx& c) {
  int f = 0;
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(&f, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(c, hipStreamSynchronize(c.stream));
  if (f) return seterr(c, IGR_EDOMAIN, "negative pressure under sqrt or non-positive density (DomainError)");
  return IGR_OK;
}

inline void destroy_graphs(Ctx& c) {
  for (auto& kv : c.graphs) {
    (void)hipGraphExecDestroy(kv.second.second);
    (void)hipGraphDestroy(kv.second.first);
  }
  c.graphs.clear();
}

// Tangent states of a context (tangent.hip): state, two stage buffers and
// varsigma per tangent direction.
struct TanBufs {
  int nt = 0;
  std::vector<double*> s, s1, s2, vs;
};
int tan_bufs(Ctx& c, int nt, TanBufs*& T);
void sens2_begin(Ctx& c, TanBufs& T, int nt);
void sens2_step(Ctx& c, TanBufs& T, int nt, double dt, double gamma, double alpha, int n_iter, int solver,
                double sgam);
void sens2_end(Ctx& c, TanBufs& T, int nt, double gamma, double alpha, int n_iter, int solver);
int sens2_steps(double dt, double t_end, double& dts);
int sens1_args(Ctx& c, TanBufs& T, int n_tan, double dt, double t_end, double gamma, double alpha, int n_iter,
               int solver, int form, double sgam, K1Args& a, int& n_steps);

// defined in capi.hip
void step2(Ctx& c, double dt, double gamma, double alpha, int n_iter, int solver);
void solve2(Ctx& c, const double* qs, double alpha, int n_iter, int solver, double* sc, double* x = nullptr,
            const double* b_in = nullptr);
void solve2_fields(Ctx& c, const double* mx, const double* my, const double* rho, double alpha, int n_iter,
                   int solver, double* sc, double* x, const double* b_in);
void make_basis(int p, double* xi, double* w, double* Dm);
void default_tiles(Ctx& c);
int grid2_blocks(const Ctx& c);
int grid2_bblocks(const Ctx& c);
void launch_fill(const Ctx& c, double* x, size_t n, double v);
K1Args k1_base(Ctx& c, double gamma, double alpha, int n_iter, int solver);
int run1(Ctx& c, const K1Args& a);
int k1_block(const Ctx& c);
int ensure_scal(Ctx& c, int n_iter);

// n_steps SSP-RK3 steps on the context state.  zero_sigma: Sigma <- 0 first
// (IGR_adjoints_2d/forward.jl:42); final_solve: Sigma solve at the final
// state (:71-73).
inline int run_steps_impl(Ctx& c, int n_steps, double dt, double gamma, double alpha, int n_iter, int solver,
                          int zero_sigma, int final_solve) {
  int rc = ensure_scal(c, n_iter);
  if (rc) return rc;
  if (c.dim == 1) {
    K1Args a = k1_base(c, gamma, alpha, n_iter, solver);
    a.op = OP_RUN; a.n_steps = n_steps; a.dt = dt; a.zero_sigma = zero_sigma; a.final_solve = final_solve;
    return run1(c, a);
  }
  if (zero_sigma) launch_fill(c, c.S, c.N, 0.0);
  if (c.use_graphs && n_steps >= 2) {
    Ctx::GraphKey key{dt, gamma, alpha, n_iter, solver};
    auto it = c.graphs.find(key);
    if (it == c.graphs.end()) {
      hipGraph_t g;
      hipGraphExec_t ge;
      HIPCHK(c, hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      step2(c, dt, gamma, alpha, n_iter, solver);
      HIPCHK(c, hipStreamEndCapture(c.stream, &g));
      HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      it = c.graphs.emplace(key, std::make_pair(g, ge)).first;
    }
    for (int s = 0; s < n_steps; ++s) HIPCHK(c, hipGraphLaunch(it->second.second, c.stream));
  } else {
    for (int s = 0; s < n_steps; ++s) step2(c, dt, gamma, alpha, n_iter, solver);
  }
  if (final_solve && alpha > 0.0) solve2(c, c.q, alpha, n_iter, solver, c.scal);
  HIPCHK(c, hipGetLastError());
  return IGR_OK;
}

}  // namespace igr
