// Argument blocks and pointwise derivative helpers of the 2D discrete
// adjoint (reverse mode of the forward map, replacing Enzyme).
#pragma once
#include "k2_common.hpp"

namespace igr {

// Hyperbolic RHS + SSP-RK3 stage, reversed.  With the stage output
// out = c_n q^n + c_prev Q + c_dt H(Q, Sigma), given qbar_out:
//   qbar_stage  (=|+=)  c_prev qbar_out + H_Q^T (c_dt qbar_out)
//   qbar_n      +=      c_n qbar_out             (if non-null)
//   sbar_out     =      sbar_in + H_Sigma^T (c_dt qbar_out)
//   gamma-bar partial   H_gamma^T (c_dt qbar_out)  (per-workgroup accumulator)
struct HypRevArgs {
  Geo2 g;
  BasisK B;
  double gamma;
  const double* q;         // stage input Q (4 fields)
  const double* S;         // Sigma_s
  const double* qbar_out;  // adjoint of the stage output (4 fields)
  double* qbar_stage;      // adjoint of Q (4 fields)
  int accumulate;          // 1: qbar_stage +=, 0: qbar_stage =
  double* qbar_n;          // adjoint of q^n (may be null)
  const double* sbar_in;   // adjoint of Sigma_s carried from later (may be null)
  double* sbar_out;        // total adjoint of Sigma_s
  double c_dt, c_prev, c_n;
  double* gbar_blk;        // per-workgroup gamma-bar accumulators
};

enum RevMode { RV_VJP = 0, RV_PCG_START = 1, RV_PCG_ITER = 2, RV_PCG_FIN = 3, RV_JAC_ITER = 4,
               RV_JAC_FIN = 5, RV_CHEB_ITER = 6, RV_CHEB_FIN = 7 };

// Reverse of one fixed-iteration solve (and the standalone SIP VJP).
struct RevArgs {
  Geo2 g;
  BasisK B;
  double alpha;
  int k = 0, n_iter = 0;
  // forward quantities of the solve being reversed
  const double* invR = nullptr;   // 1/rho
  const double* invM = nullptr;   // 1/diag(A)
  const double* P = nullptr;      // PCG directions p_k, k < n       (stride N)
  const double* AP = nullptr;     // A p_k                            (stride N)
  const double* R = nullptr;      // residuals r_k, k <= n            (stride N)
  const double* X = nullptr;      // Jacobi x_k (k < n) / Chebyshev x_k at k-1 (k >= 1)   (stride N)
  const double* AX = nullptr;     // Jacobi A x_k                     (stride N)
  const double* b = nullptr;      // IGR source b (Jacobi/Chebyshev)
  const double* sc = nullptr;     // forward scalars rz | pAp | a | beta
  const double* x0 = nullptr;     // warm start
  const double* q = nullptr;      // stage state (4 fields): mx, my, rho, E
  // adjoint state
  const double* xbar = nullptr;   // adjoint of the solve output (Jacobi/Chebyshev: of x_{k+1} / x_k)
  double* xbar_out = nullptr;     // Jacobi/Chebyshev: adjoint of x_k / x_{k-1}
  const double* rbar_in = nullptr;
  double* rbar_out = nullptr;
  const double* pbar_in = nullptr;
  double* pbar_out = nullptr;
  double* irbar = nullptr;        // accumulated adjoint of 1/rho
  double* imbar = nullptr;        // accumulated adjoint of 1/M
  double* bbar = nullptr;         // accumulated adjoint of b (Jacobi/Chebyshev)
  double* rs = nullptr;           // reverse scalars
  double* x0bar = nullptr;        // adjoint of the warm start (output)
  double* qbar = nullptr;         // adjoint of the stage state (4 fields, +=)
  double* abar_blk = nullptr;     // per-workgroup alpha-bar accumulators
  double omega = 0.0;             // Jacobi
  double theta = 1.0, cc1 = 0.0, cc2 = 0.0;   // Chebyshev: d_k = cc1 d_{k-1} + cc2 z_k (k > 1), z_1/theta
  int first = 0;
  // standalone VJP
  const double* v = nullptr;      // vector the operator was applied to
  const double* wbar = nullptr;   // adjoint of A v
  double* y = nullptr;            // A^T wbar
  double* partials = nullptr;
  unsigned* counter = nullptr;
  const double* part_in = nullptr;
  double* part_out = nullptr;
  int ngroups = 0;   // workgroups of the reverse PCG start / iteration kernels (set by the launcher)
  int np_in = 0;     // partial sums the previous reverse PCG kernel wrote (set by the launcher)
};

// reverse-scalar layout per solve: for iteration k
//   rs[5k+0] beta-bar_k, rs[5k+1] a-bar_k, rs[5k+2] pAp-bar_k,
//   rs[5k+3] rz-bar total of rz_{k+1}, rs[5k+4] the part of rz-bar of rz_k
//   contributed by iteration k (a_k and beta_k); for k = 0 it is the total.
struct RS {
  static __host__ __device__ int bbar(int k) { return 5 * k + 0; }
  static __host__ __device__ int abar(int k) { return 5 * k + 1; }
  static __host__ __device__ int papbar(int k) { return 5 * k + 2; }
  static __host__ __device__ int rzbar_next(int k) { return 5 * k + 3; }
  static __host__ __device__ int carry(int k) { return 5 * k + 4; }
};

// forward scalar layout (as written by the forward PCG kernels)
struct FS {
  int n;
  __host__ __device__ FS(int n_) : n(n_) {}
  __host__ __device__ int rz(int k) const { return k; }
  __host__ __device__ int pap(int k) const { return n + 1 + k; }
  __host__ __device__ int a(int k) const { return 2 * n + 1 + k; }
  __host__ __device__ int beta(int k) const { return 3 * n + 1 + k; }
};

// Objective value / gradient kernels.
struct ObjArgs {
  Geo2 g;
  BasisK B;
  int obj;
  double prm[4];
  double Lx, Ly;
  double gamma;
  const double* q;
  double* out;        // [0] = J, [1] = KE, [2] = mass  (value kernel)
  double* qbar;       // gradient (4 fields), written
  double* gbar;       // d J / d gamma (gradient kernel, added by the last workgroup)
  double* partials;
  unsigned* counter;
};

// ---- reverse of the pointwise Euler physics at one node -----------------
// Accumulates qb[0..3] (mx, my, rho, E), sb (Sigma) and gb (gamma) from the
// adjoints of the x/y fluxes Fxb, Fyb and of the x/y wave speeds lxb, lyb.
__device__ __forceinline__ void euler_node_rev(double gam, const double q[5], const double Fxb[4],
                                               const double Fyb[4], double lxb, double lyb, double qb[4],
                                               double& sb, double& gb) {
  const double mx = q[0], my = q[1], r = q[2], E = q[3], S = q[4];
  const double ir = 1.0 / r;
  const double KE = (mx * mx + my * my) / (2.0 * r);
  const double P = (gam - 1.0) * (E - KE);
  const double pb = P + S;
  const double ux = mx * ir, uy = my * ir;
  double uxb = 0.0, uyb = 0.0, pbb = 0.0, Pb = 0.0;
  // x flux
  qb[0] += Fxb[0] * ux + Fxb[2];
  uxb += Fxb[0] * mx + Fxb[1] * my + Fxb[3] * (E + pb);
  pbb += Fxb[0] + Fxb[3] * ux;
  qb[1] += Fxb[1] * ux;
  qb[3] += Fxb[3] * ux;
  // y flux
  qb[0] += Fyb[0] * uy;
  uyb += Fyb[0] * mx + Fyb[1] * my + Fyb[3] * (E + pb);
  qb[1] += Fyb[1] * uy + Fyb[2];
  pbb += Fyb[1] + Fyb[3] * uy;
  qb[3] += Fyb[3] * uy;
  // wave speeds |u| + c, c = sqrt(gam P / rho); d|u| follows signbit
  if (lxb != 0.0 || lyb != 0.0) {
    uxb += lxb * (signbit(ux) ? -1.0 : 1.0);
    uyb += lyb * (signbit(uy) ? -1.0 : 1.0);
    const double arg = gam * P * ir;
    const double c = sqrt(arg);
    const double argb = (lxb + lyb) / (2.0 * c);
    gb += argb * P * ir;
    Pb += argb * gam * ir;
    qb[2] -= argb * arg * ir;
  }
  // u = m / rho
  qb[0] += uxb * ir;
  qb[1] += uyb * ir;
  qb[2] -= (uxb * ux + uyb * uy) * ir;
  // pb = P + Sigma
  sb += pbb;
  Pb += pbb;
  // P = (gam - 1)(E - KE)
  gb += Pb * (E - KE);
  qb[3] += (gam - 1.0) * Pb;
  const double KEb = -(gam - 1.0) * Pb;
  qb[0] += KEb * mx * ir;
  qb[1] += KEb * my * ir;
  qb[2] -= KEb * KE * ir;
}

}  // namespace igr
