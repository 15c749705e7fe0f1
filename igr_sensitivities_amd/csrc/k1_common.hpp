// Shared 1D definitions: kernel argument block and pointwise Euler physics.
#pragma once
#include "igr_device.hpp"

namespace igr {

constexpr int K1_MAXTAN = 4;

enum K1Op { OP_RUN = 0, OP_SOLVE = 1, OP_RHS = 2, OP_HYP = 3, OP_APPLY = 4, OP_STORE = 5 };

// One 1D instance (one workgroup).  All pointers are device memory.
struct K1Args {
  int op = OP_RUN;
  int lds = 0;             // OP_RUN / OP_STORE: state and scratch in dynamic LDS (k1_lds_doubles of them)
  int Ne = 1;
  double J = 0.5, invJ = 2.0, dx = 1.0, eta = 16.0;
  BasisK B;
  double gamma = 1.4, alpha = 0.0;
  int n_iter = 0, solver = 0;
  double cheb_lmax = 2.1;
  int n_steps = 0;
  double dt = 0.0;
  int n_steps2 = 0;        // OP_RUN: a second run of n_steps2 steps of dt2 after the first
  double dt2 = 0.0;
  int zero_sigma = 1, final_solve = 0;
  double* q = nullptr;     // state (3N): mu, rho, E
  double* q1 = nullptr;
  double* q2 = nullptr;
  double* S = nullptr;     // Sigma (N), warm start in/out
  double* dq = nullptr;    // RHS output (3N)
  double* invR = nullptr;
  double* invM = nullptr;
  double* b = nullptr;
  double* r = nullptr;
  double* p = nullptr;
  double* z = nullptr;
  double* Ap = nullptr;
  double* u = nullptr;
  double* g = nullptr;
  double* F = nullptr;     // 3N
  double* fstar = nullptr; // 3Ne
  double* fjump = nullptr; // Ne
  double* favg = nullptr;
  double* fsig = nullptr;
  double* trace = nullptr; // 2*n_iter PCG scalars (OP_SOLVE)
  // forward tangents (OP_RUN with n_tan > 0): state, stage buffers, varsigma
  int n_tan = 0, form = 2;
  double sgam = 0.0;
  double* ts[K1_MAXTAN] = {nullptr, nullptr, nullptr, nullptr};
  double* ts1[K1_MAXTAN] = {nullptr, nullptr, nullptr, nullptr};
  double* ts2[K1_MAXTAN] = {nullptr, nullptr, nullptr, nullptr};
  double* tvs[K1_MAXTAN] = {nullptr, nullptr, nullptr, nullptr};
  double* bs = nullptr;    // tangent elliptic right-hand side (N)
  // optional copies of every solver iterate (the adjoint's recompute):
  // PCG p_k, A p_k, r_k and (rz, pAp, a, beta); Jacobi x_k, A x_k;
  // Chebyshev x_k (at k-1), r_k
  double* stP = nullptr;
  double* stAP = nullptr;
  double* stR = nullptr;
  double* stsc = nullptr;
  double* snaps = nullptr;        // OP_STORE: records (mu, rho, E, Sigma), 4N doubles each
  const double* x_in = nullptr;   // OP_APPLY
  double* y_out = nullptr;
  int* flag = nullptr;
};

// 1D discrete adjoint (k1_adjoint.hip): one stored SSP-RK3 step and the
// adjoint work of one instance.
enum K1AdjOp { OPA_STORE = 0, OPA_REV = 1, OPA_OBJ = 2 };
struct K1AdjArgs {
  K1Args f;                    // forward arguments; f.q = q^n (Q[0])
  int op = OPA_STORE;
  double dt = 0.0;
  double* Q[3] = {nullptr, nullptr, nullptr};            // q^n, q1, q2 (3N each)
  double* Sg[4] = {nullptr, nullptr, nullptr, nullptr};  // Sigma_0 .. Sigma_3
  double* invR[3] = {nullptr, nullptr, nullptr};         // per stage: 1/rho, 1/M, b
  double* invM[3] = {nullptr, nullptr, nullptr};
  double* b[3] = {nullptr, nullptr, nullptr};
  double* P[3] = {nullptr, nullptr, nullptr};            // per stage: solver iterates
  double* AP[3] = {nullptr, nullptr, nullptr};
  double* R[3] = {nullptr, nullptr, nullptr};
  double* sc[3] = {nullptr, nullptr, nullptr};
  double* out = nullptr;       // q^{n+1} (OPA_STORE)
  double* QBnext = nullptr;    // adjoint of q^{n+1} (3N; the objective gradient for OPA_OBJ)
  double* QBn = nullptr;       // adjoint of q^n (3N)
  double* QBa = nullptr;
  double* QBb = nullptr;
  double* Xbar = nullptr;      // N each
  double* Scarry = nullptr;
  double* w = nullptr;
  double* y = nullptr;
  double* rb = nullptr;
  double* pb = nullptr;
  double* xb = nullptr;
  double* Ab = nullptr;        // 3N
  double* irbar = nullptr;
  double* imbar = nullptr;
  double* bbar = nullptr;
  double* dsc = nullptr;       // [0] J, [3] gamma-bar, [4] alpha-bar
  int has_sbar = 0;
  int obj = 0;
};

// Doubles of dynamic LDS an LDS-resident single-instance run needs: the state
// and both stage buffers (9N), Sigma and ten scratch vectors (11N), the
// fluxes (3N) and six face values per element.
__host__ __device__ inline size_t k1_lds_doubles(int Ne, int np) {
  const size_t N = (size_t)Ne * np;
  return 23 * N + 6 * (size_t)Ne;
}

#pragma clang fp contract(off)
// polytropic_flux, IGR_adjoints_1d/euler.jl:26-35
__device__ __forceinline__ void flux1(double g, double m, double r, double E, double S, double f[3]) {
  const double P = (g - 1.0) * (E - m * m / (2.0 * r));
  const double pb = P + S;
  f[0] = m * m / r + pb; f[1] = m; f[2] = (E + pb) * m / r;
}
// max_wavespeed, euler.jl:52-56 (sound_speed :42-45; negative argument = DomainError)
__device__ __forceinline__ double wave1(double g, const double q[4], int* flag) {
  const double u = q[0] / q[1];
  const double P = (g - 1.0) * (q[2] - q[0] * q[0] / (2.0 * q[1]));
  const double arg = g * P / q[1];
  if (!(arg >= 0.0)) flag_domain(flag);
  return fabs(u) + sqrt(arg);
}
// sensitivity_flux, IGR_adjoints_1d/sensitivity.jl:36-50; q = (mu, rho, E, Sigma),
// s = (s_mu, s_rho, s_E, s_Sigma)
__device__ __forceinline__ void sflux1(double g, const double q[4], const double s[4], double sgam, double f[3]) {
  const double m = q[0], r = q[1], E = q[2], S = q[3];
  const double sm = s[0], sr = s[1], sE = s[2], sS = s[3];
  const double u = m / r;
  const double su = (sm - u * sr) / r;
  const double P = (g - 1.0) * (E - m * m / (2.0 * r));
  const double e = E / r - u * u / 2.0;
  const double sP = (g - 1.0) * (sE - sr * (u * u) / 2.0 - r * u * su) + sgam * r * e;
  f[0] = sm * u + m * su + sP + sS;
  f[1] = sm;
  f[2] = su * (E + P + S) + u * (sE + sP + sS);
}
#pragma clang fp contract(fast)

}  // namespace igr
