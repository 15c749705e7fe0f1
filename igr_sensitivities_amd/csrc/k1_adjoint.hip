// 1D discrete adjoint for gfx950, FP64: the reverse mode of the 1D forward
// map (IGR_adjoints_1d/: hyperbolic.jl:23-83, elliptic.jl:32-467,
// forward.jl:17-33 and the SSP-RK3 loop of enzyme_adjoint.jl:34-63), which
// the reference differentiates with Enzyme (enzyme_adjoint.jl:255-314).
//
// Like the 1D forward, one workgroup handles one instance and synchronises
// with workgroup barriers only.  Three operations per launch:
//   OPA_STORE  one SSP-RK3 step from (q^n, Sigma_0) keeping the stage
//              states, Sigma of every stage and every solver iterate
//              (the same arithmetic as the forward step, bit for bit);
//   OPA_REV    the reverse of that stored step: adjoint of q^{n+1} (and of
//              Sigma_3) -> adjoint of q^n (and of Sigma_0), with the
//              gamma- and alpha-adjoints accumulated;
//   OPA_OBJ    the objective at the current state and its gradient (the
//              seed of the reverse sweep).
#include "igr_ctx.hpp"
#include "k1_common.hpp"
#include "k1_forward.hpp"

namespace igr {

// reverse of the pointwise 1D Euler physics at one node: flux1 (euler.jl:26-35)
// with adjoint Fb, and the wave speed |u| + c (euler.jl:42-56) with adjoint lb;
// q = (mu, rho, E, Sigma).  d|u| follows signbit, as Julia's abs does.
__device__ __forceinline__ void euler_node_rev1(double g, const double q[4], const double Fb[3], double lb,
                                                double qb[3], double& sb, double& gb) {
  const double m = q[0], r = q[1], E = q[2], S = q[3];
  const double u = m / r;
  const double KE = m * m / (2.0 * r);
  const double P = (g - 1.0) * (E - KE);
  const double pb = P + S;
  double ub = 0.0, Pb = 0.0;
  // f0 = m^2/rho + pb, f1 = m, f2 = (E + pb) m / rho
  qb[0] += Fb[0] * 2.0 * u + Fb[1] + Fb[2] * (E + pb) / r;
  qb[1] += -Fb[0] * u * u - Fb[2] * (E + pb) * u / r;
  qb[2] += Fb[2] * u;
  const double pbb = Fb[0] + Fb[2] * u;
  if (lb != 0.0) {
    ub += lb * (signbit(u) ? -1.0 : 1.0);
    const double arg = g * P / r;
    const double argb = lb / (2.0 * sqrt(arg));
    gb += argb * P / r;
    Pb += argb * g / r;
    qb[1] -= argb * arg / r;
  }
  qb[0] += ub / r;
  qb[1] -= ub * u / r;
  sb += pbb;
  Pb += pbb;
  gb += Pb * (E - KE);
  qb[2] += (g - 1.0) * Pb;
  const double KEb = -(g - 1.0) * Pb;
  qb[0] += KEb * u;
  qb[1] -= KEb * KE / r;
}

template <int NP>
struct K1R {
  const K1AdjArgs& A;
  K1Args& fa;         // forward arguments in LDS; stage pointers are switched in place
  const int N, Ne, tid, nt;
  __device__ K1R(const K1AdjArgs& a_, K1Args& f) : A(a_), fa(f), N(a_.f.Ne * NP), Ne(a_.f.Ne), tid(threadIdx.x),
                                                   nt(blockDim.x) {}
  __device__ double D(int i, int k) const { return fa.B.D[i * NP + k]; }

  __device__ void add_scalar(int slot, double v) {
    __shared__ double sred[24];
    double s[1] = {v};
    block_sum<1>(s, sred);
    if (tid == 0) A.dsc[slot] += s[0];
    __syncthreads();
  }

  // point the forward helpers at stage s's stored 1/rho, 1/M, b
  __device__ void use_stage(int s) {
    __syncthreads();
    if (tid == 0) {
      fa.invR = A.invR[s]; fa.invM = A.invM[s]; fa.b = A.b[s];
      fa.stP = A.P[s]; fa.stAP = A.AP[s]; fa.stR = A.R[s]; fa.stsc = A.sc[s];
    }
    __syncthreads();
  }

  // ---------------------------------------------------------------- store
  __device__ void store_step() {
    const double dt = A.dt;
    const double cd[3] = {dt, 1.0 / 4.0 * dt, 2.0 / 3.0 * dt};
    for (int s = 0; s < 3; ++s) {
      use_stage(s);
      if (tid == 0) fa.S = A.Sg[s + 1];
      __syncthreads();
      K1<NP> k(fa);
      if (fa.alpha > 0.0) {
        for (int n = tid; n < N; n += nt) A.Sg[s + 1][n] = A.Sg[s][n];
        __syncthreads();
        k.solve(A.Q[s], 0);
      } else {
        for (int n = tid; n < N; n += nt) A.Sg[s + 1][n] = 0.0;
        __syncthreads();
      }
      k.hyp_stage(A.Q[s], s + 1, cd[s], s < 2 ? A.Q[s + 1] : A.out);
    }
  }

  // ------------------------------------------------- hyperbolic + stage
  // out = c_n q^n + c_prev Q + c_dt H(Q, Sigma); given out-bar:
  // Q-bar (=|+=) c_prev out-bar + H_Q^T (c_dt out-bar); q^n-bar += c_n out-bar;
  // Sigma-bar = sbar_in + H_Sigma^T(...); gamma-bar accumulated.
  __device__ void hyp_rev(const double* Q, const double* S, const double* qbo, double c_dt, double c_prev, double c_n,
                          double* qbs, int acc, double* qbn, const double* sbin, double* sbout);

  // ------------------------------------------------------------ SIP VJP
  // y = A w (A is symmetric), and the adjoints of <w, A(1/rho, alpha) v>:
  // irbar += d/d(1/rho), alpha-bar += d/d alpha.
  __device__ void apply_rev(const double* w, const double* v, double* y) {
    K1<NP> k(fa);
    k.apply(w, y);
    const double al = fa.alpha, invJ = fa.invJ, J = fa.J;
    const double* ir = fa.invR;
    double ab = 0.0;
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      double dw = 0.0, dv = 0.0;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) { dw += D(i, kk) * w[e * NP + kk]; dv += D(i, kk) * v[e * NP + kk]; }
      double irb = fa.B.w[i] * J * w[n] * v[n] + al * fa.B.w[i] * invJ * dv * dw;
      if (i == 0 || i == NP - 1) {
        const bool own_L = (i == NP - 1);
        const int nn = own_L ? ((e == Ne - 1 ? 0 : e + 1) * NP) : ((e == 0 ? Ne - 1 : e - 1) * NP + NP - 1);
        const double jw = own_L ? (w[n] - w[nn]) : (w[nn] - w[n]);
        const double jv = own_L ? (v[n] - v[nn]) : (v[nn] - v[n]);
        const double irL = own_L ? ir[n] : ir[nn], irR = own_L ? ir[nn] : ir[n];
        irb += -al * invJ * jv * dw / 2.0 - al * invJ * jw * dv / 2.0;
        const bool L_chosen = !(irR > irL);
        if (own_L == L_chosen) irb += al * fa.eta / fa.dx * jw * jv;
      }
      A.irbar[n] += irb;
      if (al != 0.0) ab += v[n] * (y[n] - fa.B.w[i] * J * w[n] * ir[n]) / al;
    }
    __syncthreads();
    add_scalar(4, ab);
  }

  // ---------------------------------------------------- solve reverse
  // Sigma_out = solve(Q_s, x0): given xbar (adjoint of Sigma_out), the state
  // adjoint goes into qbs (+=) and the warm-start adjoint into x0bar.
  __device__ void solve_rev(int s, const double* Qs, const double* x0, const double* xbar, double* qbs,
                            double* x0bar) {
    use_stage(s);
    K1<NP> k(fa);
    const int n_it = fa.n_iter;
    const double* iM = fa.invM;
    double *w = A.w, *y = A.y, *rb = A.rb, *pb = A.pb, *xb = A.xb;
    for (int n = tid; n < N; n += nt) { A.irbar[n] = 0.0; A.imbar[n] = 0.0; A.bbar[n] = 0.0; }
    __syncthreads();
    const double eps = 1e-30;
    if (fa.solver == 0) {
      const double* sc = A.sc[s];
      const double* P = A.P[s];
      const double* AP = A.AP[s];
      const double* R = A.R[s];
      for (int n = tid; n < N; n += nt) { pb[n] = 0.0; rb[n] = 0.0; }
      __syncthreads();
      double rzb_next = 0.0;
      for (int kk = n_it - 1; kk >= 0; --kk) {
        const double* Pk = P + (size_t)kk * N;
        const double* Qk = AP + (size_t)kk * N;
        const double* Rk1 = R + (size_t)(kk + 1) * N;
        const double rzk = sc[kk], rzk1 = sc[kk + 1], pap = sc[n_it + 1 + kk];
        const double ak = sc[2 * n_it + 1 + kk], bk = sc[3 * n_it + 1 + kk];
        const double bb = k.dot(pb, Pk);
        const double rzb1 = rzb_next + bb / (rzk + eps);
        const double carry = -bb * rzk1 / ((rzk + eps) * (rzk + eps));
        for (int n = tid; n < N; n += nt) {
          const double z1 = Rk1[n] * iM[n];
          const double zb = pb[n] + rzb1 * Rk1[n];
          rb[n] = rb[n] + rzb1 * z1 + zb * iM[n];
          A.imbar[n] += zb * Rk1[n];
        }
        __syncthreads();
        const double abk = -k.dot(rb, Qk) + k.dot(xbar, Pk);
        const double rzbk = carry + abk / (pap + eps);
        const double papb = -abk * rzk / ((pap + eps) * (pap + eps));
        for (int n = tid; n < N; n += nt) w[n] = -ak * rb[n] + papb * Pk[n];
        __syncthreads();
        apply_rev(w, Pk, y);
        for (int n = tid; n < N; n += nt) pb[n] = bk * pb[n] + ak * xbar[n] + papb * Qk[n] + y[n];
        __syncthreads();
        rzb_next = rzbk;
      }
      // r0 = b - A x0, z0 = r0/M, p0 = z0, rz0 = <r0, z0>
      for (int n = tid; n < N; n += nt) {
        const double z0 = R[n] * iM[n];
        const double zb = pb[n] + rzb_next * R[n];
        const double r0b = rb[n] + rzb_next * z0 + zb * iM[n];
        A.imbar[n] += zb * R[n];
        A.bbar[n] += r0b;
        w[n] = -r0b;
      }
      __syncthreads();
      apply_rev(w, x0, y);
      for (int n = tid; n < N; n += nt) x0bar[n] = xbar[n] + y[n];
      __syncthreads();
    } else if (fa.solver == 1) {
      // x_{k+1} = x_k + (2/3)(b - A x_k)/M
      const double om = 2.0 / 3.0;
      const double* X = A.P[s];
      const double* AX = A.AP[s];
      const double* cur = xbar;
      for (int kk = n_it - 1; kk >= 0; --kk) {
        double* nxt = (cur == xb) ? rb : xb;
        for (int n = tid; n < N; n += nt) w[n] = -(om * iM[n]) * cur[n];
        __syncthreads();
        apply_rev(w, X + (size_t)kk * N, y);
        for (int n = tid; n < N; n += nt) {
          const double xb1 = cur[n];
          A.imbar[n] += om * (fa.b[n] - AX[(size_t)kk * N + n]) * xb1;
          A.bbar[n] += om * iM[n] * xb1;
          nxt[n] = xb1 + y[n];
        }
        __syncthreads();
        cur = nxt;
      }
      for (int n = tid; n < N; n += nt) x0bar[n] = cur[n];
      __syncthreads();
    } else {
      // Chebyshev: d_1 = z/theta, d_k = c1 d_{k-1} + c2 z, x_k = x_{k-1} + d_k, r_k = b - A x_k
      const double lmin = 1.0, lmax = fa.cheb_lmax;
      const double theta = (lmax + lmin) / 2.0, delta = (lmax - lmin) / 2.0;
      const double sigma = theta / fmax(delta, 1e-10);
      double* cc = A.sc[s];      // the forward's coefficients, k = 1..n: cc[k], cc[n + 1 + k]
      if (tid == 0) {
        double rc = 1.0 / sigma;
        for (int kk = 1; kk <= n_it; ++kk) {
          double c1 = 0.0, c2 = 0.0;
          if (kk > 1) { const double rn = 1.0 / (2.0 * sigma - rc); c1 = rn * rc; c2 = 2.0 * rn / delta; rc = rn; }
          cc[kk] = c1; cc[n_it + 1 + kk] = c2;
        }
      }
      const double* X = A.P[s];
      const double* R = A.R[s];
      for (int n = tid; n < N; n += nt) { rb[n] = 0.0; pb[n] = 0.0; xb[n] = xbar[n]; }
      __syncthreads();
      for (int kk = n_it; kk >= 1; --kk) {
        const double c1 = cc[kk], c2 = cc[n_it + 1 + kk];
        for (int n = tid; n < N; n += nt) w[n] = -rb[n];
        __syncthreads();
        apply_rev(w, X + (size_t)(kk - 1) * N, y);
        for (int n = tid; n < N; n += nt) {
          A.bbar[n] += rb[n];
          const double Xb = xb[n] + y[n];
          const double Db = pb[n] + Xb;
          const double zb = (kk > 1) ? c2 * Db : Db / theta;
          pb[n] = (kk > 1) ? c1 * Db : 0.0;
          rb[n] = zb * iM[n];
          A.imbar[n] += zb * R[(size_t)(kk - 1) * N + n];
          xb[n] = Xb;
        }
        __syncthreads();
      }
      for (int n = tid; n < N; n += nt) { A.bbar[n] += rb[n]; w[n] = -rb[n]; }
      __syncthreads();
      apply_rev(w, x0, y);
      for (int n = tid; n < N; n += nt) x0bar[n] = xb[n] + y[n];
      __syncthreads();
    }
    prepare_rev(Qs, qbs);
  }

  // reverse of 1/rho, the IGR source b = w J 2 alpha u_x^2 and diag(A)
  // (compute_igr_rhs!, elliptic.jl:126-154; compute_sip_diagonal, :206-253)
  __device__ void prepare_rev(const double* Qs, double* qbs) {
    const double al = fa.alpha, invJ = fa.invJ, J = fa.J;
    const double* ir = fa.invR;
    const double* iM = fa.invM;
    double* Mb = A.y;      // adjoint of diag(A)
    double* dub = A.w;     // adjoint of u_x
    const double* m = Qs;
    const double* rho = Qs + N;
    double ab = 0.0;
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      Mb[n] = -A.imbar[n] * iM[n] * iM[n];
      double du = 0.0;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) du += D(i, kk) * (m[e * NP + kk] / rho[e * NP + kk]);
      du *= invJ;
      const double bbn = A.bbar[n];
      dub[n] = bbn * fa.B.w[i] * J * 4.0 * al * du;
      if (al != 0.0) ab += bbn * fa.B.w[i] * J * 2.0 * (du * du) + Mb[n] * (1.0 / iM[n] - fa.B.w[i] * J * ir[n]) / al;
    }
    __syncthreads();
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      double ub = 0.0, irb = A.irbar[n] + fa.B.w[i] * J * Mb[n];
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) {
        ub += invJ * D(kk, i) * dub[e * NP + kk];
        irb += al * fa.B.w[i] * invJ * (D(i, kk) * D(i, kk)) * Mb[e * NP + kk];
      }
      if (i == NP - 1 || i == 0) {
        const bool own_L = (i == NP - 1);
        const int nn = own_L ? ((e == Ne - 1 ? 0 : e + 1) * NP) : ((e == 0 ? Ne - 1 : e - 1) * NP + NP - 1);
        irb += (own_L ? -1.0 : 1.0) * al * invJ * (own_L ? D(NP - 1, NP - 1) : D(0, 0)) * Mb[n];
        const double irL = own_L ? ir[n] : ir[nn], irR = own_L ? ir[nn] : ir[n];
        const bool L_chosen = !(irR > irL);
        if (own_L == L_chosen) irb += al * (Mb[n] + Mb[nn]) * fa.eta / fa.dx;
      }
      const double u = m[n] / rho[n];
      qbs[n] += ub / rho[n];
      qbs[N + n] += -ir[n] * ir[n] * irb - ub * u / rho[n];
    }
    __syncthreads();
    add_scalar(4, ab);
  }

  // --------------------------------------------------------- step reverse
  __device__ void rev_step() {
    const double dt = A.dt;
    const bool solve = fa.alpha > 0.0;
    for (int n = tid; n < 3 * N; n += nt) A.QBn[n] = 0.0;
    __syncthreads();
    double* qbo[3] = {A.QBa, A.QBb, A.QBnext};
    double* qbs[3] = {A.QBn, A.QBa, A.QBb};
    const double cdt[3] = {dt, 1.0 / 4.0 * dt, 2.0 / 3.0 * dt};
    const double cpr[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    const double cn[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
    bool sbar_valid = A.has_sbar && solve;
    for (int s = 2; s >= 0; --s) {
      use_stage(s);
      hyp_rev(A.Q[s], A.Sg[s + 1], qbo[s], cdt[s], cpr[s], cn[s], qbs[s], s == 0, s > 0 ? A.QBn : nullptr,
              sbar_valid ? A.Scarry : nullptr, A.Xbar);
      if (solve) {
        solve_rev(s, A.Q[s], A.Sg[s], A.Xbar, qbs[s], A.Scarry);
        sbar_valid = true;
      }
    }
    if (!solve) {
      for (int n = tid; n < N; n += nt) A.Scarry[n] = 0.0;
      __syncthreads();
    }
  }

  // ------------------------------------------------------------ objective
  // IGR_adjoints_1d/enzyme_adjoint.jl:65-71 (l2 density), :110-117 (l2
  // pressure); kinetic energy / mass and the sin-weighted momentum.
  __device__ void objective(const double* q) {
    const double g = fa.gamma, J1 = fa.J;
    const double* m = q;
    const double* r = q + N;
    const double* E = q + 2 * N;
    const double Ld = Ne * fa.dx;
    __shared__ double sred[24];
    double acc[2] = {0.0, 0.0};
    auto xnode = [&](int i, int e) { return e * fa.dx + fa.J * (1.0 + fa.B.xi[i]); };
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      const double wJ = fa.B.w[i] * J1;
      if (A.obj == 0) acc[0] += wJ * (r[n] * r[n]);
      else if (A.obj == 1) { const double P = (g - 1.0) * (E[n] - m[n] * m[n] / (2.0 * r[n])); acc[0] += wJ * (P * P); }
      else if (A.obj == 2) { acc[0] += wJ * (m[n] * m[n]) / (2.0 * r[n]); acc[1] += wJ * r[n]; }
      else acc[0] += wJ * m[n] * sin(2.0 * M_PI * xnode(i, e) / Ld);
    }
    block_sum<2>(acc, sred);
    if (tid == 0) { sred[20] = acc[0]; sred[21] = acc[1]; }
    __syncthreads();
    const double S0 = sred[20], S1 = sred[21];
    __syncthreads();
    const double Jv = (A.obj == 2) ? S0 / S1 : S0;
    if (tid == 0) A.dsc[0] = Jv;
    double gb = 0.0;
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      const double wJ = fa.B.w[i] * J1;
      double mb = 0.0, rb = 0.0, Eb = 0.0;
      if (A.obj == 0) rb = 2.0 * wJ * r[n];
      else if (A.obj == 1) {
        const double KE = m[n] * m[n] / (2.0 * r[n]);
        const double P = (g - 1.0) * (E[n] - KE);
        const double Pb = 2.0 * wJ * P;
        gb += Pb * (E[n] - KE);
        Eb = (g - 1.0) * Pb;
        mb = -(g - 1.0) * Pb * m[n] / r[n];
        rb = (g - 1.0) * Pb * KE / r[n];
      } else if (A.obj == 2) {
        const double Ab = 1.0 / S1, Mb = -S0 / (S1 * S1);
        mb = Ab * wJ * m[n] / r[n];
        rb = -Ab * wJ * (m[n] * m[n]) / (2.0 * r[n] * r[n]) + Mb * wJ;
      } else {
        mb = wJ * sin(2.0 * M_PI * xnode(i, e) / Ld);
      }
      A.QBnext[n] = mb; A.QBnext[N + n] = rb; A.QBnext[2 * N + n] = Eb;
    }
    __syncthreads();
    add_scalar(3, gb);
  }
};

// Built without FMA contraction, like the forward it reverses.
#pragma clang fp contract(off)
template <int NP>
__device__ void K1R<NP>::hyp_rev(const double* Q, const double* S, const double* qbo, double c_dt, double c_prev,
                              double c_n, double* qbs, int acc, double* qbn, const double* sbin, double* sbout) {
  const double g = fa.gamma, invJ = fa.invJ;
  double* Ab = A.Ab;
  for (int n = tid; n < N; n += nt)
#pragma unroll
    for (int v = 0; v < 3; ++v) Ab[v * N + n] = c_dt * qbo[v * N + n];
  __syncthreads();
  const double wl = 1.0 / fa.B.w[NP - 1], w0 = 1.0 / fa.B.w[0];
  double gb = 0.0;
  auto wave = [&](const double* qq) {
    const double Pn = (g - 1.0) * (qq[2] - qq[0] * qq[0] / (2.0 * qq[1]));
    return fabs(qq[0] / qq[1]) + sqrt(g * Pn / qq[1]);
  };
  for (int n = tid; n < N; n += nt) {
    const int i = n % NP, e = n / NP;
    const double qo[4] = {Q[n], Q[N + n], Q[2 * N + n], S[n]};
    double Fb[3], qb[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      double sx = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) sx += D(k, i) * Ab[v * N + e * NP + k];
      Fb[v] = -invJ * sx;
    }
    double lb = 0.0;
    if (i == NP - 1 || i == 0) {
      const bool own_L = (i == NP - 1);
      const int nn = own_L ? ((e == Ne - 1 ? 0 : e + 1) * NP) : ((e == 0 ? Ne - 1 : e - 1) * NP + NP - 1);
      const double qn[3] = {Q[nn], Q[N + nn], Q[2 * N + nn]};
      const double lo = wave(qo), ln = wave(qn);
      const double lamL = own_L ? lo : ln, lamR = own_L ? ln : lo;
      const double lam = fmax_first(lamL, lamR);
      double lsum = 0.0;
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        const double aL = own_L ? Ab[v * N + n] : Ab[v * N + nn];
        const double aR = own_L ? Ab[v * N + nn] : Ab[v * N + n];
        const double fb = -invJ * wl * aL + invJ * w0 * aR;
        const double qL = own_L ? qo[v] : qn[v], qR = own_L ? qn[v] : qo[v];
        lsum += fb * (qL - qR) / 2.0;
        if (own_L) { Fb[v] += invJ * wl * Ab[v * N + n] + fb / 2.0; qb[v] += lam * fb / 2.0; }
        else { Fb[v] += -invJ * w0 * Ab[v * N + n] + fb / 2.0; qb[v] -= lam * fb / 2.0; }
      }
      const bool L_chosen = !(lamR > lamL);
      if (own_L == L_chosen) lb = lsum;
    }
    double sb = 0.0;
    euler_node_rev1(g, qo, Fb, lb, qb, sb, gb);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double val = c_prev * qbo[v * N + n] + qb[v];
      if (acc) qbs[v * N + n] += val;
      else qbs[v * N + n] = val;
      if (qbn) qbn[v * N + n] += c_n * qbo[v * N + n];
    }
    sbout[n] = (sbin ? sbin[n] : 0.0) + sb;
  }
  __syncthreads();
  add_scalar(3, gb);
}
#pragma clang fp contract(fast)

template <int NP>
__global__ __launch_bounds__(1024) void k1_adjoint(K1AdjArgs a) {
  __shared__ __align__(16) unsigned char fa_raw[sizeof(K1Args)];
  K1Args& fa = *reinterpret_cast<K1Args*>(fa_raw);
  if (threadIdx.x == 0) __builtin_memcpy(fa_raw, &a.f, sizeof(K1Args));
  __syncthreads();
  K1R<NP> k(a, fa);
  if (a.op == OPA_STORE) k.store_step();
  else if (a.op == OPA_REV) k.rev_step();
  else k.objective(a.f.q);
}

void launch1_adjoint(const Ctx& c, const K1AdjArgs& a, int block) {
  switch (c.np) {
    case 2: hipLaunchKernelGGL(k1_adjoint<2>, dim3(1), dim3(block), 0, c.stream, a); break;
    case 3: hipLaunchKernelGGL(k1_adjoint<3>, dim3(1), dim3(block), 0, c.stream, a); break;
    case 4: hipLaunchKernelGGL(k1_adjoint<4>, dim3(1), dim3(block), 0, c.stream, a); break;
    case 5: hipLaunchKernelGGL(k1_adjoint<5>, dim3(1), dim3(block), 0, c.stream, a); break;
  }
}

}  // namespace igr
