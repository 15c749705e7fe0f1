// The 1D whole-instance solver building blocks (struct K1): SIP apply, dot
// products, IGR source and diagonal, the fixed-iteration solvers, the tangent
// right-hand sides and the hyperbolic RHS fused with the SSP-RK3 stage.
// One workgroup runs them for one instance, synchronising with workgroup
// barriers; used by the forward (k1_forward.hip) and the adjoint
// (k1_adjoint.hip) kernels.
#pragma once
#include "igr_ctx.hpp"
#include "k1_common.hpp"

namespace igr {

// Work arrays of one instance: the argument block's (global memory) or
// carved from dynamic LDS.  Held by value, so the array addresses stay in
// registers instead of being reloaded from the argument block after stores.
struct K1Bufs {
  double *q, *q1, *q2, *S, *invR, *invM, *b, *r, *p, *z, *Ap, *u, *g, *F, *fstar, *fjump, *favg, *fsig, *bs;
};

// LDS: the work arrays live in `lds` (k1_lds_doubles(Ne, NP) doubles of
// dynamic LDS) instead of the argument block's global arrays.  A template
// parameter, so that every access through them compiles to LDS instructions
// rather than generic (flat) ones.
template <int NP, bool LDS = false>
struct K1 {
  const K1Args& a;
  const int N, Ne, tid, nt;
  K1Bufs w;
  __device__ __forceinline__ K1(const K1Args& args, double* lds = nullptr)
      : a(args), N(args.Ne * NP), Ne(args.Ne), tid(threadIdx.x), nt(blockDim.x) {
    if constexpr (!LDS) {
      (void)lds;
      w = K1Bufs{args.q, args.q1, args.q2, args.S, args.invR, args.invM, args.b, args.r, args.p, args.z, args.Ap,
                 args.u, args.g, args.F, args.fstar, args.fjump, args.favg, args.fsig, args.bs};
      return;
    }
    const size_t n = N, e = Ne;
    w.q = lds; w.q1 = lds + 3 * n; w.q2 = lds + 6 * n; w.S = lds + 9 * n;
    w.invR = lds + 10 * n; w.invM = lds + 11 * n; w.b = lds + 12 * n; w.r = lds + 13 * n; w.p = lds + 14 * n;
    w.z = lds + 15 * n; w.Ap = lds + 16 * n; w.u = lds + 17 * n; w.g = lds + 18 * n; w.bs = lds + 19 * n;
    w.F = lds + 20 * n; w.fstar = lds + 23 * n; w.fjump = lds + 23 * n + 3 * e; w.favg = lds + 23 * n + 4 * e;
    w.fsig = lds + 23 * n + 5 * e;
  }

  __device__ __forceinline__ double D(int i, int k) const { return a.B.D[i * NP + k]; }

  // y = A x (apply_sip!, elliptic.jl:32-116) in gather form.  x must be
  // complete (barrier) before the call; y is complete when it returns.
  __device__ __forceinline__ void apply(const double* x, double* y) {
    apply_with(x, [&](int n, double v) { y[n] = v; });
  }
  // The same apply with epi(n, (A x)_n) run by the thread that owns node n
  // in place of the store.  The second phase reads x at the own node only, so
  // epi may overwrite x[n].  END_SYNC = false leaves out the closing barrier
  // for a caller whose next step is a barrier anyway.
  template <bool END_SYNC = true, class Epi>
  __device__ __forceinline__ void apply_with(const double* x, Epi epi) {
    const double al = a.alpha, invJ = a.invJ;
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      double dxi = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) dxi += D(i, k) * x[e * NP + k];
      w.g[n] = al * a.B.w[i] * invJ * dxi * w.invR[n];
      if (i == 0) {   // face e: between element e-1 (L) and e (R)
        const int eL = e == 0 ? Ne - 1 : e - 1;
        double dL = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) dL += D(NP - 1, k) * x[eL * NP + k];
        const double irL = w.invR[eL * NP + NP - 1], irR = w.invR[n];
        w.fjump[e] = x[eL * NP + NP - 1] - x[n];
        w.favg[e] = invJ * (irL * dL + irR * dxi) / 2.0;
        w.fsig[e] = a.eta * fmax(irL, irR) / a.dx;
      }
    }
    __syncthreads();
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      const int eR = e == Ne - 1 ? 0 : e + 1;
      double v = a.B.w[i] * a.J * x[n] * w.invR[n];
      double st = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) st += D(k, i) * w.g[e * NP + k];
      v += st;
      // symmetry terms of the left face (this element is R) and right face (L)
      v += -al * (w.invR[e * NP] / 2.0) * invJ * D(0, i) * w.fjump[e];
      v += -al * (w.invR[e * NP + NP - 1] / 2.0) * invJ * D(NP - 1, i) * w.fjump[eR];
      if (i == 0) v += al * w.favg[e] - al * w.fsig[e] * w.fjump[e];
      if (i == NP - 1) v += -al * w.favg[eR] + al * w.fsig[eR] * w.fjump[eR];
      epi(n, v);
    }
    if (END_SYNC) __syncthreads();
  }

  // Workgroup total of one value per thread, returned to every thread, in
  // the order of block_sum (butterfly within each wave, then the wave sums in
  // wave order), so it equals dot() bit for bit.  One barrier; `slot` (LDS,
  // MAX_WAVES doubles) must not be written again before the next barrier.
  __device__ __forceinline__ double block_total(double v, double* slot) {
    v = wave_sum(v);
    const int nw = (nt + WAVE - 1) / WAVE;
    if ((tid & (WAVE - 1)) == 0) slot[tid / WAVE] = v;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < nw; ++q) t += slot[q];
    return t;
  }

  __device__ __forceinline__ double dot(const double* u, const double* v) {
    __shared__ double sred[24];
    double s[1] = {0.0};
    for (int n = tid; n < N; n += nt) s[0] += u[n] * v[n];
    block_sum<1>(s, sred);
    if (tid == 0) sred[20] = s[0];
    __syncthreads();
    const double r = sred[20];
    __syncthreads();
    return r;
  }

  // 1/rho, IGR source b and diag(A) of stage state qs
  // (compute_igr_rhs!, elliptic.jl:126-154; compute_sip_diagonal, :206-253)
  __device__ __forceinline__ void prepare(const double* qs) {
    const double* m = qs;
    const double* rho = qs + N;
    const double al = a.alpha, invJ = a.invJ;
    for (int n = tid; n < N; n += nt) {
      w.invR[n] = 1.0 / rho[n];
      w.u[n] = m[n] / rho[n];
    }
    __syncthreads();
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      double du = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) du += D(i, k) * w.u[e * NP + k];
      du *= invJ;
      w.b[n] = a.B.w[i] * a.J * (2.0 * al * (du * du));
      const double ir = w.invR[n];
      double Md = a.B.w[i] * a.J * ir;
#pragma unroll
      for (int k = 0; k < NP; ++k) Md += al * a.B.w[k] * invJ * (D(k, i) * D(k, i)) * w.invR[e * NP + k];
      if (i == NP - 1) {
        const int eR = e == Ne - 1 ? 0 : e + 1;
        const double sig = a.eta * fmax(ir, w.invR[eR * NP]) / a.dx;
        Md += -al * invJ * D(NP - 1, NP - 1) * ir;
        Md += al * sig;
      }
      if (i == 0) {
        const int eL = e == 0 ? Ne - 1 : e - 1;
        const double sig = a.eta * fmax(w.invR[eL * NP + NP - 1], ir) / a.dx;
        Md += al * invJ * D(0, 0) * ir;
        Md += al * sig;
      }
      w.invM[n] = 1.0 / Md;
    }
    __syncthreads();
  }

  // fixed-iteration solve A x = b with the prepared 1/rho and diag(A)
  // (pcg_fixed! :307-339, jacobi_fixed! :359-368, chebyshev_fixed! :389-430)
  __device__ __forceinline__ void iterate(double* x, const double* bb, int trace) {
    const int nI = a.n_iter;
    if (a.solver == 0) {
      // Four barriers per iteration: <p, Ap> is summed in the apply's
      // epilogue and <r, z> in the update, each with one block_total (two
      // slot arrays, so a total is never overwritten while being read).
      __shared__ double slotA[MAX_WAVES], slotB[MAX_WAVES];
      apply(x, w.Ap);
      double part = 0.0;
      for (int n = tid; n < N; n += nt) {
        const double r = bb[n] - w.Ap[n];
        const double z = r * w.invM[n];
        w.r[n] = r;
        w.p[n] = z;
        part += r * z;
        if (a.stR) a.stR[n] = r;
      }
      double rz = block_total(part, slotB);
      for (int it = 0; it < nI; ++it) {
        double pp = 0.0;
        apply_with<false>(w.p, [&](int n, double v) {
          w.Ap[n] = v;
          pp += w.p[n] * v;
        });
        const double pAp = block_total(pp, slotA);
        const double ak = rz / (pAp + 1e-30);
        double rr = 0.0;
        for (int n = tid; n < N; n += nt) {
          if (a.stP) { a.stP[(size_t)it * N + n] = w.p[n]; a.stAP[(size_t)it * N + n] = w.Ap[n]; }
          x[n] = x[n] + ak * w.p[n];
          const double r = w.r[n] - ak * w.Ap[n];
          const double z = r * w.invM[n];
          w.r[n] = r;
          w.z[n] = z;
          rr += r * z;
          if (a.stR) a.stR[(size_t)(it + 1) * N + n] = r;
        }
        const double rzn = block_total(rr, slotB);
        const double bk = rzn / (rz + 1e-30);
        if (trace && tid == 0 && a.trace) { a.trace[it] = ak; a.trace[nI + it] = bk; }
        if (tid == 0 && a.stsc) {   // rz_k (k <= n) | pAp_k | a_k | beta_k
          a.stsc[it] = rz; a.stsc[nI + 1 + it] = pAp; a.stsc[2 * nI + 1 + it] = ak; a.stsc[3 * nI + 1 + it] = bk;
          if (it == nI - 1) a.stsc[nI] = rzn;
        }
        for (int n = tid; n < N; n += nt) w.p[n] = w.z[n] + bk * w.p[n];
        __syncthreads();
        rz = rzn;
      }
      if (nI == 0 && tid == 0 && a.stsc) a.stsc[0] = rz;
    } else if (a.solver == 1) {
      // damped Jacobi: the update is the apply's epilogue (two barriers per iteration)
      for (int it = 0; it < nI; ++it) {
        apply_with(x, [&](int n, double v) {
          if (a.stP) { a.stP[(size_t)it * N + n] = x[n]; a.stAP[(size_t)it * N + n] = v; }
          x[n] = x[n] + (2.0 / 3.0) * (bb[n] - v) * w.invM[n];
        });
      }
    } else {
      const double lmin = 1.0, lmax = a.cheb_lmax;
      const double theta = (lmax + lmin) / 2.0, delta = (lmax - lmin) / 2.0;
      const double sigma = theta / fmax(delta, 1e-10);
      // r = b - A x as the apply's epilogue (three barriers per iteration)
      apply_with(x, [&](int n, double v) {
        w.r[n] = bb[n] - v;
        if (a.stR) a.stR[n] = w.r[n];
      });
      double rc = 1.0 / sigma;
      for (int k = 1; k <= nI; ++k) {
        double c1 = 0.0, c2 = 0.0;
        if (k > 1) { const double rn = 1.0 / (2.0 * sigma - rc); c1 = rn * rc; c2 = 2.0 * rn / delta; rc = rn; }
        for (int n = tid; n < N; n += nt) {
          const double z = w.r[n] * w.invM[n];
          const double d = (k == 1) ? z / theta : c1 * w.p[n] + c2 * z;
          w.p[n] = d;
          x[n] = x[n] + d;
          if (a.stP) a.stP[(size_t)(k - 1) * N + n] = x[n];
        }
        __syncthreads();
        apply_with(x, [&](int n, double v) {
          w.r[n] = bb[n] - v;
          if (a.stR && k < nI) a.stR[(size_t)k * N + n] = w.r[n];
        });
      }
    }
  }

  // solve_elliptic! (elliptic.jl:451-467) on state qs, warm start w.S
  __device__ __forceinline__ void solve(const double* qs, int trace) {
    prepare(qs);
    iterate(w.S, w.b, trace);
  }

  // right-hand side of the tangent elliptic equation into w.bs
  // (form 1: sensitivity.jl:168-256, SIP-like faces; form 2: :272-314, pointwise)
  __device__ __forceinline__ void sens_rhs(const double* qs, const double* ss) {
    const double al = a.alpha, invJ = a.invJ, J = a.J;
    const double* m = qs;
    const double* r = qs + N;
    const double* sm = ss;
    const double* sr = ss + N;
    const double* S = w.S;
    if (a.form == 2) {
      for (int n = tid; n < N; n += nt) {
        const int i = n % NP, e = n / NP;
        double ux = 0.0, sux = 0.0, dso = 0.0, dS = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          const int nk = e * NP + k;
          const double uj = m[nk] / r[nk];
          const double suj = (sm[nk] - uj * sr[nk]) / r[nk];
          ux += D(i, k) * uj; sux += D(i, k) * suj;
          dso += D(i, k) * (sr[nk] / r[nk]);
          dS += D(i, k) * S[nk];
        }
        ux *= invJ; sux *= invJ; dso *= invJ; dS *= invJ;
        const double so = sr[n] / r[n];
        const double R = al * (2.0 * ux * (2.0 * sux + so * ux) - dso * dS / r[n]);
        w.bs[n] = a.B.w[i] * J * R;
      }
      __syncthreads();
      return;
    }
    // form 1: faces first (kappa = s_rho / rho^2 diffusion of Sigma)
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      if (i != 0) continue;
      const int eL = e == 0 ? Ne - 1 : e - 1;
      const int nl = eL * NP + NP - 1;
      const double kL = sr[nl] / (r[nl] * r[nl]), kR = sr[n] / (r[n] * r[n]);
      double dL = 0.0, dR = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) { dL += D(NP - 1, k) * S[eL * NP + k]; dR += D(0, k) * S[e * NP + k]; }
      w.fjump[e] = S[nl] - S[n];
      w.favg[e] = invJ * (kL * dL + kR * dR) / 2.0;
      w.fsig[e] = a.eta * fmax_first(kL, kR) / a.dx;
    }
    // kappa * dSigma at every node (volume stiffness-like term)
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      double dS = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) dS += D(i, k) * S[e * NP + k];
      w.g[n] = al * a.B.w[i] * invJ * (sr[n] / (r[n] * r[n])) * dS;
    }
    __syncthreads();
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      const int eR = e == Ne - 1 ? 0 : e + 1;
      double ux = 0.0, sux = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int nk = e * NP + k;
        const double uj = m[nk] / r[nk];
        const double suj = (sm[nk] - uj * sr[nk]) / r[nk];
        ux += D(i, k) * uj; sux += D(i, k) * suj;
      }
      ux *= invJ; sux *= invJ;
      const double src = 4.0 * al * ux * sux + sr[n] * S[n] / (r[n] * r[n]);
      double v = a.B.w[i] * J * src;
      double st = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) st += w.g[e * NP + k] * D(k, i);
      v += st;
      const int n0 = e * NP, n1 = e * NP + NP - 1;
      const double k0 = sr[n0] / (r[n0] * r[n0]), k1 = sr[n1] / (r[n1] * r[n1]);
      v += -al * (k0 / 2.0) * invJ * D(0, i) * w.fjump[e];
      v += -al * (k1 / 2.0) * invJ * D(NP - 1, i) * w.fjump[eR];
      if (i == 0) v += al * w.favg[e] - al * w.fsig[e] * w.fjump[e];
      if (i == NP - 1) v += -al * w.favg[eR] + al * w.fsig[eR] * w.fjump[eR];
      w.bs[n] = v;
    }
    __syncthreads();
  }

  // sensitivity hyperbolic RHS (sensitivity.jl:91-150) + SSP-RK3 stage of a tangent
  __device__ __forceinline__ void sens_hyp_stage(const double* qs, const double* ss, const double* svs, const double* sn, int mode,
                                 double cdt, double* out);

  // hyperbolic RHS (hyperbolic.jl:23-83) fused with the SSP-RK3 stage update
  __device__ __forceinline__ void hyp_stage(const double* qs, int mode, double cdt, double* out);
};

#pragma clang fp contract(off)
template <int NP, bool LDS>
__device__ __forceinline__ void K1<NP, LDS>::hyp_stage(const double* qs, int mode, double cdt, double* out) {
  const double g = a.gamma, invJ = a.invJ;
  for (int n = tid; n < N; n += nt) {
    double f[3];
    flux1(g, qs[n], qs[N + n], qs[2 * N + n], w.S[n], f);
    w.F[n] = f[0]; w.F[N + n] = f[1]; w.F[2 * N + n] = f[2];
    const int i = n % NP, e = n / NP;
    if (i == 0) {     // LLF at face e (between e-1 and e), llf_flux euler.jl:64-78
      const int eL = e == 0 ? Ne - 1 : e - 1;
      const int nl = eL * NP + NP - 1;
      const double qL[4] = {qs[nl], qs[N + nl], qs[2 * N + nl], w.S[nl]};
      const double qR[4] = {qs[n], qs[N + n], qs[2 * N + n], w.S[n]};
      const double lam = fmax_first(wave1(g, qL, a.flag), wave1(g, qR, a.flag));
      double fL[3], fR[3];
      flux1(g, qL[0], qL[1], qL[2], qL[3], fL);
      flux1(g, qR[0], qR[1], qR[2], qR[3], fR);
#pragma unroll
      for (int v = 0; v < 3; ++v) w.fstar[v * Ne + e] = (fL[v] + fR[v]) / 2.0 + lam * (qL[v] - qR[v]) / 2.0;
    }
  }
  __syncthreads();
  for (int n = tid; n < N; n += nt) {
    const int i = n % NP, e = n / NP;
    const int eR = e == Ne - 1 ? 0 : e + 1;
    double acc[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      acc[v] = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) acc[v] -= invJ * D(i, k) * w.F[v * N + e * NP + k];
    }
    if (i == NP - 1)
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[v] -= invJ * (w.fstar[v * Ne + eR] - w.F[v * N + n]) / a.B.w[NP - 1];
    if (i == 0)
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[v] += invJ * (w.fstar[v * Ne + e] - w.F[v * N + n]) / a.B.w[0];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const size_t gi = (size_t)v * N + n;
      double o;
      if (mode == 0) o = acc[v];
      else if (mode == 1) o = qs[gi] + cdt * acc[v];
      else if (mode == 2) o = 3.0 / 4.0 * w.q[gi] + 1.0 / 4.0 * qs[gi] + cdt * acc[v];
      else o = 1.0 / 3.0 * w.q[gi] + 2.0 / 3.0 * qs[gi] + cdt * acc[v];
      out[gi] = o;
    }
  }
  __syncthreads();
}
template <int NP, bool LDS>
__device__ __forceinline__ void K1<NP, LDS>::sens_hyp_stage(const double* qs, const double* ss, const double* svs, const double* sn,
                                       int mode, double cdt, double* out) {
  const double g = a.gamma, invJ = a.invJ;
  for (int n = tid; n < N; n += nt) {
    const double qq[4] = {qs[n], qs[N + n], qs[2 * N + n], w.S[n]};
    const double sv[4] = {ss[n], ss[N + n], ss[2 * N + n], svs[n]};
    double f[3];
    sflux1(g, qq, sv, a.sgam, f);
    w.F[n] = f[0]; w.F[N + n] = f[1]; w.F[2 * N + n] = f[2];
    const int i = n % NP, e = n / NP;
    if (i == 0) {     // sensitivity_llf_flux at face e, primal wave speed (sensitivity.jl:58-78)
      const int eL = e == 0 ? Ne - 1 : e - 1;
      const int nl = eL * NP + NP - 1;
      const double qL[4] = {qs[nl], qs[N + nl], qs[2 * N + nl], w.S[nl]};
      const double sL[4] = {ss[nl], ss[N + nl], ss[2 * N + nl], svs[nl]};
      const double lam = fmax_first(wave1(g, qL, a.flag), wave1(g, qq, a.flag));
      double fL[3], fR[3];
      sflux1(g, qL, sL, a.sgam, fL);
      sflux1(g, qq, sv, a.sgam, fR);
#pragma unroll
      for (int v = 0; v < 3; ++v) w.fstar[v * Ne + e] = (fL[v] + fR[v]) / 2.0 + lam * (sL[v] - sv[v]) / 2.0;
    }
  }
  __syncthreads();
  for (int n = tid; n < N; n += nt) {
    const int i = n % NP, e = n / NP;
    const int eR = e == Ne - 1 ? 0 : e + 1;
    double acc[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      acc[v] = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) acc[v] -= invJ * D(i, k) * w.F[v * N + e * NP + k];
    }
    if (i == NP - 1)
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[v] -= invJ * (w.fstar[v * Ne + eR] - w.F[v * N + n]) / a.B.w[NP - 1];
    if (i == 0)
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[v] += invJ * (w.fstar[v * Ne + e] - w.F[v * N + n]) / a.B.w[0];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const size_t gi = (size_t)v * N + n;
      double o;
      if (mode == 0) o = acc[v];
      else if (mode == 1) o = ss[gi] + cdt * acc[v];
      else if (mode == 2) o = 3.0 / 4.0 * sn[gi] + 1.0 / 4.0 * ss[gi] + cdt * acc[v];
      else o = 1.0 / 3.0 * sn[gi] + 2.0 / 3.0 * ss[gi] + cdt * acc[v];
      out[gi] = o;
    }
  }
  __syncthreads();
}
#pragma clang fp contract(fast)

}  // namespace igr
