// 2D forward tangent (sensitivity) kernels, gfx950, FP64 — the linearized
// IGR-Euler system of IGR_adjoints_2d/sensitivity.jl, advanced together
// with the primal state.
//
//   k2_sens_rhs   right-hand side b of the sensitivity elliptic equation
//                 (compute_sensitivity_elliptic_rhs!, sensitivity.jl:236-292):
//                 pointwise after in-element derivatives, so no halo.  The
//                 solve itself reuses the primal SIP/PCG kernels with this b.
//   k2_sens_hyp   sensitivity hyperbolic RHS (sensitivity.jl:131-215) with
//                 the LLF dissipation at the primal wave speed, fused with the
//                 SSP-RK3 stage combination of the tangent state.  Reference
//                 evaluation order without FMA contraction, like k2_hyp.
#include "igr_ctx.hpp"
#include "k2_common.hpp"
#include "launch_rec.hpp"

namespace igr {


// No implicit contraction (explicit fma()): the single and batched
// instantiations round alike.
#pragma clang fp contract(off)
template <int NP, int TX, int TY>
__device__ __forceinline__ void k2_sens_rhs_body(const SensArgs& a, const int blk, const int nblk) {
  constexpr int NP2 = NP * NP, T = TX * TY;
  __shared__ double uu[T * NP2], vv[T * NP2], su[T * NP2], sv[T * NP2], Ss[T * NP2], sr[T * NP2];
  __shared__ double sD[NP * NP], sW[NP];     // basis (stage_basis)
  const Geo2 g = a.g;
  const BasisK& B = a.B;
  TileInfo<NP, TX, TY> t(g, blk, nblk);
  const int tid = threadIdx.x;
  const int s = tid / NP2, n = tid % NP2, i = n % NP, j = n / NP;
  const int lx = s % TX, ly = s / TX;
  const bool valid = lx < t.nvx && ly < t.nvy;
  const size_t N = (size_t)g.Nex * g.Ney * NP2;
  const size_t gi = valid ? t.gidx(s, n) : 0;
  if (valid) {
    // sensitivity.jl:257-265: u = mu/rho, s_u = (zeta - u sigma)/rho
    const double r = a.q[2 * N + gi];
    const double ux = a.q[gi] / r, uy = a.q[N + gi] / r;
    const double sg = a.s[2 * N + gi];
    uu[tid] = ux; vv[tid] = uy;
    su[tid] = (a.s[gi] - ux * sg) / r;
    sv[tid] = (a.s[N + gi] - uy * sg) / r;
    Ss[tid] = a.S[gi];
    sr[tid] = sg / r;
  }
  stage_basis<NP>(B, sD, sW);
  __syncthreads();
  if (!valid) return;
  double uxx = 0.0, uxy = 0.0, uyx = 0.0, uyy = 0.0, sxx = 0.0, sxy = 0.0, syx = 0.0, syy = 0.0;
  double Sx = 0.0, Sy = 0.0, dxs = 0.0, dys = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int kj = s * NP2 + k + NP * j, ik = s * NP2 + i + NP * k;
    const double Di = sD[i * NP + k], Dj = sD[j * NP + k];
    uxx = fma(Di, uu[kj], uxx); uyx = fma(Di, vv[kj], uyx);
    uxy = fma(Dj, uu[ik], uxy); uyy = fma(Dj, vv[ik], uyy);
    sxx = fma(Di, su[kj], sxx); syx = fma(Di, sv[kj], syx);
    sxy = fma(Dj, su[ik], sxy); syy = fma(Dj, sv[ik], syy);
    Sx = fma(Di, Ss[kj], Sx); Sy = fma(Dj, Ss[ik], Sy);
    dxs = fma(Di, sr[kj], dxs); dys = fma(Dj, sr[ik], dys);
  }
  uxx *= g.invJx; uyx *= g.invJx; uxy *= g.invJy; uyy *= g.invJy;
  sxx *= g.invJx; syx *= g.invJx; sxy *= g.invJy; syy *= g.invJy;
  Sx *= g.invJx; Sy *= g.invJy;
  dxs *= g.invJx; dys *= g.invJy;
  const double al = a.alpha;
  const double divu = uxx + uyy, divs = sxx + syy;
  const double Rp = al * fma(uyy, uyy, fma(2.0 * uxy, uyx, fma(uxx, uxx, divu * divu)));
  const double Rs = al * fma(2.0 * divu, divs, 2.0 * fma(uyy, syy, fma(uyx, sxy, fma(uxy, syx, uxx * sxx))));
  const double r = a.q[2 * N + gi];
  const double bpt = fma(sr[tid], Rp, Rs) - al * fma(dxs, Sx, dys * Sy) / r;
  a.out[gi] = sW[i] * sW[j] * g.Jx * g.Jy * bpt;
}
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_sens_rhs(SensArgs a) {
  k2_sens_rhs_body<NP, TX, TY>(a, blockIdx.x, gridDim.x);
}
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_sens_rhs_batch(const SensArgs* __restrict__ tab,
                                                                  const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const SensArgs& a = tab[w.x];
  k2_sens_rhs_body<NP, TX, TY>(a, w.y, tiles_of<TX, TY>(a.g));
}


#pragma clang fp contract(off)
// sensitivity_flux_x / _y with _pressure_sensitivity (sensitivity.jl:34-76).
// q = (mx, my, rho, E, Sigma), sv = (zx, zy, sigma, eta, varsigma)
__device__ __forceinline__ void sflux2(int dir, double gam, const double q[5], const double sv[5], double sgam,
                                       double f[4]) {
  const double mx = q[0], my = q[1], r = q[2], E = q[3], S = q[4];
  const double zx = sv[0], zy = sv[1], sg = sv[2], et = sv[3], vs = sv[4];
  const double ux = mx / r, uy = my / r;
  const double sux = (zx - ux * sg) / r;
  const double suy = (zy - uy * sg) / r;
  const double P = (gam - 1.0) * (E - (mx * mx + my * my) / (2.0 * r));
  double Phi = (gam - 1.0) * (et - sg * (ux * ux + uy * uy) / 2.0 - r * (ux * sux + uy * suy));
  // gamma source of a d/dgamma tangent, as in the 1D tangent (IGR_adjoints_1d/sensitivity.jl:41);
  // the reference's 2D tangent has none (sensitivity.jl:15-16): sgam = 0
  if (sgam != 0.0) Phi = Phi + sgam * r * (E / r - (ux * ux + uy * uy) / 2.0);
  const double Pt = Phi + vs;
  if (dir == 0) {
    f[0] = zx * ux + mx * sux + Pt; f[1] = zx * uy + mx * suy; f[2] = zx;
    f[3] = (et + Pt) * ux + (E + P + S) * sux;
  } else {
    f[0] = zx * uy + mx * suy; f[1] = zy * uy + my * suy + Pt; f[2] = zy;
    f[3] = (et + Pt) * uy + (E + P + S) * suy;
  }
}
__device__ __forceinline__ double pwave(int dir, double gam, const double q[5], int* flag) {
  const double P = (gam - 1.0) * (q[3] - (q[0] * q[0] + q[1] * q[1]) / (2.0 * q[2]));
  const double arg = gam * P / q[2];
  if (!(arg >= 0.0)) flag_domain(flag);
  return fabs(q[dir] / q[2]) + sqrt(arg);
}

template <int NP, int TX, int TY>
__device__ __forceinline__ void k2_sens_hyp_body(const SensArgs& a, const int blk, const int nblk) {
  constexpr int NP2 = NP * NP, T = TX * TY, NH = 2 * (TX + TY), NT = T * NP2;
  __shared__ double qs[5][T * NP2], ss[5][T * NP2];
  __shared__ double Fx[4][T * NP2], Fy[4][T * NP2];
  __shared__ double hq[5][NH * NP], hs[5][NH * NP];
  __shared__ double fxr[4][T * NP], fxl[4][T * NP], fyt[4][T * NP], fyb[4][T * NP];
  __shared__ double sD[NP * NP];     // basis (stage_basis)

  const Geo2 g = a.g;
  const BasisK& B = a.B;
  const double gam = a.gamma;
  TileInfo<NP, TX, TY> t(g, blk, nblk);
  const int tid = threadIdx.x;
  const int s = tid / NP2, n = tid % NP2, i = n % NP, j = n / NP;
  const int lx = s % TX, ly = s / TX;
  const bool valid = lx < t.nvx && ly < t.nvy;
  const size_t N = (size_t)g.Nex * g.Ney * NP2;
  const size_t gme = valid ? t.gidx(s, n) : 0;

  double qo[5] = {0.0, 0.0, 1.0, 1.0, 0.0}, so[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double sn[4] = {0.0, 0.0, 0.0, 0.0};   // s^n of stages 2 and 3, loaded with the stage values
  if (valid) {
#pragma unroll
    for (int v = 0; v < 4; ++v) { qo[v] = a.q[v * N + gme]; so[v] = a.s[v * N + gme]; }
    qo[4] = a.S[gme]; so[4] = a.vs[gme];
    if (a.mode >= 2) {
#pragma unroll
      for (int v = 0; v < 4; ++v) sn[v] = a.sn[v * N + gme];
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) { qs[v][tid] = qo[v]; ss[v][tid] = so[v]; }
  for (int idx = tid; idx < NH * NP; idx += NT) {
    const int h = idx / NP, k = idx % NP;
    size_t gi;
    if (t.halo_face_global(h, k, gi)) {
#pragma unroll
      for (int v = 0; v < 4; ++v) { hq[v][idx] = a.q[v * N + gi]; hs[v][idx] = a.s[v * N + gi]; }
      hq[4][idx] = a.S[gi]; hs[4][idx] = a.vs[gi];
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) { hq[v][idx] = (v == 2 || v == 3) ? 1.0 : 0.0; hs[v][idx] = 0.0; }
    }
  }
  double fxo[4], fyo[4];
  sflux2(0, gam, qo, so, a.sgam, fxo);
  sflux2(1, gam, qo, so, a.sgam, fyo);
#pragma unroll
  for (int v = 0; v < 4; ++v) { Fx[v][tid] = fxo[v]; Fy[v][tid] = fyo[v]; }
  stage_basis<NP>(B, sD, nullptr);
  __syncthreads();

  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  auto nb = [&](int sl, int node, int hk, double qn[5], double sn[5]) {
    if (sl < T) {
#pragma unroll
      for (int v = 0; v < 5; ++v) { qn[v] = qs[v][sl * NP2 + node]; sn[v] = ss[v][sl * NP2 + node]; }
    } else {
      const int h = sl - T;
#pragma unroll
      for (int v = 0; v < 5; ++v) { qn[v] = hq[v][h * NP + hk]; sn[v] = hs[v][h * NP + hk]; }
    }
  };
  // sensitivity_llf_flux_x/_y (sensitivity.jl:85-118)
  auto sllf = [&](int dir, const double qL[5], const double qR[5], const double sL[5], const double sR[5],
                  double f[4]) {
    const double lam = fmax_first(pwave(dir, gam, qL, a.flag), pwave(dir, gam, qR, a.flag));
    double fL[4], fR[4];
    sflux2(dir, gam, qL, sL, a.sgam, fL);
    sflux2(dir, gam, qR, sR, a.sgam, fR);
#pragma unroll
    for (int v = 0; v < 4; ++v) f[v] = (fL[v] + fR[v]) / 2.0 + lam * (sL[v] - sR[v]) / 2.0;
  };
  if (valid) {
#pragma unroll
    for (int k = 0; k < NP; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        acc[v] -= g.invJx * sD[i * NP + k] * Fx[v][s * NP2 + k + NP * j] +
                  g.invJy * sD[j * NP + k] * Fy[v][s * NP2 + i + NP * k];
    if (i == NP - 1) {
      double qn[5], sn[5], f[4];
      nb(t.nb_right(s), NP * j, j, qn, sn);
      sllf(0, qo, qn, so, sn, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fxr[v][s * NP + j] = f[v];
    }
    if (i == 0 && lx == 0) {
      double qn[5], sn[5], f[4];
      nb(t.nb_left(s), (NP - 1) + NP * j, j, qn, sn);
      sllf(0, qn, qo, sn, so, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fxl[v][s * NP + j] = f[v];
    }
    if (j == NP - 1) {
      double qn[5], sn[5], f[4];
      nb(t.nb_top(s), i, i, qn, sn);
      sllf(1, qo, qn, so, sn, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fyt[v][s * NP + i] = f[v];
    }
    if (j == 0 && ly == 0) {
      double qn[5], sn[5], f[4];
      nb(t.nb_bottom(s), i + NP * (NP - 1), i, qn, sn);
      sllf(1, qn, qo, sn, so, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fyb[v][s * NP + i] = f[v];
    }
  }
  __syncthreads();
  if (!valid) return;
  if (i == NP - 1)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] -= g.invJx * (fxr[v][s * NP + j] - fxo[v]) / B.w[NP - 1];
  if (i == 0) {
    const int sL = t.nb_left(s);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double f = (sL < T) ? fxr[v][sL * NP + j] : fxl[v][s * NP + j];
      acc[v] += g.invJx * (f - fxo[v]) / B.w[0];
    }
  }
  if (j == NP - 1)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] -= g.invJy * (fyt[v][s * NP + i] - fyo[v]) / B.w[NP - 1];
  if (j == 0) {
    const int sB = t.nb_bottom(s);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double f = (sB < T) ? fyt[v][sB * NP + i] : fyb[v][s * NP + i];
      acc[v] += g.invJy * (f - fyo[v]) / B.w[0];
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const size_t gi = v * N + gme;
    double o;
    if (a.mode == 0) o = acc[v];
    else if (a.mode == 1) o = so[v] + a.cdt * acc[v];
    else if (a.mode == 2) o = 3.0 / 4.0 * sn[v] + 1.0 / 4.0 * so[v] + a.cdt * acc[v];
    else o = 1.0 / 3.0 * sn[v] + 2.0 / 3.0 * so[v] + a.cdt * acc[v];
    a.out[gi] = o;
  }
}
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_sens_hyp(SensArgs a) {
  k2_sens_hyp_body<NP, TX, TY>(a, blockIdx.x, gridDim.x);
}
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_sens_hyp_batch(const SensArgs* __restrict__ tab,
                                                                  const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const SensArgs& a = tab[w.x];
  k2_sens_hyp_body<NP, TX, TY>(a, w.y, tiles_of<TX, TY>(a.g));
}

#pragma clang fp contract(fast)

template <int NP, int TX, int TY>
struct LaunchSens {
  static dim3 grid(const Geo2& g) { return dim3(((g.Nex + TX - 1) / TX) * ((g.Ney + TY - 1) / TY)); }
  static void rhs(const SensArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k2_sens_rhs<NP, TX, TY>), grid(a.g), dim3(TX * TY * NP * NP), 0, st, a);
  }
  static void hyp(const SensArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k2_sens_hyp<NP, TX, TY>), grid(a.g), dim3(TX * TY * NP * NP), 0, st, a);
  }
  static void rhs_batch(const SensArgs* tab, const int2* work, int nb, hipStream_t st) {
    hipLaunchKernelGGL((k2_sens_rhs_batch<NP, TX, TY>), dim3(nb), dim3(TX * TY * NP * NP), 0, st, tab, work);
  }
  static void hyp_batch(const SensArgs* tab, const int2* work, int nb, hipStream_t st) {
    hipLaunchKernelGGL((k2_sens_hyp_batch<NP, TX, TY>), dim3(nb), dim3(TX * TY * NP * NP), 0, st, tab, work);
  }
};
#define IGR_SENS_DISPATCH(np, tx, ty, CALL)                                     \
  switch (np) {                                                                 \
    case 2: LaunchSens<2, 8, 8>::CALL; break;                                   \
    case 3: LaunchSens<3, 4, 4>::CALL; break;                                   \
    case 4:                                                                     \
      if ((tx) == 8 && (ty) == 4) LaunchSens<4, 8, 4>::CALL;                    \
      else if ((tx) == 2 && (ty) == 2) LaunchSens<4, 2, 2>::CALL;               \
      else LaunchSens<4, 4, 4>::CALL;                                           \
      break;                                                                    \
    case 5: LaunchSens<5, 2, 4>::CALL; break;                                   \
  }
void launch2_sens_rhs(const Ctx& c, const SensArgs& a) {
  if (g_rec) { g_rec->recs.push_back({R_SENS_RHS, 0, g_rec->sens.size()}); g_rec->sens.push_back(a); return; }
  IGR_SENS_DISPATCH(c.np, c.htile_x, c.htile_y, rhs(a, c.stream));
}
void launch2_sens_hyp(const Ctx& c, const SensArgs& a) {
  if (g_rec) { g_rec->recs.push_back({R_SENS_HYP, 0, g_rec->sens.size()}); g_rec->sens.push_back(a); return; }
  IGR_SENS_DISPATCH(c.np, c.htile_x, c.htile_y, hyp(a, c.stream));
}
void launch2_sens_rhs_batch(int np, int tx, int ty, const SensArgs* tab, const int2* work, int nblocks,
                            hipStream_t st) {
  IGR_SENS_DISPATCH(np, tx, ty, rhs_batch(tab, work, nblocks, st));
}
void launch2_sens_hyp_batch(int np, int tx, int ty, const SensArgs* tab, const int2* work, int nblocks,
                            hipStream_t st) {
  IGR_SENS_DISPATCH(np, tx, ty, hyp_batch(tab, work, nblocks, st));
}

}  // namespace igr
