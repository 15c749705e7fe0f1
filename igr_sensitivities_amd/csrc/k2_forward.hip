// 2D forward kernels for gfx950 (MI355X), FP64.
//
// Data layout in HBM: one array per field (SoA), nodes in the reference's
// column-major order, so one element of p=3 is 16 contiguous doubles
// (128 B) and a wavefront of 64 lanes reads 4 whole elements per field.
//
// Work decomposition: a workgroup owns a TX x TY tile of elements and one
// thread per node (TX*TY*NP^2 threads; 4x4 tiles of p=3 give 256).  The
// tile plus a one-element halo ring (no corners) is staged in LDS.  Tiles
// are assigned XCD-contiguously (xcd_tile) so halo reads of neighbouring
// tiles hit the same L2.
//
// Kernels:
//   k2_sip<MODE>   matrix-free SIP apply y = A(rho) x  (apply_sip!,
//                  IGR_adjoints_2d/elliptic.jl:30-145) in gather form, fused
//                  with what surrounds it in each solver:
//                    INIT_*  IGR source b (compute_igr_rhs!, :159-197),
//                            diag(A) (compute_sip_diagonal, :205-264),
//                            1/rho, 1/M, r0 = b - A x0 and <r0, z0>
//                    PCG_A   p = z + beta p on the fly (own + halo nodes),
//                            Ap = A p, <p, Ap>       (pcg_fixed!, :329-347)
//                    JACOBI  x' = x + w (b - A x)/M  (jacobi_fixed!, :354-361)
//                    CHEB    d' = c1 d + c2 z, x' = x + d', r' = b - A x'
//                            (chebyshev_fixed!, :370-397)
//   k2_pcg_b       x += a p, r -= a Ap, <r, z>   (pcg_fixed!, :339-343)
//   k2_hyp         hyperbolic RHS (compute_hyperbolic_rhs!,
//                  IGR_adjoints_2d/hyperbolic.jl:27-119) fused with the
//                  SSP-RK3 stage combination (forward.jl:51-69), each face
//                  LLF flux computed once per tile (euler.jl:97-133).
// Dot products finish with the deterministic last-workgroup reduction.
#include <algorithm>
#include <type_traits>
#include <hip/hip_ext.h>
#include "igr_ctx.hpp"
#include "k2_common.hpp"
#include "launch_rec.hpp"

namespace igr {

// ------------------------------------------------------------------------
// SIP family
// ------------------------------------------------------------------------
// No implicit contraction: every fused multiply-add is written as fma(), so
// the single and batched instantiations of a body round alike (with
// contraction left to the compiler, a*b + c*d may fuse either product, and
// the choice can differ between instantiations of the same source).
#pragma clang fp contract(off)
// GU: chunks of the previous kernel's partial sums loaded together; a batched
// launch sums one instance's few partials and keeps the registers instead
template <int NP, int TX, int TY, int MODE, int GU = 8>
__device__ __forceinline__ void k2_sip_body(const SipArgs& a, const int blk, const int nblk) {
  constexpr int NP2 = NP * NP, T = TX * TY, NS = T + 2 * (TX + TY), NT = T * NP2;
  constexpr bool INIT = (MODE == M_INIT_PCG || MODE == M_INIT_CHEB || MODE == M_INIT_JAC);
  // LDS bank layout.  A 64-bit LDS read is served in two passes of 32 lanes
  // over a bank space of 32 doubles (same-address lanes broadcast), a 64-bit
  // write in four passes of 16 lanes over 16 doubles.  A face pass reads the
  // same node of 32 different element slots, which with a slot stride of 16
  // doubles all fall on one bank; an odd slot stride SS spreads them over 32.
  // The face terms are stored face-major, one row of FR doubles per face
  // node (F, q) with the tile elements along the row, so the face pass
  // writes consecutive doubles and the owner's reads of its 4 NP terms fall
  // on distinct banks.
  constexpr int SS = NP2 % 2 == 0 ? NP2 + 1 : NP2, FR = T + 4;
  __shared__ double vs[NS * SS];
  __shared__ double irs[NS * SS];
  __shared__ double gx[T * NP2], gy[T * NP2];
  __shared__ double fP[4 * NP * FR], fQ[4 * NP * FR], fS[4 * NP * FR];   // face-node terms
  __shared__ double fG[INIT ? 4 * NP * FR : 1];                         // penalty coefficients
  __shared__ double us[INIT ? T * NP2 : 1], ws_[INIT ? T * NP2 : 1];

  const Geo2 g = a.g;
  const BasisK& B = a.B;
  TileInfo<NP, TX, TY> t(g, blk, nblk);
  const int tid = threadIdx.x;
  const int s = tid / NP2, n = tid % NP2, i = n % NP, j = n / NP;
  const int lx = s % TX, ly = s / TX;
  const bool valid = lx < t.nvx && ly < t.nvy;
  const size_t gme = valid ? t.gidx(s, n) : 0;

  // basis in LDS (stage_basis), read per lane below
  __shared__ double sD[NP * NP], sW[NP];
  __shared__ unsigned arrive;   // waves done with the partial sums (block_partials_store_last)
  if (tid == 0) arrive = 0;

  double own_d = 0.0;
  double beta = 0.0, sc_tot = 0.0;
  // PCG_A: the previous kernel's partial sums are requested first
  double pre[1] = {0.0};
  if constexpr (MODE == M_PCG_A) grid_prefetch<1>(a.part_in, a.np_in, pre);

  // Tile + halo loads.  Every slot's node index first, then all the loads
  // with no branch between them (an empty slot reads node 0 and is zeroed
  // afterwards), so they are in flight together; a load inside a per-slot
  // branch waits for the previous slot's data before it is issued.
  constexpr int NL = (NS * NP2 + NT - 1) / NT;
  size_t gis[NL];
  bool oks[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    size_t gi = 0;
    oks[l] = t.round_global(l, s, n, gi);
    gis[l] = gi;
  }
  // PCG_A: r, 1/M, p_old, 1/rho; CHEB: r, 1/M, d, x, 1/rho; INIT: x, rho, mx, my;
  // APPLY: x, rho; JACOBI: x, 1/rho.  p_old and d are not used in a first
  // iteration; r stands in for them there, so every load has a valid address.
  const double* pold = a.first ? a.r : a.pold;
  const double* dold = a.first ? a.r : a.d;
  double L0[NL], L1[NL], L2[NL], L3[NL], L4[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const size_t gi = gis[l];
    L2[l] = 0.0; L3[l] = 0.0; L4[l] = 0.0;
    if constexpr (MODE == M_PCG_A) {
      L0[l] = a.r[gi]; L1[l] = a.invM[gi]; L2[l] = pold[gi]; L3[l] = a.invR[gi];
    } else if constexpr (MODE == M_APPLY) {
      L0[l] = a.x[gi]; L1[l] = a.rho[gi];
    } else if constexpr (INIT) {
      L0[l] = a.x[gi]; L1[l] = a.rho[gi]; L2[l] = a.mx[gi]; L3[l] = a.my[gi];
    } else if constexpr (MODE == M_JACOBI) {
      L0[l] = a.x[gi]; L1[l] = a.invR[gi];
    } else {   // M_CHEB
      L0[l] = a.r[gi]; L1[l] = a.invM[gi]; L2[l] = dold[gi]; L3[l] = a.x[gi]; L4[l] = a.invR[gi];
    }
  }
  if constexpr (MODE == M_PCG_A) {
    // beta_{it-1} = rz_it / (rz_{it-1} + 1e-30), then p = z + beta p_old in LDS
    const int nit = a.n_iter, it = a.it;
    const double rz_prev = a.first ? 0.0 : a.sc[it - 1];
    double tot[1];
    grid_finish<1, GU>(a.part_in, a.np_in, pre, tot);
    if (!a.first) beta = tot[0] / (rz_prev + 1e-30);
    sc_tot = tot[0];   // stored with the outputs, after the last barrier
    (void)nit;
  }
  // ---- stage the vector, 1/rho (and velocities for INIT) in LDS
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const int idx = tid + l * NT;
    if (idx >= NS * NP2) continue;
    const int sl = idx / NP2;
    double v = 0.0, ir = 0.0;
    if (oks[l]) {
      if constexpr (MODE == M_PCG_A) {
        const double lz = L0[l] * L1[l];
        v = a.first ? lz : fma(beta, L2[l], lz);
        ir = L3[l];
      } else if constexpr (MODE == M_APPLY) {
        v = L0[l]; ir = 1.0 / L1[l];
      } else if constexpr (INIT) {
        v = L0[l];
        const double rh = L1[l];
        ir = 1.0 / rh;
        if (sl < T) { us[idx] = L2[l] / rh; ws_[idx] = L3[l] / rh; }
      } else if constexpr (MODE == M_JACOBI) {
        v = L0[l]; ir = L1[l];
      } else {   // M_CHEB
        const double z = L0[l] * L1[l];
        const double dn = a.first ? z / a.theta : fma(a.c1, L2[l], a.c2 * z);
        v = L3[l] + dn;
        ir = L4[l];
        if (idx == tid) own_d = dn;
      }
    }
    const int o = sl * SS + (idx - sl * NP2);
    vs[o] = v;
    irs[o] = ir;
  }
  stage_basis<NP>(B, sD, sW);
  __syncthreads();
  // PCG_A after the first iteration: x += a_{it-1} p_{it-1} at the own node
  // (the previous PCG_B left x alone); x is needed only with the outputs, so
  // it is requested here, out of the entry burst, and arrives under the
  // derivative and face phases
  // (so is a_{it-1}: read where it is used, the compiler made it a per-lane
  // load after the last phase, whose latency the kernel then waited out)
  double x_own = 0.0, a_prev = 0.0;
  if constexpr (MODE == M_PCG_A) {
    x_own = (a.first ? a.r : a.xout)[gme];
    a_prev = a.first ? 0.0 : a.sc[2 * a.n_iter + 1 + a.it - 1];
  }

  const double al = a.alpha;
  const double vown = vs[s * SS + n], ir = irs[s * SS + n];
  double bval = 0.0;
  if (valid) {
    double dxi = 0.0, deta = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      dxi = fma(sD[i * NP + k], vs[s * SS + k + NP * j], dxi);
      deta = fma(sD[j * NP + k], vs[s * SS + i + NP * k], deta);
    }
    gx[s * NP2 + n] = al * sW[i] * sW[j] * g.Jy * g.invJx * ir * dxi;
    gy[s * NP2 + n] = al * sW[i] * sW[j] * g.Jx * g.invJy * ir * deta;
    if constexpr (INIT) {    // IGR source R = alpha[(div u)^2 + tr((Du)^2)]
      double uxx = 0.0, uxy = 0.0, uyx = 0.0, uyy = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        uxx = fma(sD[i * NP + k], us[s * NP2 + k + NP * j], uxx);
        uyx = fma(sD[i * NP + k], ws_[s * NP2 + k + NP * j], uyx);
        uxy = fma(sD[j * NP + k], us[s * NP2 + i + NP * k], uxy);
        uyy = fma(sD[j * NP + k], ws_[s * NP2 + i + NP * k], uyy);
      }
      uxx *= g.invJx; uyx *= g.invJx; uxy *= g.invJy; uyy *= g.invJy;
      const double dv = uxx + uyy;
      const double R = al * fma(uyy, uyy, fma(2.0 * uxy, uyx, fma(uxx, uxx, dv * dv)));
      bval = sW[i] * sW[j] * g.Jx * g.Jy * R;
      if (a.b_in) bval = a.b_in[gme];
    }
  }
  // Face terms.  An element has 4 NP face nodes (left, right, bottom, top
  // faces, NP nodes each); the tile's T 4 NP face nodes are spread over the
  // workgroup face-major (item w: face F = w / (T NP), then node q along
  // the face, then element, fastest), so a wave works on one face direction and its
  // D rows, strides and left/right roles are compile-time constants (a
  // per-lane mix of the four faces spent most of this phase on addressing).
  // Each face node (F, q) of element se gets its consistency term P, penalty
  // term Q and symmetry factor S in LDS at (F NP + q) FR + se; the owner node adds its own after the
  // barrier.  Faces F = 0, 2 (left, bottom) have the owner on their right
  // side, F = 1, 3 on their left.
  constexpr int NF = 4 * NP, NFI = T * NF, NFR = (NFI + NT - 1) / NT;
  auto face_node = [&](auto Fc, const int se, const int q) {
    constexpr int F = decltype(Fc)::value;
    constexpr bool xf = F < 2, ownL = (F & 1) != 0;
    constexpr int dro = ownL ? NP - 1 : 0, drn = ownL ? 0 : NP - 1;   // normal index, own / neighbour
    constexpr int stp = xf ? 1 : NP;
    double P = 0.0, Q = 0.0, S = 0.0, G = 0.0;
    const int lxe = se % TX, lye = se / TX;
    if (lxe < t.nvx && lye < t.nvy) {
      const int sN = F == 0 ? t.nb_left(se) : F == 1 ? t.nb_right(se) : F == 2 ? t.nb_bottom(se) : t.nb_top(se);
      const int bo = se * SS + (xf ? NP * q : q), bn = sN * SS + (xf ? NP * q : q);
      double dno = 0.0, dnn = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        dno = fma(B.D[dro * NP + k], vs[bo + stp * k], dno);
        dnn = fma(B.D[drn * NP + k], vs[bn + stp * k], dnn);
      }
      const double vo = vs[bo + stp * dro], iro = irs[bo + stp * dro];
      const double xn = vs[bn + stp * drn], irn = irs[bn + stp * drn];
      const double wf = xf ? sW[q] * g.Jy : sW[q] * g.Jx;
      const double invJ = xf ? g.invJx : g.invJy, h = xf ? g.dx : g.dy;
      const double jump = ownL ? vo - xn : xn - vo;
      const double irLs = ownL ? iro : irn, irRs = ownL ? irn : iro;
      const double dLs = ownL ? dno : dnn, dRs = ownL ? dnn : dno;
      const double avg = invJ * fma(irLs, dLs, irRs * dRs) / 2.0;
      G = g.eta * fmax(irLs, irRs) / h;
      P = al * avg * wf;
      Q = al * G * jump * wf;
      S = -al * (iro / 2.0) * invJ * jump * wf;
    }
    const int f = (F * NP + q) * FR + se;
    fP[f] = P;
    fQ[f] = Q;
    fS[f] = S;
    if constexpr (INIT) fG[f] = G;
  };
#pragma unroll
  for (int r = 0; r < NFR; ++r) {
    const int w = tid + r * NT;
    if (NFI % NT != 0 && w >= NFI) continue;
    const int F = w / (T * NP), rem = w - F * (T * NP);
    const int q = rem / T, se = rem - q * T;
    switch (F) {
      case 0: face_node(std::integral_constant<int, 0>{}, se, q); break;
      case 1: face_node(std::integral_constant<int, 1>{}, se, q); break;
      case 2: face_node(std::integral_constant<int, 2>{}, se, q); break;
      default: face_node(std::integral_constant<int, 3>{}, se, q); break;
    }
  }
  __syncthreads();

  // this element's term of face node (F, q): row F NP + q, column s
  const int fb = s;
  constexpr int R1 = NP * FR, R2 = 2 * NP * FR, R3 = 3 * NP * FR;
  double y = 0.0;
  if (valid) {
    double fc = 0.0;        // this node's own face terms (consistency + penalty)
    if (i == 0) { fc += fP[fb + j * FR]; fc -= fQ[fb + j * FR]; }
    if (i == NP - 1) { fc -= fP[fb + R1 + j * FR]; fc += fQ[fb + R1 + j * FR]; }
    if (j == 0) { fc += fP[fb + R2 + i * FR]; fc -= fQ[fb + R2 + i * FR]; }
    if (j == NP - 1) { fc -= fP[fb + R3 + i * FR]; fc += fQ[fb + R3 + i * FR]; }
    y = sW[i] * sW[j] * g.Jx * g.Jy * vown * ir;
    double st = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      st = fma(sD[k * NP + i], gx[s * NP2 + k + NP * j], st);
      st = fma(sD[k * NP + j], gy[s * NP2 + i + NP * k], st);
    }
    y += st;
    y += fma(fS[fb + j * FR], sD[i], fS[fb + R1 + j * FR] * sD[(NP - 1) * NP + i]);
    y += fma(fS[fb + R2 + i * FR], sD[j], fS[fb + R3 + i * FR] * sD[(NP - 1) * NP + j]);
    y += fc;
  }

  if constexpr (MODE == M_APPLY) {
    if (valid) store_wt(&a.y[gme], y);
  } else if constexpr (INIT) {
    double red_v[1] = {0.0};
    double o_iM = 0.0, o_r = 0.0;
    if (valid) {
      // diag(A): mass, x/y stiffness self-coupling, face terms
      double Md = sW[i] * sW[j] * g.Jx * g.Jy * ir;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const double D2 = sD[k * NP + i] * sD[k * NP + i];
        Md = fma(al * sW[k] * sW[j] * g.Jy * g.invJx * D2, irs[s * SS + k + NP * j], Md);
      }
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const double D2 = sD[k * NP + j] * sD[k * NP + j];
        Md = fma(al * sW[i] * sW[k] * g.Jx * g.invJy * D2, irs[s * SS + i + NP * k], Md);
      }
      if (i == NP - 1) Md = fma(fma(al, fG[fb + R1 + j * FR], -al * g.invJx * B.D[(NP - 1) * NP + NP - 1] * ir), sW[j] * g.Jy, Md);
      if (i == 0) Md = fma(fma(al, fG[fb + j * FR], al * g.invJx * B.D[0] * ir), sW[j] * g.Jy, Md);
      if (j == NP - 1) Md = fma(fma(al, fG[fb + R3 + i * FR], -al * g.invJy * B.D[(NP - 1) * NP + NP - 1] * ir), sW[i] * g.Jx, Md);
      if (j == 0) Md = fma(fma(al, fG[fb + R2 + i * FR], al * g.invJy * B.D[0] * ir), sW[i] * g.Jx, Md);
      const double iM = 1.0 / Md;
      o_iM = iM;
      if (!(ir > 0.0) || !isfinite(ir)) flag_domain(a.flag);
      if constexpr (MODE == M_INIT_PCG) {
        const double r0 = bval - y;
        o_r = r0;
        red_v[0] = r0 * (r0 * iM);
      } else if constexpr (MODE == M_INIT_CHEB) {
        o_r = bval - y;
      }
    }
    if constexpr (MODE == M_INIT_PCG) block_partials_store_last<1>(red_v, a.part_out, nblk, blk, &arrive);
    if (valid) {
      store_wt(&a.invR_out[gme], ir);
      store_wt(&a.invM_out[gme], o_iM);
      if constexpr (MODE == M_INIT_PCG) {
        store_wt(&a.rout[gme], o_r);
      } else if constexpr (MODE == M_INIT_CHEB) {
        store_wt(&a.rout[gme], o_r);
        store_wt(&a.b_out[gme], bval);
      } else {
        store_wt(&a.b_out[gme], bval);
      }
    }
  } else if constexpr (MODE == M_PCG_A) {
    double red_v[1] = {valid ? vown * y : 0.0};
    block_partials_store_last<1>(red_v, a.part_out, nblk, blk, &arrive);
    if (valid) {
      store_wt(&a.pnew[gme], vown);
      store_wt(&a.y[gme], y);
      if (!a.first) store_wt(&a.xout[gme], fma(a_prev, L2[0], x_own));
    }
    if (blk == 0 && tid == 0) {
      a.sc[a.it] = sc_tot;
      if (!a.first) a.sc[3 * a.n_iter + 1 + a.it - 1] = beta;
    }
  } else if constexpr (MODE == M_JACOBI) {
    if (valid) {
      if (a.y) store_wt(&a.y[gme], y);     // A x_k, kept by the adjoint's recompute
      store_wt(&a.xout[gme], fma(a.omega * (a.b[gme] - y), a.invM[gme], vown));
    }
  } else if constexpr (MODE == M_CHEB) {
    if (valid) {
      store_wt(&a.xout[gme], vown);
      store_wt(&a.dout[gme], own_d);
      store_wt(&a.rout[gme], a.b[gme] - y);
    }
  }
}
template <int NP, int TX, int TY, int MODE>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_sip(SipArgs a) {
  k2_sip_body<NP, TX, TY, MODE>(a, blockIdx.x, gridDim.x);
}
template <int NP, int TX, int TY, int MODE>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_sip_batch(const SipArgs* __restrict__ tab,
                                                                  const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const SipArgs& a = tab[w.x];
  k2_sip_body<NP, TX, TY, MODE, 1>(a, w.y, tiles_of<TX, TY>(a.g));
}


// x += a p ; r -= a Ap ; z = r/M ; <r, z>   (pcg_fixed!, elliptic.jl:339-343)
// Workgroup blk of nblk processes tiles blk, blk + nblk, ...; the next
// tile's loads are issued before this tile's stores.  The first tile's loads
// come before the grid total of the previous kernel's partial sums.  Only the
// last iteration's update (XU) touches x: x += a_k p_k of the earlier
// iterations is done by the next iteration's PCG_A, which holds p_k at its
// own nodes anyway (same arithmetic; this kernel then streams 4 fields, not 7).
template <int NP, int TX, int TY, bool XU, int GU = 8>
__device__ __forceinline__ void k2_pcg_b_body(const SipArgs& a, const int blk, const int nblk) {
  constexpr int NP2 = NP * NP;
  __shared__ unsigned arrive;   // waves done with the partial sums (block_partials_store_last)
  const int ntiles = tiles_of<TX, TY>(a.g);
  const int tid = threadIdx.x, s = tid / NP2, n = tid % NP2;
  if (tid == 0) arrive = 0;     // published by grid_finish's barrier
  const int lx = s % TX, ly = s / TX;
  double* sc = a.sc;
  const int nit = a.n_iter, it = a.it;
  // the partial sums first, then this thread's fields with no branch around
  // the loads (an invalid thread reads node 0 and stores nothing)
  double pre[1];
  grid_prefetch<1>(a.part_in, a.np_in, pre);
  const double* rin = a.r ? a.r : a.rout;
  bool valid;
  size_t gi;
  double x = 0.0, p = 0.0, r, ap, im;
  auto issue = [&](int tt) {
    const TileInfo<NP, TX, TY> t(a.g, tt, ntiles);
    valid = lx < t.nvx && ly < t.nvy;
    gi = valid ? t.gidx(s, n) : 0;
    if constexpr (XU) { x = a.xout[gi]; p = a.pold[gi]; }
    r = rin[gi]; ap = a.y[gi]; im = a.invM[gi];
  };
  issue(blk);
  const double rz_it = sc[it];
  double tot[1];
  grid_finish<1, GU>(a.part_in, a.np_in, pre, tot);
  const double ak = rz_it / (tot[0] + 1e-30);
  double acc = 0.0;
  for (int tt = blk; tt < ntiles; tt += nblk) {
    const double xn = fma(ak, p, x);
    const double rn = fma(-ak, ap, r);
    const bool v = valid;
    const size_t g0 = gi;
    acc += v ? rn * (rn * im) : 0.0;
    if (tt + nblk < ntiles) issue(tt + nblk);      // uniform branch: the workgroup's next tile
    if (v) {
      if constexpr (XU) store_wt(&a.xout[g0], xn);
      store_wt(&a.rout[g0], rn);
    }
  }
  double red_v[1] = {acc};
  block_partials_store_last<1>(red_v, a.part_out, nblk, blk, &arrive);
  if (blk == 0 && tid == 0) { sc[nit + 1 + it] = tot[0]; sc[2 * nit + 1 + it] = ak; }
}
template <int NP, int TX, int TY, bool XU>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_pcg_b(SipArgs a) {
  k2_pcg_b_body<NP, TX, TY, XU>(a, blockIdx.x, gridDim.x);
}
template <int NP, int TX, int TY, bool XU>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_pcg_b_batch(const SipArgs* __restrict__ tab,
                                                                    const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const SipArgs& a = tab[w.x];
  k2_pcg_b_body<NP, TX, TY, XU, 1>(a, w.y, a.ngroups);
}


// After the last PCG iteration: rz_n and beta_{n-1} into the scalar record.
__device__ __forceinline__ void k2_pcg_final_body(const double* part, int nb, double* sc, int nit) {
  __shared__ double red[16];
  double tot[1];
  grid_total<1>(part, nb, tot, red);
  if (threadIdx.x == 0 && nit > 0) {
    sc[nit] = tot[0];
    sc[3 * nit + 1 + nit - 1] = tot[0] / (sc[nit - 1] + 1e-30);
  }
  if (threadIdx.x == 0 && nit == 0) sc[0] = tot[0];
}
__global__ void k2_pcg_final(const double* part, int nb, double* sc, int nit) { k2_pcg_final_body(part, nb, sc, nit); }
// one workgroup per instance
__global__ void k2_pcg_final_batch(const FinArgs* __restrict__ tab) {
  const FinArgs& f = tab[blockIdx.x];
  k2_pcg_final_body(f.part, f.nb, f.sc, f.nit);
}


// ------------------------------------------------------------------------
// Hyperbolic RHS + SSP-RK3 stage update.  Built without FMA contraction and
// in the reference's evaluation order, so dq equals the CPU restatement
// bit for bit.
// ------------------------------------------------------------------------
#pragma clang fp contract(off)
template <int NP, int TX, int TY>
__device__ __forceinline__ void k2_hyp_body(const HypArgs& a, const int blk, const int nblk) {
  constexpr int NP2 = NP * NP, T = TX * TY, NT = T * NP2;
  // Halo face nodes: the left and right strips (x faces) first, then the
  // bottom and top strips (y faces); the y-strip nodes start on a wave
  // boundary where the workgroup has room, so each staging wave evaluates
  // one flux direction.
  constexpr int NHX = 2 * TY * NP, NHY = 2 * TX * NP;
  constexpr int HY0A = (NHX + WAVE - 1) / WAVE * WAVE;
  constexpr int HY0 = HY0A + NHY <= NT ? HY0A : NHX;
  static_assert(NHX + NHY <= NT, "one halo face node per thread");
  // Stage state, fluxes and face wave speeds |u| + c, |v| + c of the tile's
  // nodes, followed by those of the halo face nodes (x strips in Fx and wsx,
  // y strips in Fy and wsy; all of them in qs).  Sigma enters only through
  // the fluxes, so it is not kept.
  __shared__ double qs[4][NT + NHX + NHY];
  __shared__ double Fx[4][NT + NHX], Fy[4][NT + NHY];
  __shared__ double wsx[NT + NHX], wsy[NT + NHY];
  // Face corrections (1/J)(f* - F)/w of the tile's face nodes: element s,
  // slot s 2NP + q for its left (bottom) face node q, s 2NP + NP + q for its
  // right (top) one.
  __shared__ double cx[4][2 * T * NP], cy[4][2 * T * NP];
  __shared__ double sD[NP2];

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

  // Own node and (some threads) one halo face node: both sets of loads are
  // issued, with no branch between them, before either is written to LDS
  // (an invalid thread or empty halo slot reads node 0 and takes the
  // default state instead); q^n of stages 2 and 3 comes with them (issued
  // after the staging barrier instead, where it enters, it took 256^2 from
  // 23.8 to 26.0 us: with four rounds of workgroups the entry loads already
  // overlap other workgroups' phases).
  // LDS bank swizzle of the tile's node slots at p = 3: node n of element s
  // sits at s NP2 + (n ^ 5a), a = (s / 2) mod 4, so the face pass (the same
  // face column or row of eight consecutive elements per 32 lanes, which
  // 16-double element slots put on one bank pair) and the volume term's
  // reads fall on distinct banks; the slot of a node within its element
  // changes, the values and their order of use do not.
  auto nslot = [](const int se, const int nn) -> int {
    if constexpr (NP == 4) {
      const int a = (se >> 1) & 3;
      return se * NP2 + (nn ^ (a | (a << 2)));
    } else {
      return se * NP2 + nn;
    }
  };
  const int me = nslot(s, n);
  const bool hx = tid < NHX, hy = tid >= HY0 && tid < HY0 + NHY;
  const int hidx = hx ? tid : tid - HY0 + NHX;
  size_t hgi = 0;
  const bool hok = (hx || hy) && t.halo_face_global(hidx / NP, hidx % NP, hgi);
  if (!hok) hgi = 0;
  const double* qnp = a.mode >= 2 ? a.qn : a.q;
  double lo[5], ln[4], lh[5];
#pragma unroll
  for (int v = 0; v < 4; ++v) lo[v] = a.q[v * N + gme];
  lo[4] = a.S[gme];
#pragma unroll
  for (int v = 0; v < 4; ++v) ln[v] = qnp[v * N + gme];
#pragma unroll
  for (int v = 0; v < 4; ++v) lh[v] = a.q[v * N + hgi];
  lh[4] = a.S[hgi];
  const double qdef[5] = {0.0, 0.0, 1.0, 1.0, 0.0};
  double qo[5];
  double qn[4];
#pragma unroll
  for (int v = 0; v < 5; ++v) qo[v] = valid ? lo[v] : qdef[v];
#pragma unroll
  for (int v = 0; v < 4; ++v) qn[v] = (valid && a.mode >= 2) ? ln[v] : 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) qs[v][me] = qo[v];
  {
    double fxo[4], fyo[4];
    flux_x2(gam, qo, fxo);
    flux_y2(gam, qo, fyo);
#pragma unroll
    for (int v = 0; v < 4; ++v) { Fx[v][me] = fxo[v]; Fy[v][me] = fyo[v]; }
  }
  const bool face = (i == 0 || i == NP - 1 || j == 0 || j == NP - 1);
  if (valid && face) {
    const double c = sound2(gam, qo, a.flag);
    wsx[me] = fabs(qo[0] / qo[2]) + c;
    wsy[me] = fabs(qo[1] / qo[2]) + c;
  }
  if (hx || hy) {
    double qh[5], fh[4];
#pragma unroll
    for (int v = 0; v < 5; ++v) qh[v] = hok ? lh[v] : qdef[v];
#pragma unroll
    for (int v = 0; v < 4; ++v) qs[v][NT + hidx] = qh[v];
    const double c = sound2(gam, qh, a.flag);
    if (hx) {
      flux_x2(gam, qh, fh);
#pragma unroll
      for (int v = 0; v < 4; ++v) Fx[v][NT + hidx] = fh[v];
      wsx[NT + hidx] = fabs(qh[0] / qh[2]) + c;
    } else {
      flux_y2(gam, qh, fh);
#pragma unroll
      for (int v = 0; v < 4; ++v) Fy[v][NT + hidx - NHX] = fh[v];
      wsy[NT + hidx - NHX] = fabs(qh[1] / qh[2]) + c;
    }
  }
  stage_basis<NP>(B, sD, nullptr);
  __syncthreads();

  // volume term: dq = -(1/Jx) Dx Fx - (1/Jy) Dy Fy, x then y per k (hyperbolic.jl:55-66)
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (valid) {
#pragma unroll
    for (int k = 0; k < NP; ++k) {
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] -= g.invJx * sD[i * NP + k] * Fx[v][nslot(s, k + NP * j)];
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] -= g.invJy * sD[j * NP + k] * Fy[v][nslot(s, i + NP * k)];
    }
  }

  // Face terms, one LLF flux per face (hyperbolic.jl:75-116).  The tile's
  // faces are spread over the workgroup face-major: the right face of every
  // element (T NP items), the top face of every element (T NP), the left
  // face of the first column (TY NP) and the bottom face of the first row
  // (TX NP), each range starting on a wave boundary where the workgroup has
  // room, so a wave evaluates faces of one direction and one kind (a
  // per-node mix of the four faces left three quarters of every wave idle
  // through the fluxes and divisions).  An item writes the correction
  // (1/J)(f* - F)/w of each of its two nodes that lies in the tile; the
  // node's owner applies it after the barrier.
  {
    constexpr int NXR = T * NP, NYT = T * NP, NXL = TY * NP, NYB = TX * NP;
    constexpr int RU1 = (NXR + WAVE - 1) / WAVE * WAVE, RU2 = (NYT + WAVE - 1) / WAVE * WAVE;
    constexpr int RU3 = (NXL + WAVE - 1) / WAVE * WAVE;
    constexpr bool AL = RU1 + RU2 + RU3 + NYB <= NT;
    constexpr int O1 = AL ? RU1 : NXR, O2 = O1 + (AL ? RU2 : NYT), O3 = O2 + (AL ? RU3 : NXL);
    constexpr int OE = O3 + NYB;
    // one face: L / R are qs slots, fL / fR the flux and wave-speed slots,
    // cL / cR the correction slots (-1: the node is a halo node)
    auto face_item = [&](auto Dc, const int L, const int R, const int fL_, const int fR_, const int cL,
                         const int cR) {
      constexpr int D = decltype(Dc)::value;
      double qL[4], qR[4], fL[4], fR[4], wL, wR;
#pragma unroll
      for (int v = 0; v < 4; ++v) { qL[v] = qs[v][L]; qR[v] = qs[v][R]; }
      if constexpr (D == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) { fL[v] = Fx[v][fL_]; fR[v] = Fx[v][fR_]; }
        wL = wsx[fL_]; wR = wsx[fR_];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) { fL[v] = Fy[v][fL_]; fR[v] = Fy[v][fR_]; }
        wL = wsy[fL_]; wR = wsy[fR_];
      }
      double f[4];
      llf_combine(fL, fR, qL, qR, fmax_first(wL, wR), f);
      const double inv = D == 0 ? g.invJx : g.invJy;
      double (*cc)[2 * T * NP] = D == 0 ? cx : cy;
      if (cL >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) cc[v][cL] = inv * (f[v] - fL[v]) / B.w[NP - 1];
      }
      if (cR >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) cc[v][cR] = inv * (f[v] - fR[v]) / B.w[0];
      }
    };
#pragma unroll
    for (int w0 = 0; w0 < OE; w0 += NT) {
      const int w = w0 + tid;
      if (w < NXR) {                                   // right face of element e, node q along it
        const int e = w / NP, q = w % NP;
        if (e % TX < t.nvx && e / TX < t.nvy) {
          const int L = nslot(e, (NP - 1) + NP * q);
          const int eR = t.nb_right(e);
          const int R = eR < T ? nslot(eR, NP * q) : NT + (eR - T) * NP + q;
          face_item(std::integral_constant<int, 0>{}, L, R, L, R, e * 2 * NP + NP + q,
                    eR < T ? eR * 2 * NP + q : -1);
        }
      } else if (w >= O1 && w < O1 + NYT) {            // top face
        const int e = (w - O1) / NP, q = (w - O1) % NP;
        if (e % TX < t.nvx && e / TX < t.nvy) {
          const int L = nslot(e, q + NP * (NP - 1));
          const int eT = t.nb_top(e);
          const int h = (eT - T) * NP + q;              // halo slot when eT >= T
          const int Rt = eT < T ? nslot(eT, q) : 0;
          face_item(std::integral_constant<int, 1>{}, L, eT < T ? Rt : NT + h, L,
                    eT < T ? Rt : NT + h - NHX, e * 2 * NP + NP + q, eT < T ? eT * 2 * NP + q : -1);
        }
      } else if (w >= O2 && w < O2 + NXL) {            // left face of the first column
        const int r = (w - O2) / NP, q = (w - O2) % NP;
        if (r < t.nvy) {
          const int e = r * TX;
          const int R = nslot(e, NP * q);
          const int L = NT + r * NP + q;                  // left strip, slot r
          face_item(std::integral_constant<int, 0>{}, L, R, L, R, -1, e * 2 * NP + q);
        }
      } else if (w >= O3 && w < OE) {                  // bottom face of the first row
        const int e = (w - O3) / NP, q = (w - O3) % NP;
        if (e < t.nvx) {
          const int R = nslot(e, q);
          const int h = (2 * TY + e) * NP + q;            // bottom strip, slot e
          face_item(std::integral_constant<int, 1>{}, NT + h, R, NT + h - NHX, R, -1, e * 2 * NP + q);
        }
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  // face corrections: x faces, then y faces (hyperbolic.jl:75-116)
  const int xsl = s * 2 * NP + (i == 0 ? 0 : NP) + j;
  const int ysl = s * 2 * NP + (j == 0 ? 0 : NP) + i;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double tx = cx[v][xsl], ty = cy[v][ysl];
    double r = acc[v];
    r = i == NP - 1 ? r - tx : (i == 0 ? r + tx : r);
    r = j == NP - 1 ? r - ty : (j == 0 ? r + ty : r);
    acc[v] = r;
  }
  // SSP-RK3 stage combination (IGR_adjoints_2d/forward.jl:53-68)
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const size_t gi = v * N + gme;
    double o;
    if (a.mode == 0) o = acc[v];
    else if (a.mode == 1) o = qo[v] + a.cdt * acc[v];
    else if (a.mode == 2) o = 3.0 / 4.0 * qn[v] + 1.0 / 4.0 * qo[v] + a.cdt * acc[v];
    else o = 1.0 / 3.0 * qn[v] + 2.0 / 3.0 * qo[v] + a.cdt * acc[v];
    store_wt(&a.out[gi], o);
  }
}
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_hyp(HypArgs a) {
  k2_hyp_body<NP, TX, TY>(a, blockIdx.x, gridDim.x);
}
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_hyp_batch(const HypArgs* __restrict__ tab,
                                                                  const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const HypArgs& a = tab[w.x];
  k2_hyp_body<NP, TX, TY>(a, w.y, tiles_of<TX, TY>(a.g));
}

#pragma clang fp contract(fast)

// Plain fill (used instead of hipMemsetAsync so the step can be captured in
// a graph).
__global__ void k_fill(double* x, size_t n, double v) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = v;
}
__global__ void k_copy(double* dst, const double* src, size_t n) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}
// One fill / copy per blockIdx.y, grid-stride over x.
__global__ void k_fill_batch(const FillArgs* __restrict__ tab) {
  const FillArgs& f = tab[blockIdx.y];
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < f.n; k += (size_t)gridDim.x * blockDim.x)
    f.x[k] = f.v;
}
__global__ void k_copy_batch(const CopyArgs* __restrict__ tab) {
  const CopyArgs& f = tab[blockIdx.y];
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < f.n; k += (size_t)gridDim.x * blockDim.x)
    f.dst[k] = f.src[k];
}

// ------------------------------------------------------------------------
// Host launchers
// ------------------------------------------------------------------------
int grid2_blocks(const Ctx& c);
int grid2_bblocks(const Ctx& c);

// Kernel-timestamp events for the next SIP/hyperbolic launch (set by
// igr_time_kernel only): hipExtLaunchKernelGGL stamps them at the start and
// end of the kernel itself, so their elapsed time excludes launch gaps.
static thread_local hipEvent_t g_kt0 = nullptr, g_kt1 = nullptr;
void set_kernel_timer(hipEvent_t t0, hipEvent_t t1) { g_kt0 = t0; g_kt1 = t1; }

template <typename... Args, typename F = void (*)(Args...)>
static void launch_k(F k, dim3 G, dim3 B, hipStream_t st, Args... args) {
  if (g_kt0) {
    hipExtLaunchKernelGGL(k, G, B, 0, st, g_kt0, g_kt1, 0, args...);
    g_kt0 = g_kt1 = nullptr;
  } else {
    hipLaunchKernelGGL(k, G, B, 0, st, args...);
  }
}

template <int NP, int TX, int TY>
struct Launch2 {
  static dim3 grid(const Geo2& g) {
    int ntx = (g.Nex + TX - 1) / TX, nty = (g.Ney + TY - 1) / TY;
    return dim3(ntx * nty);
  }
  static dim3 block() { return dim3(TX * TY * NP * NP); }
  static void sip(int mode, const SipArgs& a, hipStream_t st) {
    dim3 G(a.ngroups), Bk = block();
    switch (mode) {
      case M_APPLY: launch_k(k2_sip<NP, TX, TY, M_APPLY>, G, Bk, st, a); break;
      case M_INIT_PCG: launch_k(k2_sip<NP, TX, TY, M_INIT_PCG>, G, Bk, st, a); break;
      case M_INIT_CHEB: launch_k(k2_sip<NP, TX, TY, M_INIT_CHEB>, G, Bk, st, a); break;
      case M_INIT_JAC: launch_k(k2_sip<NP, TX, TY, M_INIT_JAC>, G, Bk, st, a); break;
      case M_PCG_A: launch_k(k2_sip<NP, TX, TY, M_PCG_A>, G, Bk, st, a); break;
      case M_JACOBI: launch_k(k2_sip<NP, TX, TY, M_JACOBI>, G, Bk, st, a); break;
      case M_CHEB: launch_k(k2_sip<NP, TX, TY, M_CHEB>, G, Bk, st, a); break;
      case M_PCG_B: launch_k(k2_pcg_b<NP, TX, TY, false>, G, Bk, st, a); break;
      case M_PCG_BX: launch_k(k2_pcg_b<NP, TX, TY, true>, G, Bk, st, a); break;
    }
  }
  static void hyp(const HypArgs& a, hipStream_t st) {
    launch_k(k2_hyp<NP, TX, TY>, grid(a.g), block(), st, a);
  }
  static void sip_batch(int mode, const SipArgs* tab, const int2* work, int nb, hipStream_t st) {
    const dim3 G(nb), Bk = block();
    switch (mode) {
      case M_APPLY: hipLaunchKernelGGL((k2_sip_batch<NP, TX, TY, M_APPLY>), G, Bk, 0, st, tab, work); break;
      case M_INIT_PCG: hipLaunchKernelGGL((k2_sip_batch<NP, TX, TY, M_INIT_PCG>), G, Bk, 0, st, tab, work); break;
      case M_INIT_CHEB: hipLaunchKernelGGL((k2_sip_batch<NP, TX, TY, M_INIT_CHEB>), G, Bk, 0, st, tab, work); break;
      case M_INIT_JAC: hipLaunchKernelGGL((k2_sip_batch<NP, TX, TY, M_INIT_JAC>), G, Bk, 0, st, tab, work); break;
      case M_PCG_A: hipLaunchKernelGGL((k2_sip_batch<NP, TX, TY, M_PCG_A>), G, Bk, 0, st, tab, work); break;
      case M_JACOBI: hipLaunchKernelGGL((k2_sip_batch<NP, TX, TY, M_JACOBI>), G, Bk, 0, st, tab, work); break;
      case M_CHEB: hipLaunchKernelGGL((k2_sip_batch<NP, TX, TY, M_CHEB>), G, Bk, 0, st, tab, work); break;
      case M_PCG_B: hipLaunchKernelGGL((k2_pcg_b_batch<NP, TX, TY, false>), G, Bk, 0, st, tab, work); break;
      case M_PCG_BX: hipLaunchKernelGGL((k2_pcg_b_batch<NP, TX, TY, true>), G, Bk, 0, st, tab, work); break;
    }
  }
  static void hyp_batch(const HypArgs* tab, const int2* work, int nb, hipStream_t st) {
    hipLaunchKernelGGL((k2_hyp_batch<NP, TX, TY>), dim3(nb), block(), 0, st, tab, work);
  }
};

// (np, tile) -> the instantiated tile shape.  p = 3 (np = 4): the SIP /
// solver family has 4x4, 8x4, 8x8 and 2x2 tiles, the hyperbolic family
// 4x4, 8x4 and 2x2 (an 8x8 hyperbolic tile would not fit the LDS).  p = 4
// (np = 5, 25 nodes per element): the SIP family has 2x4 (200 threads),
// 4x4 (400), 5x5 (625) and 8x5 (1000) tiles, the others 2x4.
#define IGR_SIP_TILE_DISPATCH(L, np, tx, ty, CALL)                      \
  switch (np) {                                                         \
    case 2: L<2, 8, 8>::CALL; break;                                    \
    case 3: L<3, 4, 4>::CALL; break;                                    \
    case 4:                                                             \
      if ((tx) == 8 && (ty) == 4) L<4, 8, 4>::CALL;                     \
      else if ((tx) == 8 && (ty) == 8) L<4, 8, 8>::CALL;                \
      else if ((tx) == 2 && (ty) == 2) L<4, 2, 2>::CALL;                \
      else L<4, 4, 4>::CALL;                                            \
      break;                                                            \
    case 5:                                                             \
      if ((tx) == 8 && (ty) == 5) L<5, 8, 5>::CALL;                     \
      else if ((tx) == 5 && (ty) == 5) L<5, 5, 5>::CALL;                \
      else if ((tx) == 4 && (ty) == 4) L<5, 4, 4>::CALL;                \
      else L<5, 2, 4>::CALL;                                            \
      break;                                                            \
  }
#define IGR_HYP_TILE_DISPATCH(L, np, tx, ty, CALL)                      \
  switch (np) {                                                         \
    case 2: L<2, 8, 8>::CALL; break;                                    \
    case 3: L<3, 4, 4>::CALL; break;                                    \
    case 4:                                                             \
      if ((tx) == 8 && (ty) == 4) L<4, 8, 4>::CALL;                     \
      else if ((tx) == 2 && (ty) == 2) L<4, 2, 2>::CALL;                \
      else L<4, 4, 4>::CALL;                                            \
      break;                                                            \
    case 5: L<5, 2, 4>::CALL; break;                                    \
  }

void launch2_sip_batch(int np, int tx, int ty, int mode, const SipArgs* tab, const int2* work, int nblocks,
                       hipStream_t st) {
  IGR_SIP_TILE_DISPATCH(Launch2, np, tx, ty, sip_batch(mode, tab, work, nblocks, st));
}
void launch2_hyp_batch(int np, int tx, int ty, const HypArgs* tab, const int2* work, int nblocks, hipStream_t st) {
  IGR_HYP_TILE_DISPATCH(Launch2, np, tx, ty, hyp_batch(tab, work, nblocks, st));
}
void launch_pcg_final_batch(const FinArgs* tab, int n, hipStream_t st) {
  hipLaunchKernelGGL(k2_pcg_final_batch, dim3(n), dim3(256), 0, st, tab);
}
void launch_fill_batch(const FillArgs* tab, int n, size_t maxn, hipStream_t st) {
  const unsigned gx = (unsigned)std::min<size_t>((maxn + 255) / 256, 2048);
  hipLaunchKernelGGL(k_fill_batch, dim3(std::max(gx, 1u), n), dim3(256), 0, st, tab);
}
void launch_copy_batch(const CopyArgs* tab, int n, size_t maxn, hipStream_t st) {
  const unsigned gx = (unsigned)std::min<size_t>((maxn + 255) / 256, 2048);
  hipLaunchKernelGGL(k_copy_batch, dim3(std::max(gx, 1u), n), dim3(256), 0, st, tab);
}

// The x/r update of PCG (PCG_B) runs grid2_bblocks(c) workgroups, every
// other SIP-family kernel one per tile (grid2_blocks(c)); each kernel is
// told how many partial sums the kernel before it wrote.
void launch2_sip(const Ctx& c, int mode, const SipArgs& a0) {
  SipArgs a = a0;
  const int nt = grid2_blocks(c), nbb = grid2_bblocks(c);
  const bool pb = mode == M_PCG_B || mode == M_PCG_BX;
  a.ngroups = pb ? nbb : nt;
  a.np_in = pb ? nt : (mode == M_PCG_A ? (a.first ? nt : nbb) : 0);
  if (g_rec) { g_rec->recs.push_back({R_SIP, mode, g_rec->sip.size()}); g_rec->sip.push_back(a); return; }
  IGR_SIP_TILE_DISPATCH(Launch2, c.np, c.tile_x, c.tile_y, sip(mode, a, c.stream));
}
void launch2_hyp(const Ctx& c, const HypArgs& a) {
  if (g_rec) { g_rec->recs.push_back({R_HYP, 0, g_rec->hyp.size()}); g_rec->hyp.push_back(a); return; }
  IGR_HYP_TILE_DISPATCH(Launch2, c.np, c.htile_x, c.htile_y, hyp(a, c.stream));
}
// workgroups of one hyperbolic-family launch
int grid2_hblocks(const Ctx& c) {
  int tx = 4, ty = 4;
  if (c.np == 2) { tx = 8; ty = 8; }
  else if (c.np == 5) { tx = 2; ty = 4; }
  else if (c.np == 4) { tx = c.htile_x; ty = c.htile_y; }
  return ((c.Nex + tx - 1) / tx) * ((c.Ney + ty - 1) / ty);
}
// workgroups of one SIP-family launch (the number of partial sums it
// writes): one per tile
int grid2_blocks(const Ctx& c) {
  int tx = 4, ty = 4;
  if (c.np == 2) { tx = 8; ty = 8; }
  else if (c.np == 4 || c.np == 5) { tx = c.tile_x; ty = c.tile_y; }
  return ((c.Nex + tx - 1) / tx) * ((c.Ney + ty - 1) / ty);
}
// workgroups of the PCG x/r update: at most c.bgroups (0: one per tile),
// each running every n-th tile with the next tile's loads issued before the
// current tile's stores
int grid2_bblocks(const Ctx& c) {
  const int nt = grid2_blocks(c);
  return c.bgroups > 0 ? std::min(nt, c.bgroups) : nt;
}
// workgroups of one reverse SIP-family launch
int grid2_rblocks(const Ctx& c) {
  int tx = 4, ty = 4;
  if (c.np == 2) { tx = 8; ty = 8; }
  else if (c.np == 5) { tx = 2; ty = 4; }
  else if (c.np == 4) { tx = c.rtile_x; ty = c.rtile_y; }
  return ((c.Nex + tx - 1) / tx) * ((c.Ney + ty - 1) / ty);
}
// workgroups of the reverse PCG iteration and its start kernel: at most
// c.rgroups (0: one per reverse tile), each running every n-th tile
int grid2_rpgroups(const Ctx& c) {
  const int nt = grid2_rblocks(c);
  return c.rgroups > 0 ? std::min(nt, c.rgroups) : nt;
}
// the partial sums are the last PCG_B's (n_iter > 0) or the INIT kernel's
void launch_pcg_final(const Ctx& c, const double* part, double* sc, int nit) {
  const int nb = nit > 0 ? grid2_bblocks(c) : grid2_blocks(c);
  if (g_rec) { g_rec->recs.push_back({R_FINAL, 0, g_rec->fin.size()}); g_rec->fin.push_back({part, nb, sc, nit}); return; }
  hipLaunchKernelGGL(k2_pcg_final, dim3(1), dim3(256), 0, c.stream, part, nb, sc, nit);
}
void launch_fill(const Ctx& c, double* x, size_t n, double v) {
  if (g_rec) { g_rec->recs.push_back({R_FILL, 0, g_rec->fill.size()}); g_rec->fill.push_back({x, n, v}); return; }
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, x, n, v);
}
void launch_copy(const Ctx& c, double* dst, const double* src, size_t n) {
  if (g_rec) { g_rec->recs.push_back({R_COPY, 0, g_rec->copy.size()}); g_rec->copy.push_back({dst, src, n}); return; }
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, dst, src, n);
}

}  // namespace igr
