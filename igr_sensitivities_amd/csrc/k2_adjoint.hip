// 2D discrete adjoint kernels (gfx950, FP64): hand-written reverse mode of
// the forward map, replacing the Enzyme reverse sweep the reference uses
// (IGR_adjoints_2d/enzyme_adjoint.jl:270-296).
//
//   k2_hyp_rev      reverse of the hyperbolic RHS + SSP-RK3 stage update
//                   (hyperbolic.jl:27-119, euler.jl:20-133, forward.jl:51-69)
//   k2_sip_rev      reverse of the Sigma solve, one launch per PCG iteration
//                   (pcg_fixed!, elliptic.jl:329-347): rebuilds the adjoint
//                   of A p_k on the fly for own and halo nodes, applies A
//                   (A is symmetric, so A^T = A), accumulates d/d(1/rho) of
//                   <w, A(rho) v>, updates the adjoint direction and computes
//                   the four dot products the next reverse iteration needs in
//                   one grid reduction.  FIN reverses r0 = b - A x0, the
//                   diagonal (compute_sip_diagonal, :205-264) and the IGR
//                   source (compute_igr_rhs!, :159-197).
//   k2_objective_*  objective value (grid reduction) and its gradient
//                   (enzyme_adjoint.jl:72-76, 108-113, 145-151, 184-190, 223-228)
// Work decomposition, tiles, halo and XCD-contiguous order as in k2_forward.
#include "igr_ctx.hpp"
#include <type_traits>
#include "k2_adjoint_common.hpp"

namespace igr {

// ------------------------------------------------------------------------
// hyperbolic + stage reverse
// ------------------------------------------------------------------------
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_hyp_rev(HypRevArgs a) {
  constexpr int NP2 = NP * NP, T = TX * TY, NH = 2 * (TX + TY), NT = T * NP2;
  __shared__ double qs[5][T * NP2];
  __shared__ double db[4][T * NP2];
  __shared__ double hq[4][NH * NP];
  __shared__ double hdb[4][NH * NP];
  __shared__ double red[NT / WAVE + 1];
  __shared__ double sD[NP * NP];     // basis (stage_basis)

  const Geo2 g = a.g;
  const BasisK& B = a.B;
  const double gam = a.gamma;
  TileInfo<NP, TX, TY> t(g);
  const int tid = threadIdx.x;
  const int s = tid / NP2, n = tid % NP2, i = n % NP, j = n / NP;
  const int lx = s % TX, ly = s / TX;
  const bool valid = lx < t.nvx && ly < t.nvy;
  const size_t N = (size_t)g.Nex * g.Ney * NP2;
  const size_t gme = valid ? t.gidx(s, n) : 0;

  double qo[5] = {0.0, 0.0, 1.0, 1.0, 0.0}, dbo[4] = {0.0, 0.0, 0.0, 0.0}, qbo[4] = {0.0, 0.0, 0.0, 0.0};
  // the accumulators this thread updates at the end, loaded now so their
  // latency overlaps the state loads (qbar_stage, qbar_n, sbar_in and
  // sbar_out are distinct buffers in every launch)
  double pf_stage[4] = {0.0, 0.0, 0.0, 0.0}, pf_n[4] = {0.0, 0.0, 0.0, 0.0}, pf_sb = 0.0;
  if (valid) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qo[v] = a.q[v * N + gme];
      qbo[v] = a.qbar_out[v * N + gme];
      dbo[v] = a.c_dt * qbo[v];
      if (a.accumulate) pf_stage[v] = a.qbar_stage[v * N + gme];
      if (a.qbar_n) pf_n[v] = a.qbar_n[v * N + gme];
    }
    qo[4] = a.S[gme];
    if (a.sbar_in) pf_sb = a.sbar_in[gme];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) qs[v][tid] = qo[v];
#pragma unroll
  for (int v = 0; v < 4; ++v) db[v][tid] = dbo[v];
  for (int idx = tid; idx < NH * NP; idx += NT) {
    const int h = idx / NP, k = idx % NP;
    size_t gi;
    if (t.halo_face_global(h, k, gi)) {
#pragma unroll
      for (int v = 0; v < 4; ++v) { hq[v][idx] = a.q[v * N + gi]; hdb[v][idx] = a.c_dt * a.qbar_out[v * N + gi]; }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) { hq[v][idx] = (v >= 2) ? 1.0 : 0.0; hdb[v][idx] = 0.0; }
    }
  }
  stage_basis<NP>(B, sD, nullptr);
  __syncthreads();

  double gb = 0.0;
  if (valid) {
    double Fxb[4], Fyb[4], qb[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double sx = 0.0, sy = 0.0;
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        sx += sD[k * NP + i] * db[v][s * NP2 + k + NP * j];
        sy += sD[k * NP + j] * db[v][s * NP2 + i + NP * k];
      }
      Fxb[v] = -g.invJx * sx;
      Fyb[v] = -g.invJy * sy;
    }
    double lxb = 0.0, lyb = 0.0;
    const double wl = 1.0 / B.w[NP - 1], w0 = 1.0 / B.w[0];
    // neighbour face node (tile LDS or halo) -> q and D-bar
    auto nb = [&](int sl, int node, int hk, double qn[4], double dn[4]) {
      if (sl < T) {
#pragma unroll
        for (int v = 0; v < 4; ++v) { qn[v] = qs[v][sl * NP2 + node]; dn[v] = db[v][sl * NP2 + node]; }
      } else {
        const int h = sl - T;
#pragma unroll
        for (int v = 0; v < 4; ++v) { qn[v] = hq[v][h * NP + hk]; dn[v] = hdb[v][h * NP + hk]; }
      }
    };
    // |u| + c by the forward kernel's expression on both sides of a face: the
    // reverse must break a tie of the two speeds the way the forward did
    auto wave = [&](const double qn[4], int dir) {
      const double Pn = (gam - 1.0) * (qn[3] - (qn[0] * qn[0] + qn[1] * qn[1]) / (2.0 * qn[2]));
      return fabs(qn[dir] / qn[2]) + sqrt(gam * Pn / qn[2]);
    };
    // x faces
    if (i == NP - 1 || i == 0) {
      const bool own_is_L = (i == NP - 1);
      double qn[4], dn[4];
      if (own_is_L) nb(t.nb_right(s), NP * j, j, qn, dn);
      else nb(t.nb_left(s), (NP - 1) + NP * j, j, qn, dn);
      const double lo = wave(qo, 0), ln = wave(qn, 0);
      const double lamL = own_is_L ? lo : ln, lamR = own_is_L ? ln : lo;
      const double lam = fmax_first(lamL, lamR);
      double lb = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double dL = own_is_L ? dbo[v] : dn[v], dR = own_is_L ? dn[v] : dbo[v];
        const double fb = -g.invJx * wl * dL + g.invJx * w0 * dR;
        const double qL = own_is_L ? qo[v] : qn[v], qR = own_is_L ? qn[v] : qo[v];
        lb += fb * (qL - qR) / 2.0;
        if (own_is_L) { Fxb[v] += g.invJx * wl * dbo[v] + fb / 2.0; qb[v] += lam * fb / 2.0; }
        else { Fxb[v] += -g.invJx * w0 * dbo[v] + fb / 2.0; qb[v] -= lam * fb / 2.0; }
      }
      const bool L_chosen = !(lamR > lamL);
      if (own_is_L == L_chosen) lxb += lb;
    }
    // y faces
    if (j == NP - 1 || j == 0) {
      const bool own_is_L = (j == NP - 1);
      double qn[4], dn[4];
      if (own_is_L) nb(t.nb_top(s), i, i, qn, dn);
      else nb(t.nb_bottom(s), i + NP * (NP - 1), i, qn, dn);
      const double lo = wave(qo, 1), ln = wave(qn, 1);
      const double lamL = own_is_L ? lo : ln, lamR = own_is_L ? ln : lo;
      const double lam = fmax_first(lamL, lamR);
      double lb = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double dL = own_is_L ? dbo[v] : dn[v], dR = own_is_L ? dn[v] : dbo[v];
        const double fb = -g.invJy * wl * dL + g.invJy * w0 * dR;
        const double qL = own_is_L ? qo[v] : qn[v], qR = own_is_L ? qn[v] : qo[v];
        lb += fb * (qL - qR) / 2.0;
        if (own_is_L) { Fyb[v] += g.invJy * wl * dbo[v] + fb / 2.0; qb[v] += lam * fb / 2.0; }
        else { Fyb[v] += -g.invJy * w0 * dbo[v] + fb / 2.0; qb[v] -= lam * fb / 2.0; }
      }
      const bool L_chosen = !(lamR > lamL);
      if (own_is_L == L_chosen) lyb += lb;
    }
    double sb = 0.0;
    euler_node_rev(gam, qo, Fxb, Fyb, lxb, lyb, qb, sb, gb);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const size_t gi = v * N + gme;
      const double val = a.c_prev * qbo[v] + qb[v];
      store_wt(&a.qbar_stage[gi], a.accumulate ? pf_stage[v] + val : val);
      if (a.qbar_n) store_wt(&a.qbar_n[gi], pf_n[v] + a.c_n * qbo[v]);
    }
    store_wt(&a.sbar_out[gme], pf_sb + sb);
  }
  double rv[1] = {gb};
  block_sum<1>(rv, red);
  if (tid == 0) a.gbar_blk[blockIdx.x] += rv[0];
}

// ------------------------------------------------------------------------
// Solve reverse (PCG) and the standalone SIP VJP
// ------------------------------------------------------------------------
// State a reverse PCG iteration workgroup carries from tile to tile: the
// reverse scalars (formed once, from the previous kernel's partial sums) and
// its running sums of the four dot products and of alpha-bar.
struct RevLoop {
  double ak = 0.0, bk = 0.0, rzb = 0.0, papb = 0.0;
  double r_bb = 0.0, r_abk = 0.0, r_carry = 0.0;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
};

// One tile of the reverse SIP family: tile `tile` of `ntiles` (a single-tile
// workgroup passes its block index and the grid size).
template <int NP, int TX, int TY, int MODE, class Pre>
__device__ __forceinline__ void k2_sip_rev_body(const RevArgs& a, const int tile, const int ntiles, RevLoop& L,
                                                Pre&& scalars) {
  constexpr int NP2 = NP * NP, T = TX * TY, NS = T + 2 * (TX + TY), NT = T * NP2;
  constexpr bool FIN = (MODE == RV_PCG_FIN || MODE == RV_JAC_FIN || MODE == RV_CHEB_FIN);
  // LDS bank layout as in the forward SIP family: element slots SS doubles
  // apart (odd), face terms face-major in rows of FR doubles
  constexpr int SS = NP2 % 2 == 0 ? NP2 + 1 : NP2, FR = T + 4;
  __shared__ double ws[NS * SS];   // adjoint vector w (A is applied to it)
  __shared__ double vs[NS * SS];   // vector v of the forward product A v
  __shared__ double irs[NS * SS];
  __shared__ double gx[T * NP2], gy[T * NP2];
  __shared__ double fC[4 * NP * FR], fS[4 * NP * FR], fR[4 * NP * FR];   // face-node terms
  __shared__ double mbs[FIN ? NS * SS : 1];                 // M-bar
  __shared__ double us[FIN ? T * NP2 : 1], vvs[FIN ? T * NP2 : 1];
  __shared__ double Gxx[FIN ? T * NP2 : 1], Gxy[FIN ? T * NP2 : 1], Gyx[FIN ? T * NP2 : 1], Gyy[FIN ? T * NP2 : 1];
  __shared__ double red[4 * (NT / WAVE + 1)];

  const Geo2 g = a.g;
  const BasisK& B = a.B;
  const double al = a.alpha;
  TileInfo<NP, TX, TY> t(g, tile, ntiles);
  const int tid = threadIdx.x;
  const int s = tid / NP2, n = tid % NP2, i = n % NP, j = n / NP;
  const int lx = s % TX, ly = s / TX;
  const bool valid = lx < t.nvx && ly < t.nvy;
  const size_t N = (size_t)g.Nex * g.Ney * NP2;
  const size_t gme = valid ? t.gidx(s, n) : 0;
  const FS fs(a.n_iter);
  const int k = a.k;
  // basis in LDS (stage_basis), read per lane below
  __shared__ double sD[NP * NP], sW[NP];

  // reverse scalars of this iteration, from the dot products the previous
  // kernel left as partial sums (every workgroup forms the same totals)
  double ak = 0.0, bk = 0.0, rzb = 0.0, papb = 0.0;
  // PCG_ITER: the previous kernel's partial sums are requested first; then
  // the tile + halo fields and this thread's own-node inputs, every address
  // first and the loads with no branch between them (an empty slot reads
  // node 0 and is zeroed afterwards), so all are in flight together

  constexpr int NL = (NS * NP2 + NT - 1) / NT;
  double l_ir[NL], l_iM[NL], l_rc[NL], l_pb[NL], l_R[NL], l_p[NL];
  bool oks[NL];
  double f_irb = 0.0, f_imb = 0.0, f_xb = 0.0, f_apk = 0.0, f_pm = 0.0, f_apm = 0.0, f_Rk = 0.0;
  if constexpr (MODE == RV_PCG_ITER) {
    size_t gis[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      size_t gi = 0;
      oks[l] = t.round_global(l, s, n, gi);
      gis[l] = gi;
    }
    // r-bar and p-bar of the next iteration do not exist in a first reverse
    // iteration: valid stand-ins are read and the values dropped
    const double* rcin = a.first ? a.invR : a.rbar_in;
    const double* pbin = a.first ? a.invR : a.pbar_in;
    const double* Rk1 = a.R + (size_t)(k + 1) * N;
    const double* Pk = a.P + (size_t)k * N;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const size_t gi = gis[l];
      l_ir[l] = a.invR[gi]; l_iM[l] = a.invM[gi]; l_rc[l] = rcin[gi]; l_pb[l] = pbin[gi];
      l_R[l] = Rk1[gi]; l_p[l] = Pk[gi];
    }
    // the reverse scalars (the previous kernel's grid totals) while the
    // loads are in flight; an empty slot's values are dropped at staging
    scalars(L);
  }
  if constexpr (MODE == RV_PCG_ITER) {
    ak = L.ak; bk = L.bk; rzb = L.rzb; papb = L.papb;
  } else if constexpr (FIN) {
    rzb = a.n_iter > 0 ? a.rs[RS::carry(0)] : 0.0;
  }
  const bool have_carry = (MODE == RV_PCG_ITER || MODE == RV_CHEB_ITER) ? !a.first : (a.n_iter > 0);

  // own-node values kept for the outputs
  double o_rb = 0.0, o_zb = 0.0, o_R = 0.0, o_pb = 0.0;
  if constexpr (MODE == RV_PCG_ITER) {
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const int idx = tid + l * NT;
      if (idx < NS * NP2) {
        const bool ok = oks[l], okc = ok && !a.first;
        const double R = ok ? l_R[l] : 0.0, iM = ok ? l_iM[l] : 0.0, p = ok ? l_p[l] : 0.0;
        const double ir = ok ? l_ir[l] : 0.0, rc = okc ? l_rc[l] : 0.0, pb = okc ? l_pb[l] : 0.0;
        const double z = R * iM;
        const double zb = pb + rzb * R;
        const double rb = rc + rzb * z + zb * iM;
        const int o = (idx / NP2) * SS + idx % NP2;
        ws[o] = -ak * rb + papb * p;
        vs[o] = p;
        irs[o] = ir;
        if (l == 0) { o_rb = rb; o_zb = zb; o_R = R; o_pb = pb; }
      }
    }
  }
  for (int idx = tid; MODE != RV_PCG_ITER && idx < NS * NP2; idx += NT) {
    const int sl = idx / NP2, nn = idx % NP2;
    double w = 0.0, v = 0.0, ir = 0.0, mb = 0.0;
    size_t gi;
    if (t.slot_global(sl, nn, gi)) {
      if constexpr (MODE == RV_VJP) {
        w = a.wbar[gi]; v = a.v[gi]; ir = a.invR[gi];
      } else if constexpr (MODE == RV_JAC_ITER) {
        // x_{k+1} = x_k + omega (b - A x_k) / M: A is applied to -omega x-bar_{k+1} / M
        ir = a.invR[gi];
        const double xb1 = a.xbar[gi];
        w = -(a.omega * a.invM[gi]) * xb1;
        v = a.X[(size_t)k * N + gi];
        if (idx == tid) o_pb = xb1;
      } else if constexpr (MODE == RV_CHEB_ITER) {
        // r_k = b - A x_k: A is applied to -r-bar_k
        ir = a.invR[gi];
        const double rb = have_carry ? a.rbar_in[gi] : 0.0;
        w = -rb;
        v = a.X[(size_t)(k - 1) * N + gi];
        if (idx == tid) o_rb = rb;
      } else if constexpr (MODE == RV_JAC_FIN || MODE == RV_CHEB_FIN) {
        ir = a.invR[gi];
        const double iM = a.invM[gi];
        const double rb = (MODE == RV_CHEB_FIN && have_carry) ? a.rbar_in[gi] : 0.0;
        w = -rb;
        v = a.x0[gi];
        mb = -a.imbar[gi] * iM * iM;
        if (sl < T) { us[idx] = a.q[gi] / a.q[2 * N + gi]; vvs[idx] = a.q[N + gi] / a.q[2 * N + gi]; }
      } else {
        ir = a.invR[gi];
        const double iM = a.invM[gi];
        const double rc = have_carry ? a.rbar_in[gi] : 0.0;
        const double pb = have_carry ? a.pbar_in[gi] : 0.0;
        const int kr = (MODE == RV_PCG_ITER) ? k + 1 : 0;
        const double Rk = a.R[(size_t)kr * N + gi];
        const double z = Rk * iM;
        const double zb = pb + rzb * Rk;
        const double rb = rc + rzb * z + zb * iM;
        if constexpr (MODE == RV_PCG_ITER) {
          const double pk = a.P[(size_t)k * N + gi];
          w = -ak * rb + papb * pk;
          v = pk;
        } else {
          w = -rb;
          v = a.x0[gi];
          mb = -(a.imbar[gi] + zb * Rk) * iM * iM;
          if (sl < T) { us[idx] = a.q[gi] / a.q[2 * N + gi]; vvs[idx] = a.q[N + gi] / a.q[2 * N + gi]; }
        }
        if (idx == tid) { o_rb = rb; o_zb = zb; o_R = Rk; o_pb = pb; }
      }
    }
    const int o = sl * SS + nn;
    ws[o] = w; vs[o] = v; irs[o] = ir;
    if constexpr (FIN) mbs[o] = mb;
  }
  stage_basis<NP>(B, sD, sW);
  __syncthreads();
  if constexpr (MODE == RV_PCG_ITER) {
    // own-node inputs used only with the outputs: requested here, out of the
    // entry burst, so that they arrive under the derivative and face phases
    // (P_{k-1}, AP_{k-1}, R_k do not exist at k = 0: valid stand-ins are read
    // and the values dropped)
    const double* APk = a.AP + (size_t)k * N;
    const double* Pm = k >= 1 ? a.P + (size_t)(k - 1) * N : a.P;
    const double* APm = k >= 1 ? a.AP + (size_t)(k - 1) * N : a.AP;
    const double* Rk = a.R + (size_t)k * N;
    f_irb = a.irbar[gme]; f_imb = a.imbar[gme]; f_xb = a.xbar[gme]; f_apk = APk[gme];
    f_pm = Pm[gme]; f_apm = APm[gme]; f_Rk = Rk[gme];
    if (!valid) { f_irb = 0.0; f_imb = 0.0; f_xb = 0.0; f_apk = 0.0; }
    if (!valid || k < 1) { f_pm = 0.0; f_apm = 0.0; f_Rk = 0.0; }
  }

  const int sL = t.nb_left(s), sR = t.nb_right(s), sB = t.nb_bottom(s), sT = t.nb_top(s);
  const double wown = ws[s * SS + n], vown = vs[s * SS + n], ir = irs[s * SS + n];
  double fc = 0.0, irb = 0.0;
  double bown = 0.0;
  if (valid) {
    double dxw = 0.0, dyw = 0.0, dxv = 0.0, dyv = 0.0;
#pragma unroll
    for (int kk = 0; kk < NP; ++kk) {
      dxw += sD[i * NP + kk] * ws[s * SS + kk + NP * j];
      dyw += sD[j * NP + kk] * ws[s * SS + i + NP * kk];
      dxv += sD[i * NP + kk] * vs[s * SS + kk + NP * j];
      dyv += sD[j * NP + kk] * vs[s * SS + i + NP * kk];
    }
    const double kx = al * sW[i] * sW[j] * g.Jy * g.invJx;
    const double ky = al * sW[i] * sW[j] * g.Jx * g.invJy;
    const double cm = sW[i] * sW[j] * g.Jx * g.Jy;
    gx[s * NP2 + n] = kx * ir * dxw;
    gy[s * NP2 + n] = ky * ir * dyw;
    irb = cm * vown * wown + kx * dxv * dxw + ky * dyv * dyw;
    if constexpr (FIN) {
      // IGR source b = c alpha [(uxx+uyy)^2 + uxx^2 + 2 uxy uyx + uyy^2]; b-bar is r0-bar = -w
      // (PCG, Chebyshev) plus what the Jacobi/Chebyshev iterations accumulated
      bown = -wown;
      if constexpr (MODE != RV_PCG_FIN) bown += a.bbar[gme];
      double uxx = 0.0, uxy = 0.0, uyx = 0.0, uyy = 0.0;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) {
        uxx += sD[i * NP + kk] * us[s * NP2 + kk + NP * j];
        uyx += sD[i * NP + kk] * vvs[s * NP2 + kk + NP * j];
        uxy += sD[j * NP + kk] * us[s * NP2 + i + NP * kk];
        uyy += sD[j * NP + kk] * vvs[s * NP2 + i + NP * kk];
      }
      uxx *= g.invJx; uyx *= g.invJx; uxy *= g.invJy; uyy *= g.invJy;
      const double f = bown * cm * al;
      Gxx[s * NP2 + n] = f * (2.0 * (uxx + uyy) + 2.0 * uxx);
      Gyy[s * NP2 + n] = f * (2.0 * (uxx + uyy) + 2.0 * uyy);
      Gxy[s * NP2 + n] = f * 2.0 * uyx;
      Gyx[s * NP2 + n] = f * 2.0 * uxy;
    }
  }
  // Face terms.  An element has 4 NP face nodes (left, right, bottom, top
  // faces, NP nodes each); the tile's T 4 NP face nodes are spread over the
  // workgroup face-major (item w: face F = w / (T NP), then node q along
  // the face, then element, fastest), so a wave works on one face direction with
  // compile-time D rows, strides and left/right roles.  Face node (F, q) of
  // element se leaves in LDS, at (F NP + q) FR + se, its part of the owner's A w (consistency +
  // penalty), its symmetry factor, and its part of d/d(1/rho_owner) of the
  // face's bilinear form in (w, v); the owner adds its own after the
  // barrier.  Faces F = 0, 2 (left, bottom) have the owner on their right
  // side, F = 1, 3 on their left.
  constexpr int NF = 4 * NP, NFI = T * NF, NFR = (NFI + NT - 1) / NT;
  auto face_node = [&](auto Fc, const int se, const int q) {
    constexpr int F = decltype(Fc)::value;
    constexpr bool xf = F < 2, own_L = (F & 1) != 0;
    constexpr int dro = own_L ? NP - 1 : 0, drn = own_L ? 0 : NP - 1;   // normal index, own / neighbour
    constexpr int stp = xf ? 1 : NP;
    double C = 0.0, Sy = 0.0, Rb = 0.0;
    const int lxe = se % TX, lye = se / TX;
    if (lxe < t.nvx && lye < t.nvy) {
      const int sN = F == 0 ? t.nb_left(se) : F == 1 ? t.nb_right(se) : F == 2 ? t.nb_bottom(se) : t.nb_top(se);
      const int bo = se * SS + (xf ? NP * q : q), bn = sN * SS + (xf ? NP * q : q);
      double dow = 0.0, dov = 0.0, dnw = 0.0, dnv = 0.0;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) {
        const double Do = B.D[dro * NP + kk], Dn = B.D[drn * NP + kk];
        dow += Do * ws[bo + stp * kk];
        dov += Do * vs[bo + stp * kk];
        dnw += Dn * ws[bn + stp * kk];
        dnv += Dn * vs[bn + stp * kk];
      }
      const int oo = bo + stp * dro, on = bn + stp * drn;
      const double wo = ws[oo], vo = vs[oo], iro = irs[oo];
      const double wn = ws[on], vn = vs[on], irn = irs[on];
      const double wf = xf ? sW[q] * g.Jy : sW[q] * g.Jx;
      const double invJ = xf ? g.invJx : g.invJy, h = xf ? g.dx : g.dy;
      const double irL = own_L ? iro : irn, irR = own_L ? irn : iro;
      const double jw = own_L ? (wo - wn) : (wn - wo);
      const double jv = own_L ? (vo - vn) : (vn - vo);
      const double dLw = own_L ? dow : dnw, dRw = own_L ? dnw : dow;
      const double dLv = own_L ? dov : dnv, dRv = own_L ? dnv : dov;
      const double sig = g.eta * fmax(irL, irR) / h;
      // apply A to w: the owner's consistency + penalty, and the symmetry factor
      const double avg = invJ * (irL * dLw + irR * dRw) / 2.0;
      C = own_L ? (-al * avg * wf + al * sig * jw * wf) : (al * avg * wf - al * sig * jw * wf);
      Sy = -al * (iro / 2.0) * invJ * jw * wf;
      // d/d(1/rho_owner) of the face's bilinear form in (w, v)
      const bool L_chosen = !(irR > irL);
      const double dv_own = own_L ? dLv : dRv, dw_own = own_L ? dLw : dRw;
      Rb = -al * wf * jw * invJ * dv_own / 2.0 - al / 2.0 * invJ * jv * wf * dw_own;
      if (own_L == L_chosen) Rb += al * g.eta / h * jv * wf * jw;
    }
    const int f = (F * NP + q) * FR + se;
    fC[f] = C;
    fS[f] = Sy;
    fR[f] = Rb;
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
  double y = 0.0, ab = 0.0;
  if (valid) {
    if (i == 0) { fc += fC[fb + j * FR]; irb += fR[fb + j * FR]; }
    if (i == NP - 1) { fc += fC[fb + R1 + j * FR]; irb += fR[fb + R1 + j * FR]; }
    if (j == 0) { fc += fC[fb + R2 + i * FR]; irb += fR[fb + R2 + i * FR]; }
    if (j == NP - 1) { fc += fC[fb + R3 + i * FR]; irb += fR[fb + R3 + i * FR]; }
    const double cm = sW[i] * sW[j] * g.Jx * g.Jy;
    y = cm * wown * ir;
    double st = 0.0;
#pragma unroll
    for (int kk = 0; kk < NP; ++kk) {
      st += sD[kk * NP + i] * gx[s * NP2 + kk + NP * j];
      st += sD[kk * NP + j] * gy[s * NP2 + i + NP * kk];
    }
    y += st;
    y += fS[fb + j * FR] * sD[i] + fS[fb + R1 + j * FR] * sD[(NP - 1) * NP + i];
    y += fS[fb + R2 + i * FR] * sD[j] + fS[fb + R3 + i * FR] * sD[(NP - 1) * NP + j];
    y += fc;
    // d/d alpha of <w, A v> = <v, K w>, K = (A - mass)/alpha
    if (al != 0.0) ab = vown * (y - cm * wown * ir) / al;
  }

  if constexpr (MODE == RV_VJP) {
    if (valid) {
      a.y[gme] = y;
      a.irbar[gme] = -ir * ir * irb;     // adjoint of rho
    }
    double rv[1] = {ab};
    block_sum<1>(rv, red);
    if (tid == 0) a.abar_blk[blockIdx.x] += rv[0];
  } else if constexpr (MODE == RV_PCG_ITER) {
    double dv[4] = {0.0, 0.0, 0.0, 0.0};
    double pbk = 0.0;
    if (valid) {
      const double xb = f_xb;
      pbk = bk * o_pb + ak * xb + papb * f_apk + y;
      if (k >= 1) {
        const double pm = f_pm, apm = f_apm;
        const double iM = l_iM[0];      // own node: slot tid of the tile loads
        dv[0] = pbk * pm;
        dv[1] = xb * pm;
        dv[2] = (o_rb + pbk * iM) * apm;
        dv[3] = (f_Rk * iM) * apm;
      }
    }
    // the four dot products for the next iteration and the alpha-bar terms,
    // summed per thread over the workgroup's tiles (the kernel reduces them)
#pragma unroll
    for (int q = 0; q < 4; ++q) L.acc[q] += dv[q];
    L.acc[4] += ab;
    if (valid) {
      store_wt(&a.irbar[gme], f_irb + irb);
      store_wt(&a.imbar[gme], f_imb + o_zb * o_R);
      store_wt(&a.rbar_out[gme], o_rb);
      store_wt(&a.pbar_out[gme], pbk);
    }
  } else if constexpr (MODE == RV_JAC_ITER) {
    if (valid) {
      const double xb1 = o_pb, iM = a.invM[gme];
      a.irbar[gme] += irb;
      a.xbar_out[gme] = xb1 + y;
      a.imbar[gme] += a.omega * (a.b[gme] - a.AX[(size_t)k * N + gme]) * xb1;
      a.bbar[gme] += a.omega * iM * xb1;
    }
    double rv[1] = {ab};
    block_sum<1>(rv, red);
    if (tid == 0) a.abar_blk[blockIdx.x] += rv[0];
  } else if constexpr (MODE == RV_CHEB_ITER) {
    if (valid) {
      const double iM = a.invM[gme];
      a.irbar[gme] += irb;
      a.bbar[gme] += o_rb;
      const double Xb = a.xbar[gme] + y;                       // adjoint of x_k
      const double Db = (have_carry ? a.pbar_in[gme] : 0.0) + Xb;   // adjoint of d_k
      const double zb = (k > 1) ? a.cc2 * Db : Db / a.theta;   // adjoint of z_k = r_{k-1} / M
      a.pbar_out[gme] = (k > 1) ? a.cc1 * Db : 0.0;
      a.rbar_out[gme] = zb * iM;
      a.imbar[gme] += zb * a.R[(size_t)(k - 1) * N + gme];
      a.xbar_out[gme] = Xb;
    }
    double rv[1] = {ab};
    block_sum<1>(rv, red);
    if (tid == 0) a.abar_blk[blockIdx.x] += rv[0];
  } else if constexpr (FIN) {
    double ab_extra = 0.0;
    if (valid) {
      a.x0bar[gme] = a.xbar[gme] + y;    // x0-bar = x-bar - A r0-bar, w = -r0-bar
      const double cm = sW[i] * sW[j] * g.Jx * g.Jy;
      // diag(A) reverse: M = cm/rho + x/y stiffness self-coupling + face terms
      const double iM = a.invM[gme];
      const double Mown = 1.0 / iM;
      const double mbo = mbs[s * SS + n];
      ab_extra += (al != 0.0) ? mbo * (Mown - cm * ir) / al : 0.0;
      double irbd = cm * mbo;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) {
        // x-stiffness of node (kk, j) couples to this node (i, j): alpha w_i w_j Jy/Jx D[i][kk]^2
        const double Dx = sD[i * NP + kk];
        irbd += al * sW[i] * sW[j] * g.Jy * g.invJx * Dx * Dx * mbs[s * SS + kk + NP * j];
        const double Dy = sD[j * NP + kk];
        irbd += al * sW[i] * sW[j] * g.Jx * g.invJy * Dy * Dy * mbs[s * SS + i + NP * kk];
      }
      auto dface = [&](bool own_L, int sn, int nnode, double invJ, double h, double wf, double Dself) {
        const double irn = irs[sn * SS + nnode], mbn = mbs[sn * SS + nnode];
        const double irL = own_L ? ir : irn, irR = own_L ? irn : ir;
        irbd += (own_L ? -1.0 : 1.0) * al * invJ * Dself * wf * mbo;
        const bool L_chosen = !(irR > irL);
        if (own_L == L_chosen) irbd += (mbo + mbn) * al * wf * g.eta / h;
      };
      if (i == NP - 1) dface(true, sR, NP * j, g.invJx, g.dx, sW[j] * g.Jy, sD[(NP - 1) * NP + NP - 1]);
      if (i == 0) dface(false, sL, (NP - 1) + NP * j, g.invJx, g.dx, sW[j] * g.Jy, sD[0]);
      if (j == NP - 1) dface(true, sT, i, g.invJy, g.dy, sW[i] * g.Jx, sD[(NP - 1) * NP + NP - 1]);
      if (j == 0) dface(false, sB, i + NP * (NP - 1), g.invJy, g.dy, sW[i] * g.Jx, sD[0]);
      // IGR source reverse: u-bar, v-bar gathered from G
      double ub = 0.0, vb = 0.0;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) {
        ub += g.invJx * sD[kk * NP + i] * Gxx[s * NP2 + kk + NP * j] + g.invJy * sD[kk * NP + j] * Gxy[s * NP2 + i + NP * kk];
        vb += g.invJx * sD[kk * NP + i] * Gyx[s * NP2 + kk + NP * j] + g.invJy * sD[kk * NP + j] * Gyy[s * NP2 + i + NP * kk];
      }
      const double rho = a.q[2 * N + gme];
      const double u = us[s * NP2 + n], v = vvs[s * NP2 + n];
      // b = cm alpha F(u): d/d alpha = b/alpha; recompute b from the G terms: b-bar * b / alpha
      {
        double uxx = 0.0, uxy = 0.0, uyx = 0.0, uyy = 0.0;
#pragma unroll
        for (int kk = 0; kk < NP; ++kk) {
          uxx += sD[i * NP + kk] * us[s * NP2 + kk + NP * j];
          uyx += sD[i * NP + kk] * vvs[s * NP2 + kk + NP * j];
          uxy += sD[j * NP + kk] * us[s * NP2 + i + NP * kk];
          uyy += sD[j * NP + kk] * vvs[s * NP2 + i + NP * kk];
        }
        uxx *= g.invJx; uyx *= g.invJx; uxy *= g.invJy; uyy *= g.invJy;
        const double dvg = uxx + uyy;
        ab_extra += bown * cm * (dvg * dvg + uxx * uxx + 2.0 * uxy * uyx + uyy * uyy);
      }
      const double irt = a.irbar[gme] + irb + irbd;
      a.qbar[gme] += ub / rho;
      a.qbar[N + gme] += vb / rho;
      a.qbar[2 * N + gme] += -ir * ir * irt - (ub * u + vb * v) / rho;
    }
    double rv[1] = {ab + ab_extra};
    block_sum<1>(rv, red);
    if (tid == 0) a.abar_blk[blockIdx.x] += rv[0];
  }
}

// The reverse SIP family.  The reverse PCG iteration may run on a bounded
// grid (the launcher's workgroup count, at most one per tile): a workgroup forms
// the iteration's reverse scalars once from the previous kernel's partial
// sums, then runs tiles blockIdx.x, blockIdx.x + gridDim.x, ..., and stores
// one partial sum per dot product for the whole workgroup, so the next
// iteration reduces gridDim.x partials instead of one per tile.  The other
// modes run one tile per workgroup.
template <int NP, int TX, int TY, int MODE, bool LOOP>
__device__ __forceinline__ void k2_sip_rev_kernel(const RevArgs& a) {
  RevLoop L;
  if constexpr (MODE == RV_PCG_ITER) {
    __shared__ unsigned arrive;   // waves done with the partial sums (block_partials_store_last)
    const int ntiles = tiles_of<TX, TY>(a.g);
    if (threadIdx.x == 0) arrive = 0;   // published by grid_finish's barrier
    // the iteration's reverse scalars from the previous kernel's partial
    // sums (the stored forward scalars are loaded with them)
    const int k = a.k;
    const FS fs(a.n_iter);
    // the partial sums are requested first; a one-tile workgroup forms the
    // grid total and the reverse scalars inside its tile, after the tile's
    // field loads are issued, so that both memory latencies overlap
    double pre4[4];
    grid_prefetch<4>(a.part_in, a.np_in, pre4);
    const double f_ab = a.abar_blk[blockIdx.x];
    L.ak = a.sc[fs.a(k)];
    L.bk = a.sc[fs.beta(k)];
    const double rzk = a.sc[fs.rz(k)], rzk1 = a.sc[fs.rz(k + 1)], papk = a.sc[fs.pap(k)];
    const double carry_next = a.first ? 0.0 : a.rs[RS::carry(k + 1)];
    auto scalars = [&](RevLoop& S) {
      double d[4];
      grid_finish<4, 4>(a.part_in, a.np_in, pre4, d);
      const double eps = 1e-30;
      const double bb = d[0];
      const double rzb = carry_next + bb / (rzk + eps);
      const double abk = d[1] - d[2] - 2.0 * rzb * d[3];
      const double papb = -abk * rzk / ((papk + eps) * (papk + eps));
      const double carry = -bb * rzk1 / ((rzk + eps) * (rzk + eps)) + abk / (papk + eps);
      // every lane holds the same totals: keep them in scalar registers
      S.rzb = uniform_d(rzb); S.papb = uniform_d(papb);
      S.r_bb = uniform_d(bb); S.r_abk = uniform_d(abk); S.r_carry = uniform_d(carry);
    };
    if constexpr (LOOP) {
      // formed before the tile loop: inside it, the loop's register budget
      // spills (256^2 20-step gradient 0.245 -> 0.299 s)
      scalars(L);
      for (int tt = blockIdx.x; tt < ntiles; tt += gridDim.x)
        k2_sip_rev_body<NP, TX, TY, MODE>(a, tt, ntiles, L, [](RevLoop&) {});
    } else {
      k2_sip_rev_body<NP, TX, TY, MODE>(a, blockIdx.x, gridDim.x, L, scalars);
    }
    double dv5[5] = {L.acc[0], L.acc[1], L.acc[2], L.acc[3], L.acc[4]};
    if (block_partials_store_last<5, 4>(dv5, a.part_out, gridDim.x, blockIdx.x, &arrive) &&
        (threadIdx.x & (WAVE - 1)) == wave_sum_lane())
      a.abar_blk[blockIdx.x] = f_ab + dv5[4];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const int k = a.k;
      a.rs[RS::bbar(k)] = L.r_bb;
      a.rs[RS::abar(k)] = L.r_abk;
      a.rs[RS::papbar(k)] = L.papb;
      a.rs[RS::rzbar_next(k)] = L.rzb;
      a.rs[RS::carry(k)] = L.r_carry;
    }
  } else {
    k2_sip_rev_body<NP, TX, TY, MODE>(a, blockIdx.x, gridDim.x, L, [](RevLoop&) {});
  }
}
// one tile per workgroup (every mode)
template <int NP, int TX, int TY, int MODE>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_sip_rev(RevArgs a) {
  k2_sip_rev_kernel<NP, TX, TY, MODE, false>(a);
}
// the reverse PCG iteration on fewer workgroups than tiles; the tile loop
// needs the compiler held to four waves per SIMD (two 8-wave workgroups per
// CU), which it otherwise spends on registers
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP, 4) void k2_sip_rev_loop(RevArgs a) {
  k2_sip_rev_kernel<NP, TX, TY, RV_PCG_ITER, true>(a);
}

// First reverse PCG step: a-bar_{n-1} = <x-bar, p_{n-1}> (the later adjoints
// are 0), on the reverse PCG iteration's grid: tiles blockIdx.x,
// blockIdx.x + gridDim.x, ..., one partial sum per workgroup.
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_pcg_rev_start(RevArgs a) {
  constexpr int NP2 = NP * NP, NT = TX * TY * NP2;
  __shared__ double red[4 * (NT / WAVE + 1)];
  const int ntiles = tiles_of<TX, TY>(a.g);
  const int tid = threadIdx.x, s = tid / NP2, n = tid % NP2;
  const int lx = s % TX, ly = s / TX;
  const size_t N = (size_t)a.g.Nex * a.g.Ney * NP2;
  const int nit = a.n_iter;
  double dv[1] = {0.0};
  for (int tt = blockIdx.x; tt < ntiles; tt += gridDim.x) {
    const TileInfo<NP, TX, TY> t(a.g, tt, ntiles);
    if (lx < t.nvx && ly < t.nvy) {
      const size_t gi = t.gidx(s, n);
      dv[0] += a.xbar[gi] * a.P[(size_t)(nit - 1) * N + gi];
    }
  }
  double d4[4] = {0.0, dv[0], 0.0, 0.0};
  block_partials_store<4>(d4, red, a.part_out, gridDim.x, blockIdx.x);
}

// ------------------------------------------------------------------------
// objectives: value (NV partial sums) and gradient
// ------------------------------------------------------------------------

template <int NP, int TX, int TY, bool GRAD>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_objective(ObjArgs a) {
  constexpr int NP2 = NP * NP, NT = TX * TY * NP2;
  __shared__ double red[2 * (NT / WAVE + 1)];
  TileInfo<NP, TX, TY> t(a.g);
  const Geo2& g = a.g;
  const int tid = threadIdx.x, s = tid / NP2, n = tid % NP2, i = n % NP, j = n / NP;
  const int lx = s % TX, ly = s / TX;
  const size_t N = (size_t)g.Nex * g.Ney * NP2;
  double acc[2] = {0.0, 0.0};
  if (lx < t.nvx && ly < t.nvy) {
    const size_t gi = t.gidx(s, n);
    const double wJ = a.B.w[i] * a.B.w[j] * g.Jx * g.Jy;
    const double x = (t.ex0 + lx) * g.dx + g.Jx * (1.0 + a.B.xi[i]);
    const double y = (t.ey0 + ly) * g.dy + g.Jy * (1.0 + a.B.xi[j]);
    const double mx = a.q[gi], my = a.q[N + gi], r = a.q[2 * N + gi], E = a.q[3 * N + gi];
    const double ke = (mx * mx + my * my) / (2.0 * r);
    double wt = 1.0;
    const double dx_ = x - a.prm[0], dy_ = y - a.prm[1];
    if (a.obj == 3 || a.obj == 5) wt = exp(-(dx_ * dx_ + dy_ * dy_) / (a.prm[2] * a.prm[2]));
    if (a.obj == 6) wt = (dx_ * dx_ + dy_ * dy_ < a.prm[2] * a.prm[2]) ? 1.0 : 0.0;
    const double gg = sin(2.0 * M_PI * x / a.Lx) * sin(2.0 * M_PI * y / a.Ly);
    const double P = (a.gamma - 1.0) * (E - ke);
    if (!GRAD) {
      switch (a.obj) {
        case 0: acc[0] = wJ * (r * r); break;
        case 1: acc[0] = wJ * (P * P); break;
        case 2: acc[0] = wJ * ke; acc[1] = wJ * r; break;
        case 3: case 6: acc[0] = wJ * wt * ke; break;
        case 4: acc[0] = wJ * mx * gg; break;
        case 5: acc[0] = wJ * r * wt; break;
      }
    } else {
      double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
      switch (a.obj) {
        case 0: b2 = 2.0 * wJ * r; break;
        case 1: {
          const double Pb = 2.0 * wJ * P;
          b3 = (a.gamma - 1.0) * Pb;
          b0 = -(a.gamma - 1.0) * Pb * mx / r;
          b1 = -(a.gamma - 1.0) * Pb * my / r;
          b2 = (a.gamma - 1.0) * Pb * ke / r;
          acc[0] = Pb * (E - ke);
          break;
        }
        case 2: {
          const double KE = a.out[1], M = a.out[2];
          b0 = wJ * mx / r / M;
          b1 = wJ * my / r / M;
          b2 = -wJ * ke / r / M - KE * wJ / (M * M);
          break;
        }
        case 3: case 6:
          b0 = wJ * wt * mx / r; b1 = wJ * wt * my / r; b2 = -wJ * wt * ke / r; break;
        case 4: b0 = wJ * gg; break;
        case 5: b2 = wJ * wt; break;
      }
      a.qbar[gi] = b0; a.qbar[N + gi] = b1; a.qbar[2 * N + gi] = b2; a.qbar[3 * N + gi] = b3;
    }
  }
  double* out = a.out;
  double* gbar = a.gbar;
  const int obj = a.obj;
  if (!GRAD) {
    grid_sum<2>(acc, red, a.partials, a.counter, gridDim.x, blockIdx.x, [out, obj](double* tot) {
      out[1] = tot[0];
      out[2] = tot[1];
      out[0] = (obj == 2) ? tot[0] / tot[1] : tot[0];
    });
  } else {
    double one[1] = {acc[0]};
    grid_sum<1>(one, red, a.partials, a.counter, gridDim.x, blockIdx.x, [gbar](double* tot) { gbar[0] += tot[0]; });
  }
}

// Sum of per-workgroup accumulators into out[0] (single workgroup, fixed order).
__global__ void k_sum_blocks(const double* blk, int nb, double* out) {
  __shared__ double red[16];
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += blk[b];
  block_sum<1>(v, red);
  if (threadIdx.x == 0) out[0] += v[0];
}
__global__ void k_axpy(double* y, const double* x, double c, size_t n) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] += c * x[k];
}

// ------------------------------------------------------------------------
template <int NP, int TX, int TY>
struct LaunchRev {
  static dim3 grid(const Geo2& g) {
    return dim3(((g.Nex + TX - 1) / TX) * ((g.Ney + TY - 1) / TY));
  }
  static dim3 block() { return dim3(TX * TY * NP * NP); }
  static void hyp(const HypRevArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k2_hyp_rev<NP, TX, TY>), grid(a.g), block(), 0, st, a);
  }
  static void rev(int mode, const RevArgs& a, hipStream_t st) {
    dim3 G = grid(a.g), Bk = block(), GP(a.ngroups);
    switch (mode) {
      case RV_VJP: hipLaunchKernelGGL((k2_sip_rev<NP, TX, TY, RV_VJP>), G, Bk, 0, st, a); break;
      case RV_PCG_START: hipLaunchKernelGGL((k2_pcg_rev_start<NP, TX, TY>), GP, Bk, 0, st, a); break;
      case RV_PCG_ITER:
        if ((int)GP.x < (int)G.x) hipLaunchKernelGGL((k2_sip_rev_loop<NP, TX, TY>), GP, Bk, 0, st, a);
        else hipLaunchKernelGGL((k2_sip_rev<NP, TX, TY, RV_PCG_ITER>), G, Bk, 0, st, a);
        break;
      case RV_PCG_FIN: hipLaunchKernelGGL((k2_sip_rev<NP, TX, TY, RV_PCG_FIN>), G, Bk, 0, st, a); break;
      case RV_JAC_ITER: hipLaunchKernelGGL((k2_sip_rev<NP, TX, TY, RV_JAC_ITER>), G, Bk, 0, st, a); break;
      case RV_JAC_FIN: hipLaunchKernelGGL((k2_sip_rev<NP, TX, TY, RV_JAC_FIN>), G, Bk, 0, st, a); break;
      case RV_CHEB_ITER: hipLaunchKernelGGL((k2_sip_rev<NP, TX, TY, RV_CHEB_ITER>), G, Bk, 0, st, a); break;
      case RV_CHEB_FIN: hipLaunchKernelGGL((k2_sip_rev<NP, TX, TY, RV_CHEB_FIN>), G, Bk, 0, st, a); break;
    }
  }
  static void obj(const ObjArgs& a, bool grad, hipStream_t st) {
    if (grad) hipLaunchKernelGGL((k2_objective<NP, TX, TY, true>), grid(a.g), block(), 0, st, a);
    else hipLaunchKernelGGL((k2_objective<NP, TX, TY, false>), grid(a.g), block(), 0, st, a);
  }
};

// SIP / solver reverse: the SIP tile (4x4, 8x4, 8x8, 2x2 at p = 3); the
// hyperbolic reverse and the objectives: the hyperbolic tile (no 8x8).
#define IGR_REV_DISPATCH(np, tx, ty, CALL)                                      \
  switch (np) {                                                                 \
    case 2: LaunchRev<2, 8, 8>::CALL; break;                                    \
    case 3: LaunchRev<3, 4, 4>::CALL; break;                                    \
    case 4:                                                                     \
      if ((tx) == 8 && (ty) == 4) LaunchRev<4, 8, 4>::CALL;                     \
      else if ((tx) == 8 && (ty) == 8) LaunchRev<4, 8, 8>::CALL;                \
      else if ((tx) == 2 && (ty) == 2) LaunchRev<4, 2, 2>::CALL;                \
      else LaunchRev<4, 4, 4>::CALL;                                            \
      break;                                                                    \
    case 5: LaunchRev<5, 2, 4>::CALL; break;                                    \
  }
#define IGR_HREV_DISPATCH(np, tx, ty, CALL)                                     \
  switch (np) {                                                                 \
    case 2: LaunchRev<2, 8, 8>::CALL; break;                                    \
    case 3: LaunchRev<3, 4, 4>::CALL; break;                                    \
    case 4:                                                                     \
      if ((tx) == 8 && (ty) == 4) LaunchRev<4, 8, 4>::CALL;                     \
      else if ((tx) == 2 && (ty) == 2) LaunchRev<4, 2, 2>::CALL;                \
      else LaunchRev<4, 4, 4>::CALL;                                            \
      break;                                                                    \
    case 5: LaunchRev<5, 2, 4>::CALL; break;                                    \
  }

void launch2_hyp_rev(const Ctx& c, const HypRevArgs& a) { IGR_HREV_DISPATCH(c.np, c.htile_x, c.htile_y, hyp(a, c.stream)); }
// workgroups of the reverse PCG iteration (and of its start kernel): at
// most c.rgroups (0: one per reverse tile)
int grid2_rpgroups(const Ctx& c);
void launch2_rev(const Ctx& c, int mode, const RevArgs& a0) {
  RevArgs a = a0;
  a.ngroups = grid2_rpgroups(c);
  a.np_in = a.ngroups;   // the start kernel and every iteration write that many partial sums
  IGR_REV_DISPATCH(c.np, c.rtile_x, c.rtile_y, rev(mode, a, c.stream));
}
void launch2_obj(const Ctx& c, const ObjArgs& a, bool grad) {
  IGR_HREV_DISPATCH(c.np, c.htile_x, c.htile_y, obj(a, grad, c.stream));
}
void launch_sum_blocks(const Ctx& c, const double* blk, int nb, double* out) {
  hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(1024), 0, c.stream, blk, nb, out);
}
void launch_axpy(const Ctx& c, double* y, const double* x, double cf, size_t n) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, y, x, cf, n);
}

}  // namespace igr
