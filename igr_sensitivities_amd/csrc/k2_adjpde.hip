// 2D continuous ("optimise-then-discretise") adjoint kernels (gfx950, FP64),
// the GPU form of IGR_adjoints_2d/adjoint_pde.jl:
//
//   k2_adj_stage     compute_adjoint_rhs_conservative! (:471-506) fused with
//                    the SSP-RK3 stage of run_adjoint_conservative (:593-607):
//                    local adjoint flux divergence (:226-265), adjoint LLF
//                    face corrections (:268-309), minus the local source
//                    (:136-212), plus the nonlocal source from Pi (:373-457),
//                    then the stage combination.  One workgroup per tile of
//                    elements with the face-node halo ring, as k2_hyp.
//   k2_adj_b         right-hand side of the Pi solve: alpha w_i w_j Jx Jy
//                    (div a_mu + grad a_E . u) (:332-357, :484-486).
//   k_adj_terminal2  L2 gradient of the objective at t = T (:535-576).
//
// The nonlocal source needs second derivatives of products with Pi; they are
// element-local (the LGL differentiation acts inside an element), so the
// products live in LDS and no extra halo is needed.
#include "igr_ctx.hpp"
#include "adjpde_common.hpp"

namespace igr {

#pragma clang fp contract(off)
// adjoint_flux_local_x / _y (adjoint_pde.jl:41-78).  q = (mx, my, rho, E, Sigma).
template <int DIR>
__device__ __forceinline__ void adj_flux2(double g, const double q[5], const double a[4], double f[4]) {
  const double mx = q[0], my = q[1], r = q[2], E = q[3], S = q[4];
  const double ux = mx / r, uy = my / r;
  const double P = (g - 1.0) * (E - (mx * mx + my * my) / (2.0 * r));
  const double H = E + P + S;
  const double Pmx = -(g - 1.0) * ux;
  const double Pmy = -(g - 1.0) * uy;
  const double Pr = (g - 1.0) * (ux * ux + uy * uy) / 2.0;
  const double PE = g - 1.0;
  const double amu = a[0] * mx + a[1] * my;
  if (DIR == 0) {
    f[0] = amu / r + a[0] * ux + Pmx * a[0] + a[2] + a[3] * (H / r) + a[3] * Pmx * mx / r;
    f[1] = a[1] * ux + Pmy * a[0] + a[3] * Pmy * mx / r;
    f[2] = -ux * amu / r + Pr * a[0] + a[3] * mx * (Pr / r - H / (r * r));
    f[3] = PE * a[0] + a[3] * (1.0 + PE) * ux;
  } else {
    f[0] = a[0] * uy + Pmx * a[1] + a[3] * Pmx * my / r;
    f[1] = amu / r + a[1] * uy + Pmy * a[1] + a[2] + a[3] * (H / r) + a[3] * Pmy * my / r;
    f[2] = -uy * amu / r + Pr * a[1] + a[3] * my * (Pr / r - H / (r * r));
    f[3] = PE * a[1] + a[3] * (1.0 + PE) * uy;
  }
}
// adjoint LLF (adjoint_pde.jl:95-98): the dissipation has the opposite sign
// of the forward flux's
__device__ __forceinline__ void adj_llf_combine(const double fL[4], const double fR[4], const double aL[4],
                                                const double aR[4], double lam, double f[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) f[v] = (fL[v] + fR[v]) / 2.0 - lam * (aL[v] - aR[v]) / 2.0;
}

// compute_adjoint_source_local! at one node (adjoint_pde.jl:163-208) from
// the physical derivatives of the primal fields; returns S_loc (4).
__device__ __forceinline__ void adj_source_node2(double g, const double qo[5], const double d[10], const double ao[4],
                                                 double Sl[4]) {
  const double gm = g - 1.0;
  const double mxx = d[0], mxy = d[1], myx = d[2], myy = d[3], rx = d[4], ry = d[5], Ex = d[6], Ey = d[7];
  const double Sx = d[8], Sy = d[9];
  const double r = qo[2], mx = qo[0], my = qo[1], en = qo[3];
  const double ux = mx / r, uy = my / r;
  const double uxx = (mxx - ux * rx) / r;
  const double uyy = (myy - uy * ry) / r;
  const double P = gm * (en - (mx * mx + my * my) / (2.0 * r));
  const double H = en + P + qo[4];
  const double Pr = gm * (ux * ux + uy * uy) / 2.0;
  const double Pbx = gm * (Ex - (mx * mxx + my * myx) / r + (mx * mx + my * my) * rx / (2.0 * (r * r))) + Sx;
  const double Pby = gm * (Ey - (mx * mxy + my * myy) / r + (mx * mx + my * my) * ry / (2.0 * (r * r))) + Sy;
  const double dPx_dmx = gm * (-mxx / r + mx * rx / (r * r));
  const double dPx_dmy = gm * (-myx / r + my * rx / (r * r));
  const double dPx_dr = gm * ((mx * mxx + my * myx) / (r * r) - (mx * mx + my * my) * rx / (r * r * r));
  const double dPy_dmx = gm * (-mxy / r + mx * ry / (r * r));
  const double dPy_dmy = gm * (-myy / r + my * ry / (r * r));
  const double dPy_dr = gm * ((mx * mxy + my * myy) / (r * r) - (mx * mx + my * my) * ry / (r * r * r));
  const double HEx = Ex + Pbx, HEy = Ey + Pby;
  const double dmx_mx = mxx / r + uxx - mx * rx / (r * r) + dPx_dmx + uyy;
  const double dmx_my = dPx_dmy + mxy / r - mx * ry / (r * r);
  const double dmx_r = mxx * (-ux / r) + mx * (ux * rx / (r * r) - uxx / r) + dPx_dr +
                       mxy * (-uy / r) + mx * (uy * ry / (r * r) - uyy / r);
  const double dmy_mx = myx / r - my * rx / (r * r) + dPy_dmx;
  const double dmy_my = uxx + myy / r + uyy - my * ry / (r * r) + dPy_dmy;
  const double dmy_r = myx * (-ux / r) + my * (ux * rx / (r * r) - uxx / r) +
                       myy * (-uy / r) + my * (uy * ry / (r * r) - uyy / r) + dPy_dr;
  const double dE_mx = dPx_dmx * ux + HEx / r + (-gm * ux) * uxx + H * (-rx / (r * r)) +
                       dPy_dmx * uy + (-gm * ux) * uyy;
  const double dE_my = dPx_dmy * ux + (-gm * uy) * uxx + dPy_dmy * uy + HEy / r +
                       (-gm * uy) * uyy + H * (-ry / (r * r));
  const double dE_r = dPx_dr * ux + HEx * (-ux / r) + Pr * uxx + H * (ux * rx / (r * r) - uxx / r) +
                      dPy_dr * uy + HEy * (-uy / r) + Pr * uyy + H * (uy * ry / (r * r) - uyy / r);
  const double dE_E = g * (uxx + uyy);
  Sl[0] = ao[0] * dmx_mx + ao[1] * dmy_mx + ao[3] * dE_mx;
  Sl[1] = ao[0] * dmx_my + ao[1] * dmy_my + ao[3] * dE_my;
  Sl[2] = ao[0] * dmx_r + ao[1] * dmy_r + ao[3] * dE_r;
  Sl[3] = ao[3] * dE_E;
}
#pragma clang fp contract(fast)

template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_adj_stage(AdjStageArgs a) {
  constexpr int NP2 = NP * NP, T = TX * TY, NH = 2 * (TX + TY), NT = T * NP2;
  __shared__ double qs[5][NT];
  __shared__ double as[4][NT];
  __shared__ double F[8][NT];     // adjoint fluxes (x: 0-3, y: 4-7), then the nonlocal products
  __shared__ double wsx[NT], wsy[NT];
  __shared__ double hq[5][NH * NP], ha[4][NH * NP];
  __shared__ double fxr[4][T * NP], fxl[4][T * NP], fyt[4][T * NP], fyb[4][T * NP];
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

  double qo[5] = {0.0, 0.0, 1.0, 1.0, 0.0}, ao[4] = {0.0, 0.0, 0.0, 0.0};
  double pio = 0.0;
  if (valid) {
#pragma unroll
    for (int v = 0; v < 4; ++v) { qo[v] = a.q[v * N + gme]; ao[v] = a.A[v * N + gme]; }
    qo[4] = a.S[gme];
    if (a.Pi) pio = a.Pi[gme];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) qs[v][tid] = qo[v];
#pragma unroll
  for (int v = 0; v < 4; ++v) as[v][tid] = ao[v];
  for (int idx = tid; idx < NH * NP; idx += NT) {
    const int h = idx / NP, k = idx % NP;
    size_t gi;
    if (t.halo_face_global(h, k, gi)) {
#pragma unroll
      for (int v = 0; v < 4; ++v) { hq[v][idx] = a.q[v * N + gi]; ha[v][idx] = a.A[v * N + gi]; }
      hq[4][idx] = a.S[gi];
    } else {
      hq[0][idx] = 0.0; hq[1][idx] = 0.0; hq[2][idx] = 1.0; hq[3][idx] = 1.0; hq[4][idx] = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) ha[v][idx] = 0.0;
    }
  }
  double fxo[4], fyo[4];
  adj_flux2<0>(gam, qo, ao, fxo);
  adj_flux2<1>(gam, qo, ao, fyo);
#pragma unroll
  for (int v = 0; v < 4; ++v) { F[v][tid] = fxo[v]; F[4 + v][tid] = fyo[v]; }
  const bool face = (i == 0 || i == NP - 1 || j == 0 || j == NP - 1);
  if (valid && face) {
    const double c = sound2(gam, qo, a.flag);
    wsx[tid] = fabs(qo[0] / qo[2]) + c;
    wsy[tid] = fabs(qo[1] / qo[2]) + c;
  }
  stage_basis<NP>(B, sD, nullptr);
  __syncthreads();

  double acc[4] = {0.0, 0.0, 0.0, 0.0}, Sl[4] = {0.0, 0.0, 0.0, 0.0};
  double prod[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, Rtr = 0.0;
  if (valid) {
    const int e0 = s * NP2;
    // volume term +(1/Jx) Dx fx + (1/Jy) Dy fy (adjoint_pde.jl:252-265)
#pragma unroll
    for (int k = 0; k < NP; ++k) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        acc[v] += g.invJx * sD[i * NP + k] * F[v][e0 + k + NP * j] + g.invJy * sD[j * NP + k] * F[4 + v][e0 + i + NP * k];
    }
    // adjoint LLF face fluxes, one evaluation per face inside the tile
    auto nbq = [&](int sl, int node, int hk, double qn[5], double an_[4]) {
      if (sl < T) {
#pragma unroll
        for (int v = 0; v < 5; ++v) qn[v] = qs[v][sl * NP2 + node];
#pragma unroll
        for (int v = 0; v < 4; ++v) an_[v] = as[v][sl * NP2 + node];
      } else {
        const int h = sl - T;
#pragma unroll
        for (int v = 0; v < 5; ++v) qn[v] = hq[v][h * NP + hk];
#pragma unroll
        for (int v = 0; v < 4; ++v) an_[v] = ha[v][h * NP + hk];
      }
    };
    if (i == NP - 1) {
      double qR[5], aR[4], fR[4], f[4];
      const int sR = t.nb_right(s);
      nbq(sR, NP * j, j, qR, aR);
      const double wR = (sR < T) ? wsx[sR * NP2 + NP * j] : fabs(qR[0] / qR[2]) + sound2(gam, qR, a.flag);
      adj_flux2<0>(gam, qR, aR, fR);
      adj_llf_combine(fxo, fR, ao, aR, fmax_first(wsx[tid], wR), f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fxr[v][s * NP + j] = f[v];
    }
    if (i == 0 && lx == 0) {
      double qL[5], aL[4], fL[4], f[4];
      nbq(t.nb_left(s), (NP - 1) + NP * j, j, qL, aL);
      const double wL = fabs(qL[0] / qL[2]) + sound2(gam, qL, a.flag);
      adj_flux2<0>(gam, qL, aL, fL);
      adj_llf_combine(fL, fxo, aL, ao, fmax_first(wL, wsx[tid]), f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fxl[v][s * NP + j] = f[v];
    }
    if (j == NP - 1) {
      double qR[5], aR[4], fR[4], f[4];
      const int sT = t.nb_top(s);
      nbq(sT, i, i, qR, aR);
      const double wR = (sT < T) ? wsy[sT * NP2 + i] : fabs(qR[1] / qR[2]) + sound2(gam, qR, a.flag);
      adj_flux2<1>(gam, qR, aR, fR);
      adj_llf_combine(fyo, fR, ao, aR, fmax_first(wsy[tid], wR), f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fyt[v][s * NP + i] = f[v];
    }
    if (j == 0 && ly == 0) {
      double qL[5], aL[4], fL[4], f[4];
      nbq(t.nb_bottom(s), i + NP * (NP - 1), i, qL, aL);
      const double wL = fabs(qL[1] / qL[2]) + sound2(gam, qL, a.flag);
      adj_flux2<1>(gam, qL, aL, fL);
      adj_llf_combine(fL, fyo, aL, ao, fmax_first(wL, wsy[tid]), f);
#pragma unroll
      for (int v = 0; v < 4; ++v) fyb[v][s * NP + i] = f[v];
    }
    // derivatives of the primal fields inside the element (adjoint_pde.jl:147-161)
    double d[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double uxx = 0.0, uxy = 0.0, uyx = 0.0, uyy = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int kx = e0 + k + NP * j, ky = e0 + i + NP * k;
      const double Dx = sD[i * NP + k], Dy = sD[j * NP + k];
      d[0] += Dx * qs[0][kx]; d[1] += Dy * qs[0][ky];
      d[2] += Dx * qs[1][kx]; d[3] += Dy * qs[1][ky];
      d[4] += Dx * qs[2][kx]; d[5] += Dy * qs[2][ky];
      d[6] += Dx * qs[3][kx]; d[7] += Dy * qs[3][ky];
      d[8] += Dx * qs[4][kx]; d[9] += Dy * qs[4][ky];
      // nodal velocity derivatives (adjoint_pde.jl:397-406)
      uxx += Dx * (qs[0][kx] / qs[2][kx]); uyx += Dx * (qs[1][kx] / qs[2][kx]);
      uxy += Dy * (qs[0][ky] / qs[2][ky]); uyy += Dy * (qs[1][ky] / qs[2][ky]);
    }
#pragma unroll
    for (int m = 0; m < 10; m += 2) { d[m] *= g.invJx; d[m + 1] *= g.invJy; }
    adj_source_node2(gam, qo, d, ao, Sl);
    if (a.Pi) {
      uxx *= g.invJx; uyx *= g.invJx; uxy *= g.invJy; uyy *= g.invJy;
      const double divu = uxx + uyy;
      const double r = qo[2];
      prod[0] = pio * divu;
      prod[1] = pio * uxx; prod[2] = pio * uyx;
      prod[3] = pio * uxy; prod[4] = pio * uyy;
      prod[5] = pio * d[8] / r; prod[6] = pio * d[9] / r;
      Rtr = divu * divu + uxx * uxx + 2.0 * uxy * uyx + uyy * uyy;
    }
  }
  __syncthreads();   // the fluxes in F are consumed; F now takes the products
  if (a.Pi) {
#pragma unroll
    for (int m = 0; m < 7; ++m) F[m][tid] = prod[m];
  }
  __syncthreads();
  if (!valid) return;

  // face corrections, x faces then y faces (adjoint_pde.jl:277-285, 299-307)
  if (i == NP - 1) {
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += g.invJx * (fxr[v][s * NP + j] - fxo[v]) / B.w[NP - 1];
  }
  if (i == 0) {
    const int sL = t.nb_left(s);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double f = (sL < T) ? fxr[v][sL * NP + j] : fxl[v][s * NP + j];
      acc[v] -= g.invJx * (f - fxo[v]) / B.w[0];
    }
  }
  if (j == NP - 1) {
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += g.invJy * (fyt[v][s * NP + i] - fyo[v]) / B.w[NP - 1];
  }
  if (j == 0) {
    const int sB = t.nb_bottom(s);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double f = (sB < T) ? fyt[v][sB * NP + i] : fyb[v][s * NP + i];
      acc[v] -= g.invJy * (f - fyo[v]) / B.w[0];
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] -= Sl[v];
  if (a.Pi) {
    // nonlocal source (adjoint_pde.jl:420-455)
    const int e0 = s * NP2;
    double dxPd = 0.0, dyPd = 0.0, dxPxx = 0.0, dyPyx = 0.0, dxPxy = 0.0, dyPyy = 0.0, dxCx = 0.0, dyCy = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int kx = e0 + k + NP * j, ky = e0 + i + NP * k;
      const double Dx = sD[i * NP + k], Dy = sD[j * NP + k];
      dxPd += Dx * F[0][kx];
      dyPd += Dy * F[0][ky];
      dxPxx += Dx * F[1][kx];
      dyPyx += Dy * F[2][ky];
      dxPxy += Dx * F[3][kx];
      dyPyy += Dy * F[4][ky];
      dxCx += Dx * F[5][kx];
      dyCy += Dy * F[6][ky];
    }
    dxPd *= g.invJx; dyPd *= g.invJy;
    dxPxx *= g.invJx; dyPyx *= g.invJy;
    dxPxy *= g.invJx; dyPyy *= g.invJy;
    dxCx *= g.invJx; dyCy *= g.invJy;
    const double Wx = dxPd + dxPxx + dyPyx;
    const double Wy = dyPd + dxPxy + dyPyy;
    const double divC = dxCx + dyCy;
    const double r = qo[2];
    acc[0] += -2.0 * Wx / r;
    acc[1] += -2.0 * Wy / r;
    acc[2] += (pio * Rtr + divC) / r + 2.0 * (Wx * qo[0] + Wy * qo[1]) / (r * r);
  }
  // SSP-RK3 stage combination (adjoint_pde.jl:594-607)
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const size_t gi = v * N + gme;
    double o;
    if (a.mode == 0) o = acc[v];
    else if (a.mode == 1) o = ao[v] + a.cdt * acc[v];
    else if (a.mode == 2) o = 3.0 / 4.0 * a.an[gi] + 1.0 / 4.0 * ao[v] + a.cdt * acc[v];
    else o = 1.0 / 3.0 * a.an[gi] + 2.0 / 3.0 * ao[v] + a.cdt * acc[v];
    a.out[gi] = o;
  }
}

// right-hand side of the Pi solve (adjoint_pde.jl:332-357 and :484-486)
template <int NP, int TX, int TY>
__global__ __launch_bounds__(TX * TY * NP * NP) void k2_adj_b(AdjBArgs a) {
  constexpr int NP2 = NP * NP, NT = TX * TY * NP2;
  __shared__ double ls[3][NT];   // a_mx, a_my, a_E
  __shared__ double sD[NP * NP], sW[NP];     // basis (stage_basis)
  const Geo2 g = a.g;
  const BasisK& B = a.B;
  TileInfo<NP, TX, TY> t(g);
  const int tid = threadIdx.x;
  const int s = tid / NP2, n = tid % NP2, i = n % NP, j = n / NP;
  const int lx = s % TX, ly = s / TX;
  const bool valid = lx < t.nvx && ly < t.nvy;
  const size_t N = (size_t)g.Nex * g.Ney * NP2;
  const size_t gme = valid ? t.gidx(s, n) : 0;
  double ux = 0.0, uy = 0.0;
  if (valid) {
    ls[0][tid] = a.A[gme];
    ls[1][tid] = a.A[N + gme];
    ls[2][tid] = a.A[3 * N + gme];
    const double r = a.q[2 * N + gme];
    ux = a.q[gme] / r;
    uy = a.q[N + gme] / r;
  } else {
    ls[0][tid] = 0.0; ls[1][tid] = 0.0; ls[2][tid] = 0.0;
  }
  stage_basis<NP>(B, sD, sW);
  __syncthreads();
  if (!valid) return;
  const int e0 = s * NP2;
  double dax = 0.0, day = 0.0, dEx = 0.0, dEy = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    dax += sD[i * NP + k] * ls[0][e0 + k + NP * j];
    day += sD[j * NP + k] * ls[1][e0 + i + NP * k];
    dEx += sD[i * NP + k] * ls[2][e0 + k + NP * j];
    dEy += sD[j * NP + k] * ls[2][e0 + i + NP * k];
  }
  dax *= g.invJx; dEx *= g.invJx;
  day *= g.invJy; dEy *= g.invJy;
  const double bb = dax + day + dEx * ux + dEy * uy;
  a.b[gme] = bb * (a.alpha * sW[i] * sW[j] * g.Jx * g.Jy);
}

// terminal condition (adjoint_pde.jl:535-576; WINDOWED_DENSITY and DISC_KE
// follow from the same quadrature sums as the objective kernel)
__global__ void k_adj_terminal2(AdjTermArgs a) {
  const int np = a.np, np2 = np * np;
  const size_t N = (size_t)a.g.Nex * a.g.Ney * np2;
  for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < N; n += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(n % np), j = (int)((n / np) % np);
    const size_t e = n / np2;
    const int ex = (int)(e % a.g.Nex), ey = (int)(e / a.g.Nex);
    const double x = ex * a.g.dx + a.g.Jx * (1.0 + a.B.xi[i]);
    const double y = ey * a.g.dy + a.g.Jy * (1.0 + a.B.xi[j]);
    const double mx = a.q[n], my = a.q[N + n], r = a.q[2 * N + n], E = a.q[3 * N + n];
    const double ux = mx / r, uy = my / r;
    double o[4] = {0.0, 0.0, 0.0, 0.0};
    const double dx_ = x - a.prm[0], dy_ = y - a.prm[1];
    switch (a.obj) {
      case 0: o[2] = 2.0 * r; break;
      case 1: {
        const double P = (a.gamma - 1.0) * (E - (mx * mx + my * my) / (2.0 * r));
        o[0] = -2.0 * (a.gamma - 1.0) * P * ux;
        o[1] = -2.0 * (a.gamma - 1.0) * P * uy;
        o[2] = (a.gamma - 1.0) * P * (ux * ux + uy * uy);
        o[3] = 2.0 * (a.gamma - 1.0) * P;
        break;
      }
      case 2: {
        const double KE = a.sums[1], Mt = a.sums[2];
        o[0] = ux / Mt;
        o[1] = uy / Mt;
        o[2] = -(mx * mx + my * my) / (2.0 * (r * r)) / Mt - KE / (Mt * Mt);
        break;
      }
      case 3: case 6: {
        const double w = (a.obj == 3) ? exp(-(dx_ * dx_ + dy_ * dy_) / (a.prm[2] * a.prm[2]))
                                      : ((dx_ * dx_ + dy_ * dy_ < a.prm[2] * a.prm[2]) ? 1.0 : 0.0);
        o[0] = w * ux;
        o[1] = w * uy;
        o[2] = -w * (ux * ux + uy * uy) / 2.0;
        break;
      }
      case 4: o[0] = sin(2.0 * M_PI * x / a.Lx) * sin(2.0 * M_PI * y / a.Ly); break;
      case 5: o[2] = exp(-(dx_ * dx_ + dy_ * dy_) / (a.prm[2] * a.prm[2])); break;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) a.a[v * N + n] = o[v];
  }
}

// ------------------------------------------------------------------------
// launchers (the hyperbolic-family tile shape of the context)
// ------------------------------------------------------------------------
template <int NP, int TX, int TY>
struct LaunchAdj {
  static dim3 grid(const Geo2& g) { return dim3(((g.Nex + TX - 1) / TX) * ((g.Ney + TY - 1) / TY)); }
  static dim3 block() { return dim3(TX * TY * NP * NP); }
  static void stage(const AdjStageArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k2_adj_stage<NP, TX, TY>), grid(a.g), block(), 0, st, a);
  }
  static void rhs(const AdjBArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((k2_adj_b<NP, TX, TY>), grid(a.g), block(), 0, st, a);
  }
};
#define IGR_ADJ_TILE_DISPATCH(np, tx, ty, CALL)                          \
  switch (np) {                                                          \
    case 2: LaunchAdj<2, 8, 8>::CALL; break;                             \
    case 3: LaunchAdj<3, 4, 4>::CALL; break;                             \
    case 4:                                                              \
      if ((tx) == 8 && (ty) == 4) LaunchAdj<4, 8, 4>::CALL;              \
      else if ((tx) == 2 && (ty) == 2) LaunchAdj<4, 2, 2>::CALL;         \
      else LaunchAdj<4, 4, 4>::CALL;                                     \
      break;                                                             \
    case 5: LaunchAdj<5, 2, 4>::CALL; break;                             \
  }

void launch2_adj_stage(const Ctx& c, const AdjStageArgs& a) {
  IGR_ADJ_TILE_DISPATCH(c.np, c.htile_x, c.htile_y, stage(a, c.stream));
}
void launch2_adj_b(const Ctx& c, const AdjBArgs& a) {
  IGR_ADJ_TILE_DISPATCH(c.np, c.htile_x, c.htile_y, rhs(a, c.stream));
}
void launch2_adj_terminal(const Ctx& c, const AdjTermArgs& a) {
  const unsigned nb = (unsigned)std::min<size_t>((c.N + 255) / 256, 4096);
  hipLaunchKernelGGL(k_adj_terminal2, dim3(std::max(nb, 1u)), dim3(256), 0, c.stream, a);
}

}  // namespace igr
