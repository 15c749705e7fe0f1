// 1D continuous ("optimise-then-discretise") adjoint for gfx950, FP64: one
// workgroup integrates the adjoint of one instance backward through all its
// stored snapshots in a single launch (the 1D meshes are small; see
// k1_forward.hip), following IGR_adjoints_1d/adjoint_pde.jl:
//   terminal condition            :479-513
//   local adjoint flux / LLF      :46-88
//   local source                  :102-176
//   local DG operator             :191-264
//   Pi right-hand side            :278-301, :423-426
//   nonlocal source               :321-387
//   SSP-RK3 in reversed time      :524-549
// The Pi solve reuses the forward's SIP apply and fixed-iteration solvers
// (struct K1) with 1/rho and diag(A) of the step's snapshot; Pi is
// warm-started from the previous solve, as in the reference.
#include "igr_ctx.hpp"
#include "adjpde_common.hpp"
#include "k1_forward.hpp"

namespace igr {

#pragma clang fp contract(off)
// adjoint_flux_local (adjoint_pde.jl:46-66); q = (mu, rho, E, Sigma)
__device__ __forceinline__ void adj_flux1(double g, const double q[4], const double a[3], double f[3]) {
  const double m = q[0], r = q[1], E = q[2], S = q[3];
  const double u = m / r;
  const double P = (g - 1.0) * (E - m * m / (2.0 * r));
  const double Pb = P + S;
  const double dP_dE = g - 1.0;
  const double dP_dm = -(g - 1.0) * u;
  const double dP_dr = (g - 1.0) * (u * u) / 2.0;
  f[0] = a[0] * (2.0 * u + dP_dm) + a[1] + a[2] * ((E + Pb) / r + dP_dm * u);
  f[1] = a[0] * (-(u * u) + dP_dr) + a[2] * (dP_dr * u - (E + Pb) * u / r);
  f[2] = a[0] * dP_dE + a[2] * (1.0 + dP_dE) * u;
}
// compute_adjoint_source_local! at one node (adjoint_pde.jl:128-171);
// d = (mu_x, rho_x, E_x, Sigma_x)
__device__ __forceinline__ void adj_source_node1(double g, const double q[4], const double d[4], const double a[3],
                                                 double Sl[3]) {
  const double m = q[0], r = q[1], En = q[2];
  const double mx = d[0], rx = d[1], Ex = d[2], Sx = d[3];
  const double u = m / r;
  const double P = (g - 1.0) * (En - m * m / (2.0 * r));
  const double Pb = P + q[3];
  const double ux = (mx - u * rx) / r;
  const double dP_dE = g - 1.0;
  const double dP_dm = -(g - 1.0) * u;
  const double dP_dr = (g - 1.0) * (u * u) / 2.0;
  const double P_x = (g - 1.0) * (Ex - m * mx / r + (m * m) * rx / (2.0 * (r * r)));
  const double drP_x = (g - 1.0) * (m * mx / (r * r) - (m * m) * rx / (r * r * r));
  const double dmP_x = -(g - 1.0) * ux;
  const double Pb_x = P_x + Sx;
  const double dm_u = 1.0 / r;
  const double dm_ux = -rx / (r * r);
  const double dr_u = -u / r;
  const double dr_ux = -mx / (r * r) + 2.0 * m * rx / (r * r * r);
  const double dm_divFm = mx * dm_u + ux + m * dm_ux + dmP_x;
  const double dm_divFE = dmP_x * u + (Ex + Pb_x) * dm_u + dP_dm * ux + (En + Pb) * dm_ux;
  const double dr_divFm = mx * dr_u + m * dr_ux + drP_x;
  const double dr_divFE = drP_x * u + (Ex + Pb_x) * dr_u + dP_dr * ux + (En + Pb) * dr_ux;
  Sl[0] = a[0] * dm_divFm + a[2] * dm_divFE;
  Sl[1] = a[0] * dr_divFm + a[2] * dr_divFE;
  Sl[2] = a[2] * ((1.0 + dP_dE) * ux);
}
#pragma clang fp contract(fast)

template <int NP>
struct K1Pde {
  const K1PdeArgs& P;
  const K1Args& a;
  K1<NP> k;
  const int N, Ne, tid, nt;
  __device__ K1Pde(const K1PdeArgs& args)
      : P(args), a(args.f), k(args.f), N(args.f.Ne * NP), Ne(args.f.Ne), tid(threadIdx.x), nt(blockDim.x) {}
  __device__ double D(int i, int j) const { return a.B.D[i * NP + j]; }

  // terminal condition at record n_last (adjoint_pde.jl:479-513)
  __device__ void terminal() {
    __shared__ double sred[40];   // block_sum<2> needs 2 * 16 waves; totals at 36, 37
    const double* q = P.snaps + (size_t)P.n_last * 4 * N;
    const double g = a.gamma;
    double KE = 0.0, Mt = 0.0;
    if (P.obj == 2) {
      double s[2] = {0.0, 0.0};
      for (int n = tid; n < N; n += nt) {
        const double wJ = a.B.w[n % NP] * a.J;
        s[0] += wJ * (q[n] * q[n]) / (2.0 * q[N + n]);
        s[1] += wJ * q[N + n];
      }
      block_sum<2>(s, sred);
      if (tid == 0) { sred[36] = s[0]; sred[37] = s[1]; }
      __syncthreads();
      KE = sred[36]; Mt = sred[37];
    }
    const double Ld = Ne * a.dx;
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      const double m = q[n], r = q[N + n], E = q[2 * N + n];
      double o[3] = {0.0, 0.0, 0.0};
      if (P.obj == 0) {
        o[1] = 2.0 * r;
      } else if (P.obj == 1) {
        const double u = m / r;
        const double Pp = (g - 1.0) * (E - (m * m) / (2.0 * r));
        o[0] = -2.0 * (g - 1.0) * Pp * u;
        o[1] = (g - 1.0) * Pp * (u * u);
        o[2] = 2.0 * (g - 1.0) * Pp;
      } else if (P.obj == 4) {
        const double x = e * a.dx + a.J * (1.0 + a.B.xi[i]);
        o[0] = sin(2.0 * M_PI * x / Ld);
      } else {
        o[0] = (m / r) / Mt;
        o[1] = -(m * m) / (2.0 * (r * r)) / Mt - KE / (Mt * Mt);
      }
      P.A[n] = o[0]; P.A[N + n] = o[1]; P.A[2 * N + n] = o[2];
      P.Pi[n] = 0.0;
    }
    __syncthreads();
  }

  // b = alpha w J (d a_mu/dx + d a_E/dx u) from stage input src (adjoint_pde.jl:278-301, 423-426)
  __device__ void pi_rhs(const double* q, const double* src) {
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      double dam = 0.0, daE = 0.0;
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        dam += D(i, j) * src[e * NP + j];
        daE += D(i, j) * src[2 * N + e * NP + j];
      }
      dam *= a.invJ; daE *= a.invJ;
      const double u = q[n] / q[N + n];
      const double bb = dam + daE * u;
      P.bPi[n] = bb * (a.alpha * a.B.w[i] * a.J);
    }
    __syncthreads();
  }

  // adjoint operator (adjoint_pde.jl:406-451) on src, fused with the stage update into out
  __device__ void stage(const double* q, const double* src, int mode, double cdt, double* out) {
    const double g = a.gamma, invJ = a.invJ;
    const double* S = q + 3 * N;
    const bool nl = a.alpha > 0.0;
    double* Pux = a.u;   // free scratch: Pi u_x and Pi Sigma_x / rho
    double* PSx = a.g;
    for (int n = tid; n < N; n += nt) {
      const double qq[4] = {q[n], q[N + n], q[2 * N + n], S[n]};
      const double aa[3] = {src[n], src[N + n], src[2 * N + n]};
      double f[3];
      adj_flux1(g, qq, aa, f);
      a.F[n] = f[0]; a.F[N + n] = f[1]; a.F[2 * N + n] = f[2];
      const int i = n % NP, e = n / NP;
      if (i == 0) {   // adjoint LLF at face e, between e-1 (L) and e (R), adjoint_pde.jl:75-88
        const int eL = e == 0 ? Ne - 1 : e - 1;
        const int nL = eL * NP + NP - 1;
        const double qL[4] = {q[nL], q[N + nL], q[2 * N + nL], S[nL]};
        const double aL[3] = {src[nL], src[N + nL], src[2 * N + nL]};
        const double lam = fmax_first(wave1(g, qL, a.flag), wave1(g, qq, a.flag));
        double fL[3];
        adj_flux1(g, qL, aL, fL);
#pragma unroll
        for (int v = 0; v < 3; ++v) a.fstar[v * Ne + e] = (fL[v] + f[v]) / 2.0 - lam * (aL[v] - aa[v]) / 2.0;
      }
      if (nl) {   // first pass of the nonlocal source (adjoint_pde.jl:331-347)
        double ux = 0.0, Sx = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          const int nj = e * NP + j;
          ux += D(i, j) * (q[nj] / q[N + nj]);
          Sx += D(i, j) * S[nj];
        }
        ux *= invJ; Sx *= invJ;
        Pux[n] = P.Pi[n] * ux;
        PSx[n] = P.Pi[n] * Sx / q[N + n];
      }
    }
    __syncthreads();
    for (int n = tid; n < N; n += nt) {
      const int i = n % NP, e = n / NP;
      const int eR = e == Ne - 1 ? 0 : e + 1;
      double acc[3] = {0.0, 0.0, 0.0}, d[4] = {0.0, 0.0, 0.0, 0.0}, ux = 0.0, dPux = 0.0, dPSx = 0.0;
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const int nj = e * NP + j;
        const double Dij = D(i, j);
#pragma unroll
        for (int v = 0; v < 3; ++v) acc[v] += invJ * Dij * a.F[v * N + nj];
        d[0] += Dij * q[nj]; d[1] += Dij * q[N + nj]; d[2] += Dij * q[2 * N + nj]; d[3] += Dij * S[nj];
        if (nl) {
          ux += Dij * (q[nj] / q[N + nj]);
          dPux += Dij * Pux[nj];
          dPSx += Dij * PSx[nj];
        }
      }
      if (i == NP - 1)
#pragma unroll
        for (int v = 0; v < 3; ++v) acc[v] += invJ * (a.fstar[v * Ne + eR] - a.F[v * N + n]) / a.B.w[NP - 1];
      if (i == 0)
#pragma unroll
        for (int v = 0; v < 3; ++v) acc[v] -= invJ * (a.fstar[v * Ne + e] - a.F[v * N + n]) / a.B.w[0];
#pragma unroll
      for (int m = 0; m < 4; ++m) d[m] *= invJ;
      const double qq[4] = {q[n], q[N + n], q[2 * N + n], S[n]};
      const double aa[3] = {src[n], src[N + n], src[2 * N + n]};
      double Sl[3];
      adj_source_node1(g, qq, d, aa, Sl);
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[v] -= Sl[v];
      if (nl) {   // second pass (adjoint_pde.jl:350-384)
        dPux *= invJ; dPSx *= invJ; ux *= invJ;
        const double r = qq[1], m = qq[0];
        acc[0] += -4.0 * dPux / r;
        acc[1] += (2.0 * P.Pi[n] * (ux * ux) + dPSx) / r + 4.0 * dPux * m / (r * r);
      }
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        const size_t gi = (size_t)v * N + n;
        double o;
        if (mode == 1) o = aa[v] + cdt * acc[v];
        else if (mode == 2) o = 3.0 / 4.0 * P.A[gi] + 1.0 / 4.0 * aa[v] + cdt * acc[v];
        else o = 1.0 / 3.0 * P.A[gi] + 2.0 / 3.0 * aa[v] + cdt * acc[v];
        out[gi] = o;
      }
    }
    __syncthreads();
  }

  __device__ void run() {
    terminal();
    const double h = P.h;
    for (int n = P.n_steps; n >= 1; --n) {
      const double* q = P.snaps + (size_t)n * 4 * N;
      if (a.alpha > 0.0) k.prepare(q);   // 1/rho and diag(A) of this snapshot (the IGR source is unused)
      const double* st[3] = {P.A, P.A1, P.A2};
      double* ou[3] = {P.A1, P.A2, P.A};
      const double cd[3] = {h, 1.0 / 4.0 * h, 2.0 / 3.0 * h};
      for (int s = 0; s < 3; ++s) {
        if (a.alpha > 0.0) {
          pi_rhs(q, st[s]);
          k.iterate(P.Pi, P.bPi, 0);
        }
        stage(q, st[s], s + 1, cd[s], ou[s]);
      }
    }
  }
};

template <int NP>
__global__ __launch_bounds__(1024) void k1_adjpde(K1PdeArgs a) {
  K1Pde<NP> w(a);
  w.run();
}

void launch1_adjpde(const Ctx& c, const K1PdeArgs& a, int block) {
  switch (c.np) {
    case 2: hipLaunchKernelGGL(k1_adjpde<2>, dim3(1), dim3(block), 0, c.stream, a); break;
    case 3: hipLaunchKernelGGL(k1_adjpde<3>, dim3(1), dim3(block), 0, c.stream, a); break;
    case 4: hipLaunchKernelGGL(k1_adjpde<4>, dim3(1), dim3(block), 0, c.stream, a); break;
    case 5: hipLaunchKernelGGL(k1_adjpde<5>, dim3(1), dim3(block), 0, c.stream, a); break;
  }
}

}  // namespace igr
