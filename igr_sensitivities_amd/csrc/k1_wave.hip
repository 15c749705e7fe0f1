// 1D runs of one small instance in one wavefront, with no workgroup barrier.
//
// An instance of N_e = 64 EPL elements (C1, the 1D blast: 128 elements at
// p = 3) is held by the 64 lanes of one wave: lane l owns elements
// l EPL .. l EPL + EPL - 1, NP nodes each, in registers.  Everything inside
// an element (derivatives, the IGR source, the diagonal, the volume terms)
// is lane-local with compile-time basis entries; the two faces at the ends
// of a lane's block take the neighbour lane's end-node values through DPP
// wave rotates (wave_ror:1 from lane l - 1, wave_rol:1 from lane l + 1, both
// periodic over the 64 lanes, which is the periodic mesh).  A face between
// two lanes is evaluated by both of them with the same operands in the same
// order, so both get the same bits.  Dot products (PCG) are DPP wave sums.
// The stage state and the solver vectors stay in registers, q^n of the step
// in lane-private LDS slots; no barrier is ever needed.
//
// The workgroup-per-instance kernel (k1_forward.hip) spends a step of this
// instance in ~330 barrier phases; this one has none.
//
// Follows IGR_adjoints_1d/: hyperbolic.jl:23-83 (RHS), elliptic.jl:32-116
// (apply_sip!), :126-154 (compute_igr_rhs!), :206-253 (compute_sip_diagonal),
// :307-430 (pcg/jacobi/chebyshev), forward.jl:50-87 (run_forward!).
#include "igr_ctx.hpp"
#include "k1_common.hpp"

namespace igr {

namespace {
// v of the neighbour lane (DPP wave rotate by one lane, both 32-bit halves)
template <int CTRL>
__device__ __forceinline__ double wave_rot(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double from_left(double v) { return wave_rot<0x13C>(v); }    // wave_ror:1: lane l - 1
__device__ __forceinline__ double from_right(double v) { return wave_rot<0x134>(v); }   // wave_rol:1: lane l + 1
}  // namespace

#pragma clang fp contract(off)

template <int NP, int EPL>
struct K1W {
  static constexpr int M = EPL * NP;   // nodes per lane
  const K1Args& a;
  double ir[M], im[M], b[M];            // 1/rho, 1/diag(A), right-hand side of the solve
  double fsig[EPL + 1];                 // penalty of face f (f < EPL: left face of element f; EPL: right end)

  __device__ __forceinline__ explicit K1W(const K1Args& args) : a(args) {}
  __device__ __forceinline__ double D(int i, int k) const { return a.B.D[i * NP + k]; }
  __device__ __forceinline__ double W(int i) const { return a.B.w[i]; }

  // 1/rho, the IGR source and diag(A) of state (mom, rho)
  // (compute_igr_rhs!, elliptic.jl:126-154; compute_sip_diagonal, :206-253)
  __device__ __forceinline__ void prepare(const double (&mom)[M], const double (&rho)[M]) {
    const double al = a.alpha, invJ = a.invJ, J = a.J;
    double u[M];
#pragma unroll
    for (int m = 0; m < M; ++m) { ir[m] = 1.0 / rho[m]; u[m] = mom[m] / rho[m]; }
    const double irL = from_left(ir[M - 1]), irR = from_right(ir[0]);
#pragma unroll
    for (int f = 0; f <= EPL; ++f) {
      const double l = f == 0 ? irL : ir[f * NP - 1];
      const double r = f == EPL ? irR : ir[f * NP];
      fsig[f] = a.eta * fmax(l, r) / a.dx;
    }
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const int m = e * NP + i;
        double du = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) du = fma(D(i, k), u[e * NP + k], du);
        du *= invJ;
        b[m] = W(i) * J * (2.0 * al * (du * du));
        double Md = W(i) * J * ir[m];
#pragma unroll
        for (int k = 0; k < NP; ++k) Md = fma(al * W(k) * invJ * (D(k, i) * D(k, i)), ir[e * NP + k], Md);
        if (i == NP - 1) {
          Md += -al * invJ * D(NP - 1, NP - 1) * ir[m];
          Md += al * fsig[e + 1];
        }
        if (i == 0) {
          Md += al * invJ * D(0, 0) * ir[m];
          Md += al * fsig[e];
        }
        im[m] = 1.0 / Md;
      }
    }
  }

  // v = A x (apply_sip!, elliptic.jl:32-116)
  __device__ __forceinline__ void apply(const double (&x)[M], double (&v)[M]) const {
    const double al = a.alpha, invJ = a.invJ, J = a.J;
    double d[M], g[M];
#pragma unroll
    for (int e = 0; e < EPL; ++e)
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) s = fma(D(i, k), x[e * NP + k], s);
        d[e * NP + i] = s;
        g[e * NP + i] = al * W(i) * invJ * s * ir[e * NP + i];
      }
    // faces: value jump and average flux (f < EPL: left face of element f)
    const double xL = from_left(x[M - 1]), dL = from_left(d[M - 1]), irL = from_left(ir[M - 1]);
    const double xR = from_right(x[0]), dR = from_right(d[0]), irR = from_right(ir[0]);
    double fjump[EPL + 1], favg[EPL + 1];
#pragma unroll
    for (int f = 0; f <= EPL; ++f) {
      const double xl = f == 0 ? xL : x[f * NP - 1], dl = f == 0 ? dL : d[f * NP - 1];
      const double il = f == 0 ? irL : ir[f * NP - 1];
      const double xr = f == EPL ? xR : x[f * NP], dr = f == EPL ? dR : d[f * NP];
      const double irr = f == EPL ? irR : ir[f * NP];
      fjump[f] = xl - xr;
      favg[f] = invJ * fma(il, dl, irr * dr) / 2.0;
    }
#pragma unroll
    for (int e = 0; e < EPL; ++e)
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const int m = e * NP + i;
        double y = W(i) * J * x[m] * ir[m];
        double st = 0.0;
#pragma unroll
        for (int k = 0; k < NP; ++k) st = fma(D(k, i), g[e * NP + k], st);
        y += st;
        // symmetry terms of the left face (this element is R) and the right face (L)
        y += -al * (ir[e * NP] / 2.0) * invJ * D(0, i) * fjump[e];
        y += -al * (ir[e * NP + NP - 1] / 2.0) * invJ * D(NP - 1, i) * fjump[e + 1];
        if (i == 0) y += al * favg[e] - al * fsig[e] * fjump[e];
        if (i == NP - 1) y += -al * favg[e + 1] + al * fsig[e + 1] * fjump[e + 1];
        v[m] = y;
      }
  }

  __device__ __forceinline__ static double wave_total(double s) {
    double t[1] = {s};
    wave_sums_last<1>(t);
    return read_lane(t[0], WAVE - 1);
  }

  // fixed-iteration solve A x = b (pcg_fixed! :307-339, jacobi_fixed! :359-368,
  // chebyshev_fixed! :389-430), warm start x
  __device__ __forceinline__ void iterate(double (&x)[M]) const {
    const int nI = a.n_iter;
    double v[M];
    if (a.solver == 0) {
      double r[M], p[M];
      apply(x, v);
      double part = 0.0;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        r[m] = b[m] - v[m];
        p[m] = r[m] * im[m];
        part += r[m] * p[m];
      }
      double rz = wave_total(part);
      for (int it = 0; it < nI; ++it) {
        apply(p, v);
        double pp = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) pp += p[m] * v[m];
        const double ak = rz / (wave_total(pp) + 1e-30);
        double rr = 0.0, z[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
          x[m] = x[m] + ak * p[m];
          r[m] = r[m] - ak * v[m];
          z[m] = r[m] * im[m];
          rr += r[m] * z[m];
        }
        const double rzn = wave_total(rr);
        const double bk = rzn / (rz + 1e-30);
#pragma unroll
        for (int m = 0; m < M; ++m) p[m] = z[m] + bk * p[m];
        rz = rzn;
      }
    } else if (a.solver == 1) {
      for (int it = 0; it < nI; ++it) {
        apply(x, v);
#pragma unroll
        for (int m = 0; m < M; ++m) x[m] = x[m] + (2.0 / 3.0) * (b[m] - v[m]) * im[m];
      }
    } else {
      const double lmin = 1.0, lmax = a.cheb_lmax;
      const double theta = (lmax + lmin) / 2.0, delta = (lmax - lmin) / 2.0;
      const double sigma = theta / fmax(delta, 1e-10);
      double r[M], dd[M];
      apply(x, v);
#pragma unroll
      for (int m = 0; m < M; ++m) { r[m] = b[m] - v[m]; dd[m] = 0.0; }
      double rc = 1.0 / sigma;
      for (int k = 1; k <= nI; ++k) {
        double c1 = 0.0, c2 = 0.0;
        if (k > 1) { const double rn = 1.0 / (2.0 * sigma - rc); c1 = rn * rc; c2 = 2.0 * rn / delta; rc = rn; }
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const double z = r[m] * im[m];
          dd[m] = (k == 1) ? z / theta : c1 * dd[m] + c2 * z;
          x[m] = x[m] + dd[m];
        }
        apply(x, v);
#pragma unroll
        for (int m = 0; m < M; ++m) r[m] = b[m] - v[m];
      }
    }
  }

  // hyperbolic RHS (hyperbolic.jl:23-83) fused with the SSP-RK3 stage
  // update, in place: qs <- stage(qs, q^n) (mode 1: qs + c dt; 2: 3/4 q^n +
  // 1/4 qs + c dt; 3: 1/3 q^n + 2/3 qs + c dt), q^n in this lane's LDS slots
  __device__ __forceinline__ void hyp_stage(double (&qs)[3][M], const double (&S)[M], const double (*qn)[M][WAVE],
                                            int mode, double cdt) const {
    const int lane = threadIdx.x;
    const double g = a.gamma, invJ = a.invJ;
    double F[3][M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      double f[3];
      flux1(g, qs[0][m], qs[1][m], qs[2][m], S[m], f);
      F[0][m] = f[0]; F[1][m] = f[1]; F[2][m] = f[2];
    }
    // LLF at the faces (llf_flux, euler.jl:64-78); the end faces take the
    // neighbour lane's end node
    double eL[4], eR[4];
#pragma unroll
    for (int v = 0; v < 3; ++v) { eL[v] = from_left(qs[v][M - 1]); eR[v] = from_right(qs[v][0]); }
    eL[3] = from_left(S[M - 1]);
    eR[3] = from_right(S[0]);
    double fstar[3][EPL + 1];
#pragma unroll
    for (int f = 0; f <= EPL; ++f) {
      double qL[4], qR[4];
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        qL[v] = f == 0 ? eL[v] : qs[v][f * NP - 1];
        qR[v] = f == EPL ? eR[v] : qs[v][f * NP];
      }
      qL[3] = f == 0 ? eL[3] : S[f * NP - 1];
      qR[3] = f == EPL ? eR[3] : S[f * NP];
      const double lam = fmax_first(wave1(g, qL, a.flag), wave1(g, qR, a.flag));
      double fL[3], fR[3];
      flux1(g, qL[0], qL[1], qL[2], qL[3], fL);
      flux1(g, qR[0], qR[1], qR[2], qR[3], fR);
#pragma unroll
      for (int v = 0; v < 3; ++v) fstar[v][f] = (fL[v] + fR[v]) / 2.0 + lam * (qL[v] - qR[v]) / 2.0;
    }
#pragma unroll
    for (int e = 0; e < EPL; ++e)
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const int m = e * NP + i;
        double acc[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          acc[v] = 0.0;
#pragma unroll
          for (int k = 0; k < NP; ++k) acc[v] -= invJ * D(i, k) * F[v][e * NP + k];
        }
        if (i == NP - 1)
#pragma unroll
          for (int v = 0; v < 3; ++v) acc[v] -= invJ * (fstar[v][e + 1] - F[v][m]) / W(NP - 1);
        if (i == 0)
#pragma unroll
          for (int v = 0; v < 3; ++v) acc[v] += invJ * (fstar[v][e] - F[v][m]) / W(0);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          if (mode == 1) qs[v][m] = qs[v][m] + cdt * acc[v];
          else if (mode == 2) qs[v][m] = 3.0 / 4.0 * qn[v][m][lane] + 1.0 / 4.0 * qs[v][m] + cdt * acc[v];
          else qs[v][m] = 1.0 / 3.0 * qn[v][m][lane] + 2.0 / 3.0 * qs[v][m] + cdt * acc[v];
        }
      }
  }
};

// OP_RUN of one instance without tangents or stored iterates: Sigma <- 0 (if
// zero_sigma), n_steps steps of dt then n_steps2 of dt2, optional final solve.
template <int NP, int EPL>
__global__ __launch_bounds__(64) void k1_wave_run(const K1Args* __restrict__ ap) {
  constexpr int M = EPL * NP;
  const K1Args& a = *ap;
  const int lane = threadIdx.x;
  const size_t N = (size_t)a.Ne * NP;
  const size_t n0 = (size_t)lane * M;
  K1W<NP, EPL> k(a);
  // q^n of the current step in lane-private LDS slots [v][m][lane] (lane
  // fastest: no bank conflicts); the stage state in registers
  __shared__ double qn[3][M][WAVE];
  double qs[3][M], S[M];
#pragma unroll
  for (int v = 0; v < 3; ++v)
#pragma unroll
    for (int m = 0; m < M; ++m) qs[v][m] = a.q[v * N + n0 + m];
#pragma unroll
  for (int m = 0; m < M; ++m) S[m] = a.zero_sigma ? 0.0 : a.S[n0 + m];
  auto solve = [&]() {
    if (a.alpha > 0.0) {
      k.prepare(qs[0], qs[1]);
      k.iterate(S);
    } else {
#pragma unroll
      for (int m = 0; m < M; ++m) S[m] = 0.0;
    }
  };
  const int total = a.n_steps + a.n_steps2;
  for (int step = 0; step < total; ++step) {
    const double dt = step < a.n_steps ? a.dt : a.dt2;
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
      for (int m = 0; m < M; ++m) qn[v][m][lane] = qs[v][m];
    solve();                                        // stage 1: q -> q1
    k.hyp_stage(qs, S, qn, 1, dt);
    solve();                                        // stage 2: q1 -> q2
    k.hyp_stage(qs, S, qn, 2, 1.0 / 4.0 * dt);
    solve();                                        // stage 3: q2 -> q
    k.hyp_stage(qs, S, qn, 3, 2.0 / 3.0 * dt);
  }
  if (a.final_solve && a.alpha > 0.0) solve();
#pragma unroll
  for (int v = 0; v < 3; ++v)
#pragma unroll
    for (int m = 0; m < M; ++m) a.q[v * N + n0 + m] = qs[v][m];
#pragma unroll
  for (int m = 0; m < M; ++m) a.S[n0 + m] = S[m];
}

#pragma clang fp contract(fast)

// Elements per lane of a one-wave run, 0 when the instance does not qualify
// (a plain run without tangents or stored iterates, N_e = 64, or N_e = 128
// up to p = 3: at p = 4 two elements per lane exceed the register file).
int k1_wave_epl(const K1Args& a, int np) {
  if (a.op != OP_RUN || a.n_tan != 0 || a.stP || a.stAP || a.stR || a.stsc || a.trace) return 0;
  if (a.Ne == 64) return 1;
  if (a.Ne == 128 && np <= 4) return 2;
  return 0;
}

void launch1_wave(hipStream_t st, int np, int epl, const K1Args* d) {
#define IGR_K1W(NPv)                                                                                   \
  case NPv:                                                                                            \
    if (epl == 1) hipLaunchKernelGGL((k1_wave_run<NPv, 1>), dim3(1), dim3(WAVE), 0, st, d);            \
    else if (NPv <= 4) hipLaunchKernelGGL((k1_wave_run<(NPv <= 4 ? NPv : 4), 2>), dim3(1), dim3(WAVE), 0, st, d); \
    break;
  switch (np) {
    IGR_K1W(2)
    IGR_K1W(3)
    IGR_K1W(4)
    IGR_K1W(5)
  }
#undef IGR_K1W
}

}  // namespace igr
