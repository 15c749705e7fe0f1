// 1D forward path for gfx950, FP64: one workgroup runs a whole 1D
// instance — every SSP-RK3 stage, its Sigma solve and the stage update —
// inside a single launch, synchronising with workgroup barriers only.
//
// 1D meshes in the reference's configurations are small (512 to 16384
// nodes), so a grid-wide launch per solver iteration would be pure launch
// latency; one persistent workgroup per instance keeps the working set in
// L1/L2 and needs no grid synchronisation.  Batches of instances run one
// workgroup each in the same launch.
//
// Follows IGR_adjoints_1d/: hyperbolic.jl:23-83 (RHS), elliptic.jl:32-116
// (apply_sip!), :126-154 (compute_igr_rhs!), :206-253 (compute_sip_diagonal),
// :307-430 (pcg/jacobi/chebyshev), forward.jl:50-87 (run_forward!).
#include "igr_ctx.hpp"
#include "k1_common.hpp"
#include "k1_forward.hpp"

namespace igr {

// Whole-run kernel: `op` selects what one launch does.
//  OP_RUN       Sigma <- 0 (if zero_sigma), n_steps SSP-RK3 steps, optional final solve
//  OP_SOLVE     one Sigma solve on the state
//  OP_RHS       compute_igr_euler_rhs! into a.dq
//  OP_HYP       hyperbolic RHS only (Sigma given) into a.dq
//  OP_APPLY     y = A(rho) x with rho = state rho
//  OP_STORE     run_forward_store (forward.jl:137-187): Sigma <- 0, a solve
//               on the initial state, then n_steps steps each followed by a
//               solve; (state, Sigma) stored after the initial solve and
//               after every step
template <int NP, bool LDS>
__device__ __forceinline__ void k1_ops(const K1Args& a, K1<NP, LDS>& k) {
  const int N = k.N;
  // the stage buffers as values: a stage-indexed choice between members of
  // `k` would become a load at a variable offset into it, which keeps the
  // whole solver object in per-lane scratch memory
  double* const Q0 = k.w.q;
  double* const Q1 = k.w.q1;
  double* const Q2 = k.w.q2;
  if (a.op == OP_STORE) {
    for (int n = k.tid; n < N; n += k.nt) k.w.S[n] = 0.0;
    __syncthreads();
    if (a.alpha > 0.0) k.solve(k.w.q, 0);
    for (int step = 0; step <= a.n_steps; ++step) {
      if (step > 0) {
        for (int s = 0; s < 3; ++s) {
          // stage s: q -> q1, q1 -> q2, q2 -> q (selects, not arrays: no scratch)
          const double* st = s == 0 ? Q0 : (s == 1 ? Q1 : Q2);
          double* ou = s == 0 ? Q1 : (s == 1 ? Q2 : Q0);
          const double cd = s == 0 ? a.dt : (s == 1 ? 1.0 / 4.0 * a.dt : 2.0 / 3.0 * a.dt);
          if (a.alpha > 0.0) k.solve(st, 0);
          else { for (int n = k.tid; n < N; n += k.nt) k.w.S[n] = 0.0; __syncthreads(); }
          k.hyp_stage(st, s + 1, cd, ou);
        }
        if (a.alpha > 0.0) k.solve(k.w.q, 0);
      }
      double* d = a.snaps + (size_t)step * 4 * N;   // record `step`: after the initial solve / after step `step`
      for (int n = k.tid; n < 3 * N; n += k.nt) d[n] = k.w.q[n];
      for (int n = k.tid; n < N; n += k.nt) d[3 * N + n] = k.w.S[n];
    }
    return;
  }
  if (a.op == OP_APPLY) {
    for (int n = k.tid; n < N; n += k.nt) k.w.invR[n] = 1.0 / k.w.q[N + n];
    __syncthreads();
    k.apply(a.x_in, a.y_out);
    return;
  }
  if (a.op == OP_SOLVE) { k.solve(k.w.q, 1); return; }
  if (a.op == OP_RHS || a.op == OP_HYP) {
    if (a.op == OP_RHS) {
      if (a.alpha > 0.0) k.solve(k.w.q, 0);
      else { for (int n = k.tid; n < N; n += k.nt) k.w.S[n] = 0.0; __syncthreads(); }
    }
    k.hyp_stage(k.w.q, 0, 0.0, a.dq);
    return;
  }
  // OP_RUN
  if (a.zero_sigma) {
    for (int n = k.tid; n < N; n += k.nt) {
      k.w.S[n] = 0.0;
      for (int t = 0; t < a.n_tan; ++t) a.tvs[t][n] = 0.0;
    }
    __syncthreads();
  }
  for (int step = 0; step < a.n_steps + a.n_steps2; ++step) {
    const double dt = step < a.n_steps ? a.dt : a.dt2;
    for (int s = 0; s < 3; ++s) {
      const double* st = s == 0 ? Q0 : (s == 1 ? Q1 : Q2);
      double* ou = s == 0 ? Q1 : (s == 1 ? Q2 : Q0);
      const double cd = s == 0 ? dt : (s == 1 ? 1.0 / 4.0 * dt : 2.0 / 3.0 * dt);
      if (a.alpha > 0.0) k.solve(st, 0);
      else { for (int n = k.tid; n < N; n += k.nt) k.w.S[n] = 0.0; __syncthreads(); }
      for (int t = 0; t < a.n_tan; ++t) {
        const double* ss = s == 0 ? a.ts[t] : (s == 1 ? a.ts1[t] : a.ts2[t]);
        double* so = s == 0 ? a.ts1[t] : (s == 1 ? a.ts2[t] : a.ts[t]);
        if (a.alpha > 0.0) { k.sens_rhs(st, ss); k.iterate(a.tvs[t], k.w.bs, 0); }
        else { for (int n = k.tid; n < N; n += k.nt) a.tvs[t][n] = 0.0; __syncthreads(); }
        k.sens_hyp_stage(st, ss, a.tvs[t], a.ts[t], s + 1, cd, so);
      }
      k.hyp_stage(st, s + 1, cd, ou);
    }
  }
  if (a.final_solve && a.alpha > 0.0) {
    k.solve(k.w.q, 0);
    for (int t = 0; t < a.n_tan; ++t) { k.sens_rhs(k.w.q, a.ts[t]); k.iterate(a.tvs[t], k.w.bs, 0); }
  }
}

// a.lds: the run keeps its state, stage buffers and scratch in dynamic LDS.
// A 1D instance of the reference's sizes (C1: 512 nodes, 100 KB) fits, and
// every solver phase then waits on LDS instead of L2 between its workgroup
// barriers.  The state and Sigma are copied in first and back at the end;
// tangents, snapshots and traces stay in global memory.  ALLOW_LDS = false
// compiles the global-memory form only (the batch kernel never uses LDS).
template <int NP, bool ALLOW_LDS>
__device__ __forceinline__ void k1_body(const K1Args& a) {
  if constexpr (ALLOW_LDS) {
    if (a.lds) {
      extern __shared__ double k1_lds[];
      K1<NP, true> k(a, k1_lds);
      for (int n = threadIdx.x; n < 3 * k.N; n += blockDim.x) k.w.q[n] = a.q[n];
      for (int n = threadIdx.x; n < k.N; n += blockDim.x) k.w.S[n] = a.S[n];
      __syncthreads();
      k1_ops<NP>(a, k);
      __syncthreads();
      for (int n = threadIdx.x; n < 3 * k.N; n += blockDim.x) a.q[n] = k.w.q[n];
      for (int n = threadIdx.x; n < k.N; n += blockDim.x) a.S[n] = k.w.S[n];
      return;
    }
  }
  K1<NP, false> k(a);
  k1_ops<NP>(a, k);
}

// Single instance: the argument block is read from device memory.  (Taken by
// value, it would be copied to per-lane scratch: the run keeps references to
// it, and the kernel-argument segment has no generic address.)
template <int NP>
__global__ __launch_bounds__(1024) void k1_run(const K1Args* __restrict__ a) { k1_body<NP, true>(*a); }
// Batch: the device array is the work list (instances sorted by decreasing
// cost); workgroup b runs entries b, b + gridDim.x, ... one after another.
template <int NP>
__global__ __launch_bounds__(1024) void k1_run_batch(const K1Args* __restrict__ args, int n) {
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    k1_body<NP, false>(args[k]);
    __syncthreads();
  }
}

// Dynamic LDS of an LDS-resident run, kept below the 160 KB of a CU with room
// for the kernel's static LDS.
constexpr size_t K1_LDS_MAX_BYTES = 150 * 1024;

// One-wave runs of small instances (k1_wave.hip)
int k1_wave_epl(const K1Args& a, int np);
void launch1_wave(hipStream_t st, int np, int epl, const K1Args* d);

template <int NP>
static void launch1_run_np(const Ctx& c, K1Args a, int block) {
  K1Args* d = reinterpret_cast<K1Args*>(c.k1args);
  if (c.k1_wave) {
    const int epl = k1_wave_epl(a, NP);
    if (epl > 0) {
      (void)hipMemcpyAsync(d, &a, sizeof(K1Args), hipMemcpyHostToDevice, c.stream);
      launch1_wave(c.stream, NP, epl, d);
      return;
    }
  }
  const size_t bytes = k1_lds_doubles(a.Ne, NP) * sizeof(double);
  size_t smem = 0;
  if ((a.op == OP_RUN || a.op == OP_STORE) && bytes <= K1_LDS_MAX_BYTES) {
    // set on every launch (once per whole run): the attribute belongs to the
    // current device, and this keeps no state shared between host threads
    const bool attr = hipFuncSetAttribute(reinterpret_cast<const void*>(k1_run<NP>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)K1_LDS_MAX_BYTES) == hipSuccess;
    (void)hipGetLastError();
    if (attr) { a.lds = 1; smem = bytes; }
  }
  // ordered on the context stream after the previous launch that read the block
  (void)hipMemcpyAsync(d, &a, sizeof(K1Args), hipMemcpyHostToDevice, c.stream);
  hipLaunchKernelGGL(k1_run<NP>, dim3(1), dim3(block), smem, c.stream, static_cast<const K1Args*>(d));
}
void launch1_run(const Ctx& c, const K1Args& a, int block) {
  switch (c.np) {
    case 2: launch1_run_np<2>(c, a, block); break;
    case 3: launch1_run_np<3>(c, a, block); break;
    case 4: launch1_run_np<4>(c, a, block); break;
    case 5: launch1_run_np<5>(c, a, block); break;
  }
}
void launch1_batch(hipStream_t st, int np, const K1Args* dargs, int n_inst, int n_wg, int block) {
  switch (np) {
    case 2: hipLaunchKernelGGL(k1_run_batch<2>, dim3(n_wg), dim3(block), 0, st, dargs, n_inst); break;
    case 3: hipLaunchKernelGGL(k1_run_batch<3>, dim3(n_wg), dim3(block), 0, st, dargs, n_inst); break;
    case 4: hipLaunchKernelGGL(k1_run_batch<4>, dim3(n_wg), dim3(block), 0, st, dargs, n_inst); break;
    case 5: hipLaunchKernelGGL(k1_run_batch<5>, dim3(n_wg), dim3(block), 0, st, dargs, n_inst); break;
  }
}

}  // namespace igr
