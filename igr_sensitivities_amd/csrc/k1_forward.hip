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
template <int NP>
__device__ void k1_ops(const K1Args& a) {
  K1<NP> k(a);
  const int N = k.N;
  if (a.op == OP_STORE) {
    auto store = [&](int rec) {
      double* d = a.snaps + (size_t)rec * 4 * N;
      for (int n = k.tid; n < 3 * N; n += k.nt) d[n] = a.q[n];
      for (int n = k.tid; n < N; n += k.nt) d[3 * N + n] = a.S[n];
    };
    for (int n = k.tid; n < N; n += k.nt) a.S[n] = 0.0;
    __syncthreads();
    if (a.alpha > 0.0) k.solve(a.q, 0);
    store(0);
    for (int step = 0; step < a.n_steps; ++step) {
      const double* st[3] = {a.q, a.q1, a.q2};
      double* ou[3] = {a.q1, a.q2, a.q};
      const double cd[3] = {a.dt, 1.0 / 4.0 * a.dt, 2.0 / 3.0 * a.dt};
      for (int s = 0; s < 3; ++s) {
        if (a.alpha > 0.0) k.solve(st[s], 0);
        else { for (int n = k.tid; n < N; n += k.nt) a.S[n] = 0.0; __syncthreads(); }
        k.hyp_stage(st[s], s + 1, cd[s], ou[s]);
      }
      if (a.alpha > 0.0) k.solve(a.q, 0);
      store(step + 1);
    }
    return;
  }
  if (a.op == OP_APPLY) {
    for (int n = k.tid; n < N; n += k.nt) a.invR[n] = 1.0 / a.q[N + n];
    __syncthreads();
    k.apply(a.x_in, a.y_out);
    return;
  }
  if (a.op == OP_SOLVE) { k.solve(a.q, 1); return; }
  if (a.op == OP_RHS || a.op == OP_HYP) {
    if (a.op == OP_RHS) {
      if (a.alpha > 0.0) k.solve(a.q, 0);
      else { for (int n = k.tid; n < N; n += k.nt) a.S[n] = 0.0; __syncthreads(); }
    }
    k.hyp_stage(a.q, 0, 0.0, a.dq);
    return;
  }
  // OP_RUN
  if (a.zero_sigma) {
    for (int n = k.tid; n < N; n += k.nt) {
      a.S[n] = 0.0;
      for (int t = 0; t < a.n_tan; ++t) a.tvs[t][n] = 0.0;
    }
    __syncthreads();
  }
  for (int step = 0; step < a.n_steps + a.n_steps2; ++step) {
    const double dt = step < a.n_steps ? a.dt : a.dt2;
    const double* st[3] = {a.q, a.q1, a.q2};
    double* ou[3] = {a.q1, a.q2, a.q};
    const double cd[3] = {dt, 1.0 / 4.0 * dt, 2.0 / 3.0 * dt};
    for (int s = 0; s < 3; ++s) {
      if (a.alpha > 0.0) k.solve(st[s], 0);
      else { for (int n = k.tid; n < N; n += k.nt) a.S[n] = 0.0; __syncthreads(); }
      for (int t = 0; t < a.n_tan; ++t) {
        const double* ss = s == 0 ? a.ts[t] : (s == 1 ? a.ts1[t] : a.ts2[t]);
        double* so = s == 0 ? a.ts1[t] : (s == 1 ? a.ts2[t] : a.ts[t]);
        if (a.alpha > 0.0) { k.sens_rhs(st[s], ss); k.iterate(a.tvs[t], a.bs, 0); }
        else { for (int n = k.tid; n < N; n += k.nt) a.tvs[t][n] = 0.0; __syncthreads(); }
        k.sens_hyp_stage(st[s], ss, a.tvs[t], a.ts[t], s + 1, cd[s], so);
      }
      k.hyp_stage(st[s], s + 1, cd[s], ou[s]);
    }
  }
  if (a.final_solve && a.alpha > 0.0) {
    k.solve(a.q, 0);
    for (int t = 0; t < a.n_tan; ++t) { k.sens_rhs(a.q, a.ts[t]); k.iterate(a.tvs[t], a.bs, 0); }
  }
}

// a.lds: the run keeps its state, stage buffers and scratch in dynamic LDS.
// A 1D instance of the reference's sizes (C1: 512 nodes, 100 KB) fits, and
// every solver phase then waits on LDS instead of L2 between its workgroup
// barriers.  The state and Sigma are copied in first and back at the end;
// tangents, snapshots and traces stay in global memory.
template <int NP>
__device__ void k1_body(const K1Args& a) {
  if (!a.lds) { k1_ops<NP>(a); return; }
  extern __shared__ double k1_lds[];
  __shared__ __align__(16) unsigned char raw[sizeof(K1Args)];
  K1Args* la = reinterpret_cast<K1Args*>(raw);
  const int N = a.Ne * NP, Ne = a.Ne;
  if (threadIdx.x == 0) {
    __builtin_memcpy(raw, &a, sizeof(K1Args));
    double* p = k1_lds;
    la->q = p; p += 3 * N;
    la->q1 = p; p += 3 * N;
    la->q2 = p; p += 3 * N;
    la->S = p; p += N;
    la->invR = p; p += N;
    la->invM = p; p += N;
    la->b = p; p += N;
    la->r = p; p += N;
    la->p = p; p += N;
    la->z = p; p += N;
    la->Ap = p; p += N;
    la->u = p; p += N;
    la->g = p; p += N;
    la->bs = p; p += N;
    la->F = p; p += 3 * N;
    la->fstar = p; p += 3 * Ne;
    la->fjump = p; p += Ne;
    la->favg = p; p += Ne;
    la->fsig = p;
  }
  __syncthreads();
  for (int n = threadIdx.x; n < 3 * N; n += blockDim.x) la->q[n] = a.q[n];
  for (int n = threadIdx.x; n < N; n += blockDim.x) la->S[n] = a.S[n];
  __syncthreads();
  k1_ops<NP>(*la);
  __syncthreads();
  for (int n = threadIdx.x; n < 3 * N; n += blockDim.x) a.q[n] = la->q[n];
  for (int n = threadIdx.x; n < N; n += blockDim.x) a.S[n] = la->S[n];
}

// Single instance: arguments by value in the kernel-argument segment.
template <int NP>
__global__ __launch_bounds__(1024) void k1_run(K1Args a) { k1_body<NP>(a); }
// Batch: the device array is the work list (instances sorted by decreasing
// cost); workgroup b runs entries b, b + gridDim.x, ... one after another.
template <int NP>
__global__ __launch_bounds__(1024) void k1_run_batch(const K1Args* __restrict__ args, int n) {
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    k1_body<NP>(args[k]);
    __syncthreads();
  }
}

// Dynamic LDS of an LDS-resident run, kept below the 160 KB of a CU with room
// for the kernel's static LDS.
constexpr size_t K1_LDS_MAX_BYTES = 150 * 1024;

template <int NP>
static void launch1_run_np(const Ctx& c, K1Args a, int block) {
  const size_t bytes = k1_lds_doubles(a.Ne, NP) * sizeof(double);
  size_t smem = 0;
  if ((a.op == OP_RUN || a.op == OP_STORE) && bytes <= K1_LDS_MAX_BYTES) {
    static bool attr = false;
    if (!attr) {
      attr = hipFuncSetAttribute(reinterpret_cast<const void*>(k1_run<NP>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)K1_LDS_MAX_BYTES) == hipSuccess;
      (void)hipGetLastError();
    }
    if (attr) { a.lds = 1; smem = bytes; }
  }
  hipLaunchKernelGGL(k1_run<NP>, dim3(1), dim3(block), smem, c.stream, a);
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
