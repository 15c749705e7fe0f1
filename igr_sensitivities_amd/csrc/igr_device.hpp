// Device-side helpers shared by the 1D and 2D kernels: basis constants
// passed by value in the kernel-argument segment (scalar loads), wavefront
// and workgroup reductions, and the deterministic "last workgroup finishes
// the reduction" pattern used for grid-wide dot products.
#pragma once
#include <hip/hip_runtime.h>

namespace igr {

constexpr int MAXNP = 5;      // p <= 4
constexpr int WAVE = 64;      // CDNA wavefront width

struct BasisK {
  double D[MAXNP * MAXNP];    // D[i*NP + k] = D_{ik}, packed with stride NP
  double w[MAXNP];
  double xi[MAXNP];
};

// v taken from another lane by a DPP lane pattern (two 32-bit DPP moves).
// A lane whose source is out of range keeps an undefined value.
template <int CTRL>
__device__ __forceinline__ double dpp_take(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double read_lane(double v, int lane) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Store of a kernel output that the next kernel on the stream reads: a
// relaxed store at agent scope, which gfx950 writes through the XCD's L2
// (sc1) instead of leaving it dirty there for the end-of-kernel writeback
// that makes it visible to the other XCDs.  No ordering is implied; the
// kernel boundary orders it as before.  Measured on the forward step: 128^2
// 2.06 -> 1.98 ms, 256^2 4.71 -> 4.65 ms (non-temporal stores instead: 256^2
// 5.2 ms).
__device__ __forceinline__ void store_wt(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A wave-uniform double moved to scalar registers (the first active lane's
// value); for values every lane holds alike, such as grid totals.
__device__ __forceinline__ double uniform_d(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Wavefront sums of NV values, in a fixed order, valid in the last lane
// (WAVE - 1).  A workgroup whose size is a multiple of 64 (every lane of
// every wave active) sums by DPP lane moves, the NV values step by step
// together: pairs and quads (quad_perm), half rows and rows of 16
// (row_half_mirror, row_mirror), then the row sums through row_bcast:15 and
// row_bcast:31 into lane 63 (the other lanes end with partial or undefined
// values).  These are VALU moves, not LDS round trips.  Otherwise (a
// partial last wave, whose absent lanes DPP cannot skip) butterfly
// shuffles, where an absent lane reads as 0; there lane 0 holds the sum
// (wave_sum_lane).
template <int NV>
__device__ __forceinline__ void wave_sums_last(double (&v)[NV]) {
  if ((blockDim.x & (WAVE - 1)) == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += dpp_take<0xB1>(v[k]);    // quad_perm [1,0,3,2]
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += dpp_take<0x4E>(v[k]);    // quad_perm [2,3,0,1]
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += dpp_take<0x141>(v[k]);   // row_half_mirror
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += dpp_take<0x140>(v[k]);   // row_mirror
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += dpp_take<0x142>(v[k]);   // row_bcast:15 (rows 1, 3 matter)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += dpp_take<0x143>(v[k]);   // row_bcast:31 (row 3 matters)
    return;
  }
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, WAVE);
  }
}

// The lane that holds wave_sums_last's sums in every wave: the last lane
// for the DPP path; lane 0 for the shuffles, the one lane whose butterfly
// partners are all present in a partial last wave (lanes 0 .. A-1 active).
__device__ __forceinline__ int wave_sum_lane() { return (blockDim.x & (WAVE - 1)) == 0 ? WAVE - 1 : 0; }

// Wavefront sum (64 lanes); every lane gets the sum (wave_sums_last's order).
__device__ __forceinline__ double wave_sum(double v) {
  double a[1] = {v};
  wave_sums_last<1>(a);
  return read_lane(a[0], wave_sum_lane());
}

// Workgroup sum of NV values; result valid in thread 0.  `red` is LDS of
// at least NV * (blockDim/64) doubles.  Fixed summation order.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  wave_sums_last<NV>(v);
  if (lane == wave_sum_lane()) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k * nw + wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += red[k * nw + w];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Grid-wide reduction of NV values.  Each workgroup writes its partial sums
// to partials[k*nblocks + bid]; the last workgroup to arrive (device-scope
// counter) sums all partials in a fixed order and calls fin(totals) on
// thread 0, then re-arms the counter.  Deterministic run to run.
template <int NV, class Fin>
__device__ __forceinline__ void grid_sum(double (&v)[NV], double* red, double* partials,
                                         unsigned* counter, int nblocks, int bid, Fin fin) {
  block_sum<NV>(v, red);
  __shared__ int s_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * nblocks + bid] = v[k];
    __threadfence();
    unsigned old = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (old == (unsigned)(nblocks - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
      s += __hip_atomic_load(&partials[(size_t)k * nblocks + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    acc[k] = s;
  }
  block_sum<NV>(acc, red);
  if (threadIdx.x == 0) {
    fin(acc);
    __hip_atomic_store(counter, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

constexpr int MAX_WAVES = 16;   // waves of a 1024-thread workgroup

// Store this workgroup's partial sums of the first NSTORE of its NV values
// at part[k*nb + bid] (no atomics, no fences: the consumer is the next kernel
// on the stream); thread 0 is left with all NV workgroup totals in v, summed
// in block_sum's order.  One barrier; the per-wave sums go through this
// helper's own LDS, so it must run once per kernel, after the kernel's last
// use of its other LDS reductions.
template <int NV, int NSTORE = NV>
__device__ __forceinline__ void block_partials_store(double (&v)[NV], double* /*red*/, double* part, int nb,
                                                     int bid) {
  __shared__ double sp[NV * MAX_WAVES];
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  wave_sums_last<NV>(v);
  if (lane == wave_sum_lane()) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sp[k * nw + wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sp[k * nw + w];
      v[k] = s;
      if (k < NSTORE) part[(size_t)k * nb + bid] = s;
    }
  }
}

// The same partial sums as block_partials_store without its barrier: each
// wave writes its sums to LDS and counts itself in *arrive (an LDS counter
// the caller sets to 0 before an earlier barrier); the wave that counts
// last adds the per-wave sums in wave order, as block_partials_store does,
// and stores the first NSTORE of them.  The other waves go on at once (to
// their stores, which a barrier would have waited for).  Returns true in
// the last wave, whose lanes then hold all NV workgroup totals in v.
template <int NV, int NSTORE = NV>
__device__ __forceinline__ bool block_partials_store_last(double (&v)[NV], double* part, int nb, int bid,
                                                          unsigned* arrive) {
  __shared__ double sp[NV * MAX_WAVES];
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  wave_sums_last<NV>(v);
  const int sl = wave_sum_lane();
  unsigned prev = 0;
  if (lane == sl) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sp[k * nw + wid] = v[k];
    prev = __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  prev = (unsigned)__builtin_amdgcn_readlane((int)prev, sl);
  const bool last = prev == (unsigned)(nw - 1);
  if (last) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sp[k * nw + w];
      v[k] = s;
      if (k < NSTORE && lane == sl) part[(size_t)k * nb + bid] = s;
    }
  }
  return last;
}

// Every workgroup sums the nb partials of the previous kernel in the same
// fixed order (per thread, then per wave, then over the waves), so all
// workgroups hold bit-identical totals.  grid_prefetch issues the loads of
// the first chunk (b = thread index) so a kernel can start them before its
// other loads; grid_finish adds any further chunks and reduces.  One
// barrier; own LDS, used once per kernel.
template <int NV>
__device__ __forceinline__ void grid_prefetch(const double* part, int nb, double (&pre)[NV]) {
  // no branch: a thread beyond nb reads entry nb - 1 and drops it in
  // grid_finish (a load inside a branch is waited for before the branch ends)
  const int b = min((int)threadIdx.x, nb - 1);
#pragma unroll
  for (int k = 0; k < NV; ++k) pre[k] = part[(size_t)k * nb + b];
}
template <int NV, int U = 1>
__device__ __forceinline__ void grid_finish(const double* part, int nb, const double (&pre)[NV], double (&out)[NV]) {
  __shared__ double sg[NV * MAX_WAVES];
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  const int bs = blockDim.x;
  // one pass over the remaining workgroups loads all NV partials of U chunks
  // together (a chunk past nb reads entry nb - 1 and is not added); each
  // value sums over b in increasing order, from 0.0, whatever U is
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = (int)threadIdx.x < nb ? 0.0 + pre[k] : 0.0;
  for (int b0 = threadIdx.x + bs; b0 < nb; b0 += U * bs) {
    double v[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = min(b0 + u * bs, nb - 1);
#pragma unroll
      for (int k = 0; k < NV; ++k) v[u][k] = part[(size_t)k * nb + b];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = b0 + u * bs < nb;
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] = in ? acc[k] + v[u][k] : acc[k];
    }
  }
  wave_sums_last<NV>(acc);
  if (lane == wave_sum_lane()) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sg[k * nw + wid] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += sg[k * nw + w];
    out[k] = s;
  }
}
template <int NV>
__device__ __forceinline__ void grid_total(const double* part, int nb, double (&out)[NV], double* /*red*/) {
  double pre[NV];
  grid_prefetch<NV>(part, nb, pre);
  grid_finish<NV>(part, nb, pre, out);
}

// Writes D (NP x NP) and, when sW is given, w (NP) of the basis into LDS.
// Thread 0 stores the entries from scalar loads of the argument block at
// compile-time offsets: a per-lane (vector) read of the argument block is
// not served by the scalar cache and costs a memory round trip.  The
// caller's next barrier publishes them.
template <int NP>
__device__ __forceinline__ void stage_basis(const BasisK& B, double* sD, double* sW) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NP * NP; ++k) sD[k] = B.D[k];
    if (sW) {
#pragma unroll
      for (int k = 0; k < NP; ++k) sW[k] = B.w[k];
    }
  }
}

// Julia max(x, y): returns x on ties.
__device__ __forceinline__ double fmax_first(double x, double y) { return (y > x) ? y : x; }

// Device flag set when a pressure under sqrt is negative or a value is not
// finite; the host turns it into IGR_EDOMAIN.
__device__ __forceinline__ void flag_domain(int* flag) {
  __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// XCD-aware tile order: workgroups are dispatched round-robin over the 8
// XCDs, so workgroup b runs on XCD b % 8.  Giving each XCD a contiguous band
// of tiles keeps neighbouring tiles (halo reads) in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  if (nb % 8) return b;
  return (b % 8) * (nb / 8) + (b / 8);
}

}  // namespace igr
