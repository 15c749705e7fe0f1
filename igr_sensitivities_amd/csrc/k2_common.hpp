// Shared 2D definitions: geometry, kernel argument blocks, tile/halo
// addressing and the pointwise Euler physics.
#pragma once
#include "igr_device.hpp"

namespace igr {

struct Geo2 {
  int Nex, Ney;
  double Jx, Jy, invJx, invJy, dx, dy;
  double eta;   // SIP penalty (p+1)^2, IGR_adjoints_2d/elliptic.jl:40
};

enum SipMode {
  M_APPLY = 0, M_INIT_PCG = 1, M_INIT_CHEB = 2, M_INIT_JAC = 3,
  M_PCG_A = 4, M_JACOBI = 5, M_CHEB = 6, M_PCG_B = 7,
  M_PCG_BX = 8   // PCG_B of the last iteration, which also updates x (the others leave it to the next PCG_A)
};

struct SipArgs {
  Geo2 g;
  BasisK B;
  double alpha;
  const double* x = nullptr;     // vector (APPLY/JACOBI/CHEB) or warm start (INIT)
  const double* rho = nullptr;   // APPLY, INIT
  const double* mx = nullptr;    // INIT
  const double* my = nullptr;    // INIT
  const double* invR = nullptr;
  const double* r = nullptr;
  const double* invM = nullptr;
  const double* pold = nullptr;
  const double* b = nullptr;
  const double* d = nullptr;
  const double* b_in = nullptr;  // INIT: use this right-hand side instead of the IGR source
  double* sc = nullptr;          // per-solve scalars (rz | pAp | a | beta)
  int first = 0, it = 0, n_iter = 0;
  int ngroups = 0;               // workgroups of the launch (set by the launcher)
  int np_in = 0;                 // partial sums the previous kernel wrote (set by the launcher)
  double c1 = 0.0, c2 = 0.0, theta = 1.0, omega = 0.0;
  double* y = nullptr;
  double* pnew = nullptr;
  double* xout = nullptr;
  double* dout = nullptr;
  double* rout = nullptr;
  double* invR_out = nullptr;
  double* invM_out = nullptr;
  double* b_out = nullptr;
  double* partials = nullptr;
  unsigned* counter = nullptr;
  const double* part_in = nullptr;   // partial sums written by the previous kernel
  double* part_out = nullptr;        // this kernel's partial sums
  int* flag = nullptr;
};

struct HypArgs {
  Geo2 g;
  BasisK B;
  double gamma;
  const double* q;    // stage input state (4 fields)
  const double* S;    // Sigma
  const double* qn;   // q^n (stages 2 and 3)
  double* out;        // dq (mode 0) or the stage output state
  int mode;           // 0: dq ; 1, 2, 3: SSP-RK3 stage
  double cdt;         // dt, (1/4)dt or (2/3)dt, as the reference forms it
  int* flag;
};

// Forward tangent (sensitivity) kernels.
struct SensArgs {
  Geo2 g;
  BasisK B;
  double gamma, alpha;
  const double* q;    // primal stage state (4 fields)
  const double* S;    // primal Sigma
  const double* s;    // tangent stage state (4 fields: zx, zy, sigma, eta)
  const double* vs;   // tangent Sigma (varsigma)
  const double* sn;   // tangent state at t^n (stages 2, 3)
  double* out;        // b (k2_sens_rhs) or the tangent stage output / ds
  int mode;           // 0: ds ; 1, 2, 3: SSP-RK3 stage
  double cdt;
  double sgam;        // gamma source sgam * rho e of a d/dgamma tangent (0: the reference's 2D tangent)
  int* flag;
};

// Batched launches over many meshes: workgroup b of the launch serves
// work[b] = (instance, workgroup of that instance), and the instance's
// arguments are entry `instance` of a device table.
template <int TX, int TY>
__host__ __device__ inline int tiles_of(const Geo2& g) {
  return ((g.Nex + TX - 1) / TX) * ((g.Ney + TY - 1) / TY);
}
struct FinArgs { const double* part; int nb; double* sc; int nit; };   // end of a PCG solve
struct FillArgs { double* x; size_t n; double v; };
struct CopyArgs { double* dst; const double* src; size_t n; };

// Tile of TX x TY elements plus a one-element halo ring (no corners).
// LDS slots: tile s = ly*TX + lx, then left[TY], right[TY], bottom[TX], top[TX].
template <int NP, int TX, int TY>
struct TileInfo {
  static constexpr int T = TX * TY, NP2 = NP * NP;
  int Nex, Ney, ex0, ey0, nvx, nvy;
  int exL, exR, eyB, eyT;   // columns / rows of the four halo strips (periodic)
  __device__ TileInfo(const Geo2& g) : TileInfo(g, blockIdx.x, gridDim.x) {}
  // workgroup blk of the nblk workgroups that cover this mesh
  __device__ TileInfo(const Geo2& g, int blk, int nblk) {
    const int ntx = (g.Nex + TX - 1) / TX;
    const int tile = xcd_tile(blk, nblk);
    const int tx = tile % ntx, ty = tile / ntx;
    Nex = g.Nex; Ney = g.Ney;
    ex0 = tx * TX; ey0 = ty * TY;
    nvx = min(TX, Nex - ex0); nvy = min(TY, Ney - ey0);
    // periodic wrap by comparison (ex0 < Nex, ex0 + nvx <= Nex): a runtime
    // integer modulo costs tens of instructions ahead of every halo load
    exL = ex0 > 0 ? ex0 - 1 : Nex - 1;
    exR = ex0 + nvx < Nex ? ex0 + nvx : ex0 + nvx - Nex;
    eyB = ey0 > 0 ? ey0 - 1 : Ney - 1;
    eyT = ey0 + nvy < Ney ? ey0 + nvy : ey0 + nvy - Ney;
  }
  __device__ size_t gidx(int s, int n) const {
    const int lx = s % TX, ly = s / TX;
    return (size_t)n + (size_t)NP2 * ((size_t)(ex0 + lx) + (size_t)Nex * (ey0 + ly));
  }
  __device__ int nb_left(int s) const { const int lx = s % TX, ly = s / TX; return lx > 0 ? s - 1 : T + ly; }
  __device__ int nb_right(int s) const { const int lx = s % TX, ly = s / TX; return lx < nvx - 1 ? s + 1 : T + TY + ly; }
  __device__ int nb_bottom(int s) const { const int lx = s % TX, ly = s / TX; return ly > 0 ? s - TX : T + 2 * TY + lx; }
  __device__ int nb_top(int s) const { const int lx = s % TX, ly = s / TX; return ly < nvy - 1 ? s + TX : T + 2 * TY + TX + lx; }
  // Element of LDS slot sl (tile slots 0..T-1, then the left, right, bottom
  // and top halo strips); false when the slot lies outside this tile's rows
  // or columns or past the last strip.  Selects only, no branches: a
  // per-lane branch tree here ran ahead of every tile load.
  __device__ bool slot_elem(int sl, int& gex, int& gey) const {
    const bool in = sl < T;
    const int lx = sl % TX, ly = sl / TX;
    const int h = sl - T;
    const bool xs = h < 2 * TY;                        // left or right strip
    const int k = xs ? (h < TY ? h : h - TY) : (h < 2 * TY + TX ? h - 2 * TY : h - 2 * TY - TX);
    gex = in ? ex0 + lx : xs ? (h < TY ? exL : exR) : ex0 + k;
    gey = in ? ey0 + ly : xs ? ey0 + k : (h < 2 * TY + TX ? eyB : eyT);
    return in ? (lx < nvx && ly < nvy) : (h < 2 * (TX + TY) && k < (xs ? nvy : nvx));
  }
  // Node n of the slot staged by the thread of tile element s in load round l
  // of a T * NP2-thread workgroup (slot s + l T): round 0 is the tile
  // element s itself, round l >= 1 the halo slot h = s + (l - 1) T.  The
  // address is 0 when the slot is empty.  With l known at compile time (an
  // unrolled loop) this is a handful of selects.
  __device__ bool round_global(int l, int s, int n, size_t& gi) const {
    int gex, gey;
    bool ok;
    if (l == 0) {
      const int lx = s % TX, ly = s / TX;
      ok = lx < nvx && ly < nvy;
      gex = ex0 + lx; gey = ey0 + ly;
    } else {
      const int h = s + (l - 1) * T;
      const bool xs = h < 2 * TY;                               // left or right strip
      const bool lo = xs ? h < TY : h < 2 * TY + TX;            // left / bottom
      const int k = h - (xs ? (lo ? 0 : TY) : (lo ? 2 * TY : 2 * TY + TX));
      gex = xs ? (lo ? exL : exR) : ex0 + k;
      gey = xs ? ey0 + k : (lo ? eyB : eyT);
      ok = h < 2 * (TX + TY) && k < (xs ? nvy : nvx);
    }
    gi = ok ? (size_t)n + (size_t)NP2 * ((size_t)gex + (size_t)Nex * (size_t)gey) : 0;
    return ok;
  }
  __device__ bool slot_global(int sl, int nn, size_t& gi) const {
    int gex, gey;
    if (!slot_elem(sl, gex, gey)) return false;
    gi = (size_t)nn + (size_t)NP2 * ((size_t)gex + (size_t)Nex * gey);
    return true;
  }
  // k-th node of the face of halo element h that touches the tile
  __device__ bool halo_face_global(int h, int k, size_t& gi) const {
    int gex, gey, nn;
    if (!slot_elem(T + h, gex, gey)) return false;
    if (h < TY) nn = (NP - 1) + NP * k;              // left neighbour: its right face
    else if (h < 2 * TY) nn = NP * k;                // right neighbour: its left face
    else if (h < 2 * TY + TX) nn = k + NP * (NP - 1); // bottom neighbour: its top face
    else nn = k;                                      // top neighbour: its bottom face
    gi = (size_t)nn + (size_t)NP2 * ((size_t)gex + (size_t)Nex * gey);
    return true;
  }
};

// ---- pointwise Euler physics, reference evaluation order, no contraction
#pragma clang fp contract(off)
// polytropic_flux_x / _y, IGR_adjoints_2d/euler.jl:31-61.  q = (mx,my,rho,E,Sigma)
__device__ __forceinline__ void flux_x2(double gam, const double q[5], double f[4]) {
  const double P = (gam - 1.0) * (q[3] - (q[0] * q[0] + q[1] * q[1]) / (2.0 * q[2]));
  const double pb = P + q[4];
  const double ux = q[0] / q[2];
  f[0] = q[0] * ux + pb; f[1] = q[1] * ux; f[2] = q[0]; f[3] = (q[3] + pb) * ux;
}
__device__ __forceinline__ void flux_y2(double gam, const double q[5], double f[4]) {
  const double P = (gam - 1.0) * (q[3] - (q[0] * q[0] + q[1] * q[1]) / (2.0 * q[2]));
  const double pb = P + q[4];
  const double uy = q[1] / q[2];
  f[0] = q[0] * uy; f[1] = q[1] * uy + pb; f[2] = q[1]; f[3] = (q[3] + pb) * uy;
}
// sound_speed, euler.jl:68-71; a negative argument is Julia's DomainError
__device__ __forceinline__ double sound2(double gam, const double q[5], int* flag) {
  const double P = (gam - 1.0) * (q[3] - (q[0] * q[0] + q[1] * q[1]) / (2.0 * q[2]));
  const double arg = gam * P / q[2];
  if (!(arg >= 0.0)) flag_domain(flag);
  return sqrt(arg);
}
// f* = (fL + fR)/2 + lam (qL - qR)/2, euler.jl:105-108
__device__ __forceinline__ void llf_combine(const double fL[4], const double fR[4], const double qL[5],
                                            const double qR[5], double lam, double f[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) f[v] = (fL[v] + fR[v]) / 2.0 + lam * (qL[v] - qR[v]) / 2.0;
}
#pragma clang fp contract(fast)

}  // namespace igr
