// The Sigma solve for callers that cannot allocate device memory: the
// kernels of igr_solve_elliptic run on caller-owned device arrays, and their
// work vectors, partial sums, scalar record and domain flag live in one
// file-scope device array instead of a context's allocations.
//
// The array is a host-side stand-in in host code: its device address is
// resolved once per device with hipGetSymbolAddress (each GPU that runs this
// code holds its own copy until the process ends) and only that address is
// passed to kernels.  Every call on a device shares that device's array, so
// calls that overlap on one GPU (two streams, two host threads) race.
#include <algorithm>
#include <mutex>
#include "host_common.hpp"

namespace igr {

// Largest mesh the scratch array serves: 262,144 nodes per field, e.g.
// 128 x 128 elements at p = 3; up to 1024 solver iterations.
constexpr size_t NA_MAX_NODES = 262144;
constexpr int NA_MAX_ITER = 1024;
constexpr size_t NA_FIELDS = 8;            // 1/rho, 1/M, b, r, r2, pA/xA, pB/dA, Ap/dB
constexpr size_t NA_PART = 2 * 4096;       // two partial-sum regions
constexpr size_t NA_SC = 4 * (size_t)NA_MAX_ITER + 64;
constexpr size_t NA_DOUBLES = NA_FIELDS * NA_MAX_NODES + NA_PART + NA_SC + 64;   // 16.1 MiB

static __device__ __align__(256) double scratch[NA_DOUBLES];

static __global__ void k_na_zero_flag(int* p) { *p = 0; }

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
double* g_na_ptr[NA_MAX_DEVICES] = {};

// Device address of this device's scratch array (resolved on first use).
int scratch_ptr(double*& out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return IGR_EHIP;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    double* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return IGR_EHIP;
    }
    g_na_ptr[dev] = p;
  }
  out = g_na_ptr[dev];
  return IGR_OK;
}
}  // namespace

}  // namespace igr

using namespace igr;

extern "C" int igr_solve_elliptic_noalloc(int p, int nex, int ney, double lx, double ly, double alpha, int n_iter,
                                          int solver, const double* mx, const double* my, const double* rho,
                                          double* sigma, int* domain_flag, void* stream) {
  if (p < 1 || p > 4 || nex < 1 || ney < 1 || !(lx > 0.0) || !(ly > 0.0) || n_iter < 0 ||
      n_iter > NA_MAX_ITER || solver < 0 || solver > 2 || !mx || !my || !rho || !sigma)
    return IGR_EINVAL;
  const size_t np = (size_t)p + 1;
  const size_t N = np * np * (size_t)nex * (size_t)ney;
  if (N > NA_MAX_NODES) return IGR_EINVAL;
  (void)hipGetLastError();   // a launch error reported below is this call's own
  double* base = nullptr;
  int rc = scratch_ptr(base);
  if (rc) return rc;

  // A host-side description of the mesh whose buffers all point into the
  // scratch array (no device allocation; the Ctx members stay empty).
  Ctx c;
  c.dim = 2; c.p = p; c.np = p + 1; c.Nex = nex; c.Ney = ney; c.Lx = lx; c.Ly = ly; c.nv = 4; c.N = N;
  c.stream = (hipStream_t)stream;
  double xi[MAXNP], w[MAXNP], Dm[MAXNP * MAXNP];
  make_basis(p, xi, w, Dm);
  for (int i = 0; i < c.np; ++i) { c.basis.xi[i] = xi[i]; c.basis.w[i] = w[i]; }
  for (int k = 0; k < c.np * c.np; ++k) c.basis.D[k] = Dm[k];
  default_tiles(c);
  const size_t nb = (size_t)grid2_blocks(c);
  if (2 * (nb + 64) > NA_PART) return IGR_EINVAL;

  double* f = base;
  double** fields[NA_FIELDS] = {&c.invR, &c.invM, &c.b, &c.r, &c.r2, &c.pA, &c.pB, &c.Ap};
  for (size_t k = 0; k < NA_FIELDS; ++k) *fields[k] = f + k * NA_MAX_NODES;
  // Jacobi and Chebyshev use xA, dA, dB where PCG uses pA, pB, Ap
  c.xA = c.pA; c.dA = c.pB; c.dB = c.Ap;
  c.partials = base + NA_FIELDS * NA_MAX_NODES;
  c.pstride = NA_PART / 2;
  c.partials_n = NA_PART;
  c.scal = c.partials + NA_PART;
  c.scal_n = NA_SC;
  int* own_flag = reinterpret_cast<int*>(c.scal + NA_SC);
  c.flag = domain_flag ? domain_flag : own_flag;
  c.counter = reinterpret_cast<unsigned*>(own_flag + 16);
  c.S = sigma;

  // the flag starts at zero (a kernel rather than a memset, so that the call
  // may be captured in a graph)
  hipLaunchKernelGGL(k_na_zero_flag, dim3(1), dim3(1), 0, c.stream, c.flag);
  solve2_fields(c, mx, my, rho, alpha, n_iter, solver, c.scal, sigma, nullptr);
  return hipGetLastError() == hipSuccess ? IGR_OK : IGR_EHIP;
}
