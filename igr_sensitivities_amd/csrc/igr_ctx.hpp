// Host-side context: one mesh, its device-resident state and workspace,
// one HIP stream.  A context is not thread-safe (the reference is serial).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <cstdio>
#include <map>
#include <utility>

#include "../../include/igr.h"
#include "igr_device.hpp"

namespace igr {

// hipMalloc that, when the device is out of memory, first frees the idle
// cached kept-iterate block of the current device and tries once more
hipError_t dev_malloc(void** p, size_t bytes);

struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
};

struct Ctx {
  int dim = 2, p = 3, np = 4, Nex = 1, Ney = 1;
  double Lx = 1.0, Ly = 1.0;
  size_t N = 0;         // nodes per field
  int nv = 4;           // conserved variables
  hipStream_t stream = nullptr;
  BasisK basis{};
  std::string err;

  // state (device): q[nv*N], Sigma[N]
  double* q = nullptr;
  double* S = nullptr;
  // RK3 stage buffers and RHS
  double* q1 = nullptr;
  double* q2 = nullptr;
  double* dq = nullptr;
  // elliptic workspace
  double* invR = nullptr;
  double* invM = nullptr;
  double* b = nullptr;
  double* r = nullptr;
  double* r2 = nullptr;
  double* pA = nullptr;
  double* pB = nullptr;
  double* Ap = nullptr;
  double* xA = nullptr;
  double* dA = nullptr;
  double* dB = nullptr;
  double* tmp = nullptr;
  double* scal = nullptr;     // per-solve scalars
  size_t scal_n = 0;
  double* partials = nullptr; // grid reduction partials
  size_t partials_n = 0;
  size_t pstride = 0;         // two partial-sum regions: partials, partials + pstride
  unsigned* counter = nullptr;
  int* flag = nullptr;        // domain flag
  double* dscalar = nullptr;  // small device scalar scratch
  double* ic_tab = nullptr;   // enthalpy-defect table of the Oseen-pair initial condition
  std::vector<void*> owned;   // every hipMalloc, freed by igr_destroy

  // 1D whole-instance kernel workspace
  double* k1F = nullptr;      // 3N fluxes
  double* k1face = nullptr;   // 6*Ne face values
  double* k1args = nullptr;   // device copy of the 1D kernel's argument block (K1Args)

  // 2D elements per workgroup (p = 3): tile_* for the SIP/solver kernels (the
  // kernels with grid reductions), rtile_* for their reverse, htile_* for the
  // hyperbolic, tangent and objective kernels; chosen by a size table or by
  // the autotuner (igr_set_option IGR_OPT_AUTOTUNE, forward families only)
  int tile_x = 4, tile_y = 4;
  int rtile_x = 4, rtile_y = 4;
  int htile_x = 4, htile_y = 4;
  int use_graphs = 1;
  int k1_wave = 1;            // 1D: one-wave kernel for small plain runs (k1_wave.hip)
  // workgroups of the PCG x/r update (k2_pcg_b): 0 = one per tile; n > 0 =
  // at most n, each running every n-th tile
  int bgroups = 0;
  // workgroups of the reverse PCG iteration (k2_sip_rev PCG_ITER) and its
  // start kernel: 0 = one per reverse tile; n > 0 = at most n, each running
  // every n-th tile and storing one partial sum per dot product
  int rgroups = 0;
  // 2D gradient: the most whole steps whose solver iterates are kept in HBM
  // (-1: as many as 85% of the free memory holds; 0: none)
  int keep_steps = -1;

  // run_forward_store snapshots: records of (state, Sigma), (nv + 1) N
  // doubles each, for the continuous adjoint (adjpde.hip)
  double* snaps = nullptr;
  size_t snaps_cap = 0;       // records allocated
  int snaps_n = -1;           // steps stored (records 0 .. snaps_n); -1: none
  double* adjw = nullptr;     // continuous-adjoint workspace (adjpde.hip)

  // captured SSP-RK3 steps, keyed by their scalar parameters
  struct GraphKey {
    double dt, gamma, alpha; int n_iter, solver;
    int kind = 0;   // 0: SSP-RK3 step; 1: step + Sigma solve (store run); 2: adjoint step
    bool operator<(const GraphKey& o) const {
      if (kind != o.kind) return kind < o.kind;
      if (dt != o.dt) return dt < o.dt;
      if (gamma != o.gamma) return gamma < o.gamma;
      if (alpha != o.alpha) return alpha < o.alpha;
      if (n_iter != o.n_iter) return n_iter < o.n_iter;
      return solver < o.solver;
    }
  };
  std::map<GraphKey, std::pair<hipGraph_t, hipGraphExec_t>> graphs;

  double* alloc(size_t n_doubles);
};

}  // namespace igr

struct igr_ctx : igr::Ctx {};
