// Argument blocks of the continuous-adjoint kernels (2D: k2_adjpde.hip,
// 1D: k1_adjpde.hip) and of the snapshot-storing forward run.
#pragma once
#include "k1_common.hpp"
#include "k2_common.hpp"

namespace igr {

struct AdjStageArgs {
  Geo2 g;
  BasisK B;
  double gamma;
  const double* q;    // primal snapshot state (4 fields)
  const double* S;    // primal snapshot Sigma
  const double* A;    // adjoint stage input (4 fields)
  const double* an;   // adjoint at the start of the step (stages 2, 3)
  const double* Pi;   // adjoint elliptic variable; null: no nonlocal source (alpha <= 0)
  double* out;        // da (mode 0) or the stage output (4 fields)
  int mode;           // 0: da ; 1, 2, 3: SSP-RK3 stage
  double cdt;         // dt, (1/4)dt or (2/3)dt
  int* flag;
};

struct AdjBArgs {
  Geo2 g;
  BasisK B;
  double alpha;
  const double* q;    // primal snapshot state
  const double* A;    // adjoint stage input
  double* b;          // right-hand side of the Pi solve
};

struct AdjTermArgs {
  Geo2 g;
  BasisK B;
  int np, obj;
  double prm[4];
  double Lx, Ly, gamma;
  const double* q;     // state at t = T
  const double* sums;  // [1] kinetic energy, [2] mass (KINETIC_ENERGY)
  double* a;           // adjoint at t = T (4 fields)
};

// 1D continuous adjoint of one instance (one workgroup, k1_adjpde.hip).
struct K1PdeArgs {
  K1Args f;                  // geometry, basis, alpha/gamma, solver and its workspace
  const double* snaps;       // records (mu, rho, E, Sigma), 4N doubles each
  int n_steps;               // backward steps (records n_steps .. 1)
  int n_last;                // record of the terminal condition (the last one stored)
  double h;                  // adjoint time step min(dt, T)
  int obj;                   // objective of the terminal condition
  double* A;                 // adjoint (3N): terminal condition in, adjoint at t = 0 out
  double* A1;                // stage buffers (3N each)
  double* A2;
  double* Pi;                // adjoint elliptic variable (N), zeroed first
  double* bPi;               // its right-hand side (N)
};

}  // namespace igr
