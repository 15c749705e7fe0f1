// Launch recorder: while g_rec is set, the 2D launch helpers append their
// kernel and arguments to it instead of launching.  The batched driver
// records one SSP-RK3 step of every instance this way and replays the
// position-k launches of all instances as one batched launch.
#pragma once
#include <vector>
#include "k2_common.hpp"

namespace igr {

enum RecKind { R_SIP = 0, R_HYP = 1, R_SENS_RHS = 2, R_SENS_HYP = 3, R_FINAL = 4, R_FILL = 5, R_COPY = 6 };

struct Rec {
  int kind, mode;
  size_t idx;   // index into the vector of this kind
};

struct Recorder {
  std::vector<Rec> recs;
  std::vector<SipArgs> sip;
  std::vector<HypArgs> hyp;
  std::vector<SensArgs> sens;
  std::vector<FinArgs> fin;
  std::vector<FillArgs> fill;
  std::vector<CopyArgs> copy;
  void clear() {
    recs.clear(); sip.clear(); hyp.clear(); sens.clear(); fin.clear(); fill.clear(); copy.clear();
  }
};

extern thread_local Recorder* g_rec;

// Batched launchers (k2_forward.hip, k2_tangent.hip): `work` has one
// (instance, workgroup) entry per workgroup of the launch.
void launch2_sip_batch(int np, int tx, int ty, int mode, const SipArgs* tab, const int2* work, int nblocks,
                       hipStream_t st);
void launch2_hyp_batch(int np, int tx, int ty, const HypArgs* tab, const int2* work, int nblocks, hipStream_t st);
void launch2_sens_rhs_batch(int np, int tx, int ty, const SensArgs* tab, const int2* work, int nblocks,
                            hipStream_t st);
void launch2_sens_hyp_batch(int np, int tx, int ty, const SensArgs* tab, const int2* work, int nblocks,
                            hipStream_t st);
void launch_pcg_final_batch(const FinArgs* tab, int n, hipStream_t st);
void launch_fill_batch(const FillArgs* tab, int n, size_t maxn, hipStream_t st);
void launch_copy_batch(const CopyArgs* tab, int n, size_t maxn, hipStream_t st);

}  // namespace igr
