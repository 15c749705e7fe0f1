// Batched ensembles: run_sensitivity (forward state + forward tangents) for
// many independent instances at once, mixing 1D and 2D meshes of different
// sizes, step counts, gamma and alpha.
//
//  - 1D instances: one persistent launch.  The instances' whole-run argument
//    blocks, sorted by decreasing cost, are the work list; each workgroup
//    runs its entries one after another, whole runs in one workgroup.
//  - 2D instances: every SSP-RK3 step of every instance is recorded (the
//    same host code as a single run, with the launch helpers recording
//    instead of launching), and the k-th launch of all instances becomes one
//    batched launch whose workgroups serve (instance, tile) work items.  A
//    PCG iteration of 32 meshes is then two launches instead of 64.  The
//    step's launches are captured in a hipGraph and replayed; instances
//    drop out of the recorded step when their step count is reached.
//
// Each instance computes exactly what its own igr_run_sensitivity computes:
// same kernels, same per-instance partial-sum order, bit for bit.
#include <algorithm>
#include <map>
#include <vector>

#include "igr_ctx.hpp"
#include "host_common.hpp"
#include "launch_rec.hpp"

namespace igr {

thread_local Recorder* g_rec = nullptr;   // set while this host thread records a batch step

void launch1_batch(hipStream_t st, int np, const K1Args* dargs, int n_inst, int n_wg, int block);

namespace {

int tile_x_of(const Ctx& c, bool h) {
  return c.np == 2 ? 8 : c.np == 3 ? 4 : (c.np == 5 && h) ? 2 : (h ? c.htile_x : c.tile_x);
}
int tile_y_of(const Ctx& c, bool h) {
  return c.np == 2 ? 8 : c.np == 3 ? 4 : (c.np == 5 && h) ? 4 : (h ? c.htile_y : c.tile_y);
}

struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  T* put(const std::vector<T>& h) {
    if (h.empty()) return nullptr;
    void* d = nullptr;
    if (dev_malloc(&d, h.size() * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(d);
    if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return (T*)d;
  }
};

// One batched launch sequence for a group of 2D instances whose recorded
// launch sequences are identical in kind and mode.
struct Program {
  struct L {
    int kind, mode;
    size_t off;
    int cnt;
    size_t maxn;
  };
  int np = 4, tx = 4, ty = 4, htx = 4, hty = 4;   // SIP-family and hyperbolic-family tiles
  std::vector<L> ls;
  SipArgs* sip = nullptr;
  HypArgs* hyp = nullptr;
  SensArgs* sens = nullptr;
  FinArgs* fin = nullptr;
  FillArgs* fill = nullptr;
  CopyArgs* copy = nullptr;
  int2* work = nullptr;      // (instance, workgroup) items of a SIP-family launch
  int nblocks = 0;
  int2* bwork = nullptr;     // and of a PCG x/r update (its own workgroup count)
  int bnblocks = 0;
  int2* hwork = nullptr;     // and of a hyperbolic-family launch
  int hnblocks = 0;

  void run(hipStream_t st) const {
    for (const L& l : ls) {
      switch (l.kind) {
        case R_SIP:
          if (l.mode == M_PCG_B || l.mode == M_PCG_BX)
            launch2_sip_batch(np, tx, ty, l.mode, sip + l.off, bwork, bnblocks, st);
          else launch2_sip_batch(np, tx, ty, l.mode, sip + l.off, work, nblocks, st);
          break;
        case R_HYP: launch2_hyp_batch(np, htx, hty, hyp + l.off, hwork, hnblocks, st); break;
        case R_SENS_RHS: launch2_sens_rhs_batch(np, htx, hty, sens + l.off, hwork, hnblocks, st); break;
        case R_SENS_HYP: launch2_sens_hyp_batch(np, htx, hty, sens + l.off, hwork, hnblocks, st); break;
        case R_FINAL: launch_pcg_final_batch(fin + l.off, l.cnt, st); break;
        case R_FILL: launch_fill_batch(fill + l.off, l.cnt, l.maxn, st); break;
        case R_COPY: launch_copy_batch(copy + l.off, l.cnt, l.maxn, st); break;
      }
    }
  }
};

std::vector<int> signature(const Recorder& r, const Ctx& c) {
  std::vector<int> s = {c.np, tile_x_of(c, false), tile_y_of(c, false), tile_x_of(c, true), tile_y_of(c, true)};
  s.reserve(3 + 2 * r.recs.size());
  for (const Rec& x : r.recs) { s.push_back(x.kind); s.push_back(x.mode); }
  return s;
}

// Build the programs for instances `ids` from their recordings `recs`.
int build_programs(const std::vector<Ctx*>& ctxs, const std::vector<int>& ids, const std::vector<Recorder>& recs,
                   DevBufs& dev, std::vector<Program>& progs) {
  std::map<std::vector<int>, std::vector<int>> groups;   // signature -> positions in ids
  for (size_t k = 0; k < ids.size(); ++k) groups[signature(recs[k], *ctxs[ids[k]])].push_back((int)k);
  for (auto& kv : groups) {
    const std::vector<int>& mem = kv.second;
    const Ctx& c0 = *ctxs[ids[mem[0]]];
    Program P;
    P.np = c0.np; P.tx = tile_x_of(c0, false); P.ty = tile_y_of(c0, false);
    P.htx = tile_x_of(c0, true); P.hty = tile_y_of(c0, true);
    std::vector<SipArgs> hs;
    std::vector<HypArgs> hh;
    std::vector<SensArgs> hn;
    std::vector<FinArgs> hf;
    std::vector<FillArgs> hl;
    std::vector<CopyArgs> hc;
    const Recorder& r0 = recs[mem[0]];
    for (size_t pos = 0; pos < r0.recs.size(); ++pos) {
      Program::L l{r0.recs[pos].kind, r0.recs[pos].mode, 0, (int)mem.size(), 0};
      switch (l.kind) {
        case R_SIP: l.off = hs.size(); break;
        case R_HYP: l.off = hh.size(); break;
        case R_SENS_RHS:
        case R_SENS_HYP: l.off = hn.size(); break;
        case R_FINAL: l.off = hf.size(); break;
        case R_FILL: l.off = hl.size(); break;
        case R_COPY: l.off = hc.size(); break;
      }
      for (int m : mem) {
        const Recorder& r = recs[m];
        const size_t i = r.recs[pos].idx;
        switch (l.kind) {
          case R_SIP: hs.push_back(r.sip[i]); break;
          case R_HYP: hh.push_back(r.hyp[i]); break;
          case R_SENS_RHS:
          case R_SENS_HYP: hn.push_back(r.sens[i]); break;
          case R_FINAL: hf.push_back(r.fin[i]); break;
          case R_FILL: hl.push_back(r.fill[i]); l.maxn = std::max(l.maxn, r.fill[i].n); break;
          case R_COPY: hc.push_back(r.copy[i]); l.maxn = std::max(l.maxn, r.copy[i].n); break;
        }
      }
      P.ls.push_back(l);
    }
    std::vector<int2> work, bwork, hwork;
    for (size_t g = 0; g < mem.size(); ++g) {
      const Ctx& c = *ctxs[ids[mem[g]]];
      // the workgroup counts of the instance's own launches
      const int nb = grid2_blocks(c), nbb = grid2_bblocks(c);
      for (int b = 0; b < nb; ++b) work.push_back(make_int2((int)g, b));
      for (int b = 0; b < nbb; ++b) bwork.push_back(make_int2((int)g, b));
      const int hb = ((c.Nex + P.htx - 1) / P.htx) * ((c.Ney + P.hty - 1) / P.hty);
      for (int b = 0; b < hb; ++b) hwork.push_back(make_int2((int)g, b));
    }
    P.nblocks = (int)work.size();
    P.bnblocks = (int)bwork.size();
    P.hnblocks = (int)hwork.size();
    P.work = dev.put(work);
    P.bwork = dev.put(bwork);
    P.hwork = dev.put(hwork);
    if (!P.work || !P.bwork || !P.hwork) return IGR_ENOMEM;
    if ((!hs.empty() && !(P.sip = dev.put(hs))) || (!hh.empty() && !(P.hyp = dev.put(hh))) ||
        (!hn.empty() && !(P.sens = dev.put(hn))) || (!hf.empty() && !(P.fin = dev.put(hf))) ||
        (!hl.empty() && !(P.fill = dev.put(hl))) || (!hc.empty() && !(P.copy = dev.put(hc))))
      return IGR_ENOMEM;
    progs.push_back(std::move(P));
  }
  return IGR_OK;
}

// Run `repeats` times (hipGraph replay when repeats >= 2).
int run_programs(const std::vector<Program>& progs, int repeats, hipStream_t st) {
  if (repeats <= 0 || progs.empty()) return IGR_OK;
  if (repeats == 1) {
    for (const Program& P : progs) P.run(st);
    return hipGetLastError() == hipSuccess ? IGR_OK : IGR_EHIP;
  }
  hipGraph_t g;
  hipGraphExec_t ge;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return IGR_EHIP;
  for (const Program& P : progs) P.run(st);
  if (hipStreamEndCapture(st, &g) != hipSuccess) return IGR_EHIP;
  if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return IGR_EHIP;
  }
  int rc = IGR_OK;
  for (int k = 0; k < repeats && rc == IGR_OK; ++k)
    if (hipGraphLaunch(ge, st) != hipSuccess) rc = IGR_EHIP;
  if (hipStreamSynchronize(st) != hipSuccess) rc = IGR_EHIP;
  (void)hipGraphExecDestroy(ge);
  (void)hipGraphDestroy(g);
  return rc;
}

struct Rec2 {
  Ctx* c;
  TanBufs* T;
  int nt, n_steps;
  double dts, gamma, alpha, sgam;
};

}  // namespace
}  // namespace igr

using namespace igr;

extern "C" {

int igr_batch_run_sensitivity(igr_batch_item* items, int n_items, int n_iter, int solver, int form) {
  if (!items || n_items < 1 || n_iter < 0 || solver < 0 || solver > 2) return IGR_EINVAL;
  for (int k = 0; k < n_items; ++k) {
    igr_batch_item& it = items[k];
    it.status = IGR_EINVAL;
    it.n_steps = 0;
    if (!it.ctx || !(it.dt > 0.0) || !(it.t_end > 0.0) || it.n_tan < 0 || it.n_tan > K1_MAXTAN) return IGR_EINVAL;
    if (it.n_tan > 0 && !it.s0) return IGR_EINVAL;
    if (it.ctx->dim == 1 && form != 1 && form != 2)
      return seterr(*it.ctx, IGR_EINVAL, "Unknown elliptic_rhs (use form 1 or 2)");
    for (int j = 0; j < k; ++j)
      if (items[j].ctx == it.ctx) return seterr(*it.ctx, IGR_EINVAL, "an instance appears twice in the batch");
  }
  // per-instance set-up on the instances' own streams
  std::vector<TanBufs*> tb(n_items, nullptr);
  for (int k = 0; k < n_items; ++k) {
    Ctx& c = *items[k].ctx;
    int rc = begin_call(c);
    if (!rc) rc = ensure_scal(c, n_iter);
    if (!rc) rc = tan_bufs(c, std::max(items[k].n_tan, 1), tb[k]);
    if (rc) return rc;
    const size_t nvN = c.nv * c.N;
    for (int t = 0; t < items[k].n_tan; ++t)
      HIPCHK(c, hipMemcpyAsync(tb[k]->s[t], items[k].s0 + t * nvN, nvN * sizeof(double), hipMemcpyHostToDevice,
                               c.stream));
  }
  for (int k = 0; k < n_items; ++k) HIPCHK(*items[k].ctx, hipStreamSynchronize(items[k].ctx->stream));

  hipStream_t s1 = nullptr, s2 = nullptr;
  Ctx& c0 = *items[0].ctx;
  HIPCHK(c0, hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
  if (hipStreamCreateWithFlags(&s2, hipStreamNonBlocking) != hipSuccess) {
    (void)hipStreamDestroy(s1);
    return seterr(c0, IGR_EHIP, "hipStreamCreate");
  }
  int rc = IGR_OK;
  {
    DevBufs dev;
    // ---- 1D: one persistent launch per (degree, workgroup size)
    std::map<std::pair<int, int>, std::vector<std::pair<double, K1Args>>> g1;
    for (int k = 0; k < n_items && rc == IGR_OK; ++k) {
      igr_batch_item& it = items[k];
      if (it.ctx->dim != 1) continue;
      K1Args a;
      int ns = 0;
      rc = sens1_args(*it.ctx, *tb[k], it.n_tan, it.dt, it.t_end, it.gamma, it.alpha, n_iter, solver, form,
                      it.sgam, a, ns);
      it.n_steps = ns;
      const double cost = (double)it.ctx->N * ns * (1 + it.n_tan);
      g1[{it.ctx->np, k1_block(*it.ctx)}].push_back({cost, a});
    }
    for (auto& kv : g1) {
      if (rc) break;
      auto& v = kv.second;
      std::stable_sort(v.begin(), v.end(), [](const std::pair<double, K1Args>& x, const std::pair<double, K1Args>& y) {
        return x.first > y.first;
      });
      std::vector<K1Args> h;
      for (auto& e : v) h.push_back(e.second);
      K1Args* d = dev.put(h);
      if (!d) { rc = IGR_ENOMEM; break; }
      const int n = (int)h.size();
      launch1_batch(s1, kv.first.first, d, n, std::min(n, 2048), kv.first.second);
      if (hipGetLastError() != hipSuccess) rc = IGR_EHIP;
    }

    // ---- 2D: recorded steps replayed as batched launches
    std::vector<Rec2> r2;
    std::vector<Ctx*> ctxs;
    for (int k = 0; k < n_items; ++k) {
      igr_batch_item& it = items[k];
      if (it.ctx->dim != 2) continue;
      Rec2 e;
      e.c = it.ctx; e.T = tb[k]; e.nt = it.n_tan; e.gamma = it.gamma; e.alpha = it.alpha; e.sgam = it.sgam;
      e.n_steps = sens2_steps(it.dt, it.t_end, e.dts);
      it.n_steps = e.n_steps;
      r2.push_back(e);
      ctxs.push_back(it.ctx);
    }
    if (!r2.empty() && rc == IGR_OK) {
      auto record = [&](const std::vector<int>& ids, int what, std::vector<Program>& progs) {
        std::vector<Recorder> recs(ids.size());
        for (size_t k = 0; k < ids.size(); ++k) {
          const Rec2& e = r2[ids[k]];
          g_rec = &recs[k];
          if (what == 0) sens2_begin(*e.c, *e.T, e.nt);
          else if (what == 1) sens2_step(*e.c, *e.T, e.nt, e.dts, e.gamma, e.alpha, n_iter, solver, e.sgam);
          else sens2_end(*e.c, *e.T, e.nt, e.gamma, e.alpha, n_iter, solver);
          g_rec = nullptr;
        }
        return build_programs(ctxs, ids, recs, dev, progs);
      };
      std::vector<int> all(r2.size());
      for (size_t k = 0; k < r2.size(); ++k) all[k] = (int)k;
      std::vector<Program> pb, pe;
      rc = record(all, 0, pb);
      if (!rc) rc = run_programs(pb, 1, s2);
      // phases between distinct step counts
      std::vector<int> counts;
      for (const Rec2& e : r2) counts.push_back(e.n_steps);
      std::sort(counts.begin(), counts.end());
      counts.erase(std::unique(counts.begin(), counts.end()), counts.end());
      int done = 0;
      for (int cnt : counts) {
        if (rc) break;
        std::vector<int> act;
        for (size_t k = 0; k < r2.size(); ++k)
          if (r2[k].n_steps >= cnt && r2[k].n_steps > done) act.push_back((int)k);
        std::vector<Program> ps;
        rc = record(act, 1, ps);
        if (!rc) rc = run_programs(ps, cnt - done, s2);
        done = cnt;
      }
      if (!rc) rc = record(all, 2, pe);
      if (!rc) rc = run_programs(pe, 1, s2);
    }
    if (hipStreamSynchronize(s1) != hipSuccess && !rc) rc = IGR_EHIP;
    if (hipStreamSynchronize(s2) != hipSuccess && !rc) rc = IGR_EHIP;
  }
  (void)hipStreamDestroy(s1);
  (void)hipStreamDestroy(s2);
  if (rc) {
    for (int k = 0; k < n_items; ++k) items[k].status = rc;
    return seterr(c0, rc, "batched run failed");
  }
  int first_bad = IGR_OK;
  for (int k = 0; k < n_items; ++k) {
    igr_batch_item& it = items[k];
    Ctx& c = *it.ctx;
    const size_t N = c.N, nvN = c.nv * N;
    for (int t = 0; t < it.n_tan; ++t) {
      if (it.s_out) HIPCHK(c, hipMemcpyAsync(it.s_out + t * nvN, tb[k]->s[t], nvN * sizeof(double),
                                             hipMemcpyDeviceToHost, c.stream));
      if (it.vs_out) HIPCHK(c, hipMemcpyAsync(it.vs_out + t * N, tb[k]->vs[t], N * sizeof(double),
                                              hipMemcpyDeviceToHost, c.stream));
    }
    it.status = end_call(c);
    if (it.status && !first_bad) first_bad = it.status;
  }
  return first_bad;
}

}  // extern "C"
