// The solver context behind the C ABI (struct lps_ctx), shared by the files that implement the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lps_amd.h"
#include "lps_internal.h"

namespace lps {
struct Dev2 {
  double *x = nullptr, *y = nullptr;
};
struct ClosureAttach;  // a learned ABC closure attached to a context (lps_closure.hip)
}  // namespace lps


struct lps_ctx {
  using Dev2 = lps::Dev2;
  using InstParams = lps::InstParams;
  using DiagPartial = lps::DiagPartial;
  using HorizonBuild = lps::HorizonBuild;
  using TileArgs = lps::TileArgs;
  using StepArgs = lps::StepArgs;
  using StepState = lps::StepState;
  using RingU = lps::RingU;
  lps_options opt{};
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // the ABC rows' kernels beside the tile kernels
  hipStream_t stream3 = nullptr;  // the partial tiles' kernels beside the full tiles' kernels
  hipStream_t stream4 = nullptr;  // the learned closure's history and network beside the tile
                                  // kernels (LPS_CLOSURE_FIRST=0; by default they run first on stream)
  hipEvent_t ev_clo = nullptr;    // the closure's network done (its corrections stored)
  hipEvent_t ev_fork[2] = {nullptr, nullptr}, ev_join[2] = {nullptr, nullptr};
  hipEvent_t ev_part[3] = {nullptr, nullptr, nullptr};  // partial theta done, theta done, partial force done
  std::string err, path;
  int32_t n = 0, n_slots = 0, n_abc = 0, M = 0;
  double E = 0, nu = 0, rho = 0, dt = 0, delta = 0, dx = 0, dy = 0;
  double lam = 0, mu = 0, kappa = 0, m_ref = 0, exclude_r = -1, E0 = -1;
  bool fast = false;
  bool stencil68_ok = false;  // the unrolled 68-bond kernels apply to this problem
  bool abc_in_part = false;   // tile path: every ABC node lies in a partial tile
  // the ABC nodes' prediction folded into the partial tiles' force pass (which then waits for
  // G_u / G_v); else a kernel of its own after G_u / G_v, the partial tiles beside them
  bool fold_pred = true;
  // with the ABC rows on their own stream: the closure's network alone at the step's start on the
  // main stream (else on the fourth stream beside the tile kernels); LPS_CLOSURE_FIRST=0/1
  bool clo_first = true;
  int64_t n_full_nodes = 0;   // nodes in the full tiles
  // the full tiles split for concurrent == 4: deep (the 3 x 3 neighbourhood full, interior nodes
  // only in the 24 x 24 window: dilatation and force in one launch) and shallow (the others)
  int32_t *deep_list = nullptr, *shallow_list = nullptr;
  int32_t n_deep = 0, n_shallow = 0;
  // instances: one for lps_create, several for lps_create_batch
  std::vector<InstParams> inst_h;
  InstParams *inst_d = nullptr;
  bool batch = false;
  int32_t n_inst = 1;
  std::vector<int32_t> tile_off, abc_n, node_n;  // per instance (tile_off: n_inst + 1 entries)
  std::vector<int64_t> slot_off, node_off;       // n_inst + 1 entries
  std::vector<double> E0_inst;
  int32_t *tile_off_d = nullptr;
  DiagPartial *dout_inst = nullptr, *hout_inst = nullptr;
  Dev2 P0;                        // batch: prediction of step 1 and theta history of the setup
  double *TH0 = nullptr;
  int64_t step = 0;
  std::vector<int32_t> perm_h;  // slot -> original index, -1 at holes
  // FNV-1a hash of the ABC inputs (ABC node list, clouds, G_u, G_v, G_theta as given), one per
  // instance: a checkpoint is only restored into a context with the same operators
  std::vector<uint64_t> abc_hash;
  HorizonBuild hb;
  std::vector<void *> allocs;
  // device arrays (per slot unless noted)
  double *x = nullptr, *y = nullptr, *m = nullptr, *c2m = nullptr;
  int8_t *kind = nullptr, *is_bc = nullptr;
  int32_t *perm_d = nullptr, *abc_slot = nullptr;
  // ABC rows, ragged (StepArgs): row r's members abc_off[r] .. abc_off[r + 1], its operator
  // block at member abc_gofs[r] of gu / gv / gth
  int32_t *abc_off = nullptr, *cloud = nullptr, *abc_gofs = nullptr;
  int8_t *abc_xf = nullptr;
  uint8_t *abc_pmap = nullptr;
  int64_t abc_members = 0, abc_block_members = 0;
  bool abc_shared_blocks = false;  // lattice images share their canonical row's block
  double *gu = nullptr, *gv = nullptr, *gth = nullptr;
  // the clouds' members packed per member for the fast path's row kernels (AbcPack): their slots,
  // each cloud entry's member index (-1 pad) and the records, rewritten at each step's start
  int32_t *pk_slot = nullptr, *pk_idx = nullptr;
  int32_t n_pk = 0, pk_stride = 0;
  double4 *pk_lvl[2] = {nullptr, nullptr};  // U, V of one level per member, by step parity
  int64_t pk_last_step = -1;  // the step whose level the records hold last (-1: none valid)
  double2 *pk_th = nullptr;
  TileArgs tile{};
  // Prediction ring: step s reads P[s % 6] (U^{s+1} before the ABC overwrite) and writes
  // P[(s + 1) % 6]. At the interior nodes (kind 1) U^{s+1} equals that prediction and the fast
  // path does not store it again: U^{s-j} of those nodes is ring_u(s, j) = P[(s + 5 - j) % 6]
  // (j <= 3 while step s runs). U[] holds the ABC nodes' values, and the interior nodes' only
  // after sync_u copies them in; u_synced: U^{s-j} holds every node for j < u_synced.
  Dev2 U[3], V[3], A[2], P[6], IC, ICD;
  int32_t u_synced = 3;
  const Dev2 &ring_u(int64_t s, int j) const { return P[(s + 5 - j) % 6]; }
  double *TH[3] = {nullptr, nullptr, nullptr};
  DiagPartial *partials = nullptr, *dout = nullptr, *hout = nullptr;
  int32_t n_partials = 0;
  // graph of 30 steps from a step that is a multiple of 30
  hipGraphExec_t graph30 = nullptr;
  // the same with the timing events inside (external event-record nodes), one per event set:
  // timing mode alternates them, reading one set's times while the other replays, as long as
  // graph_timing holds (cleared when the runtime cannot time such events)
  hipGraphExec_t graph30t[2] = {nullptr, nullptr};
  bool graph_timing = true, capturing = false;
  // timing: 2 sets of 30 steps x 4 events. ev_set receives the next steps (ev_used so far); the
  // other set may hold a replay's events not yet read (ev_prev_used steps)
  bool timing = false;
  bool timing_force_only = false;  // lps_set_timing(ctx, 2 or 3): only two events per step
  bool timing_closure = false;     // lps_set_timing(ctx, 3): those two around the closure's network
  hipEvent_t *clo_ev = nullptr;    // the current step's events while it is launched (mode 3)
  std::vector<hipEvent_t> ev;
  int32_t ev_set = 0, ev_used = 0, ev_prev_used = 0;
  double t_pass[4] = {0, 0, 0, 0};  // dilatation, force (nodes), force (ABC rows), whole step
  int64_t t_count = 0;
  // learned ABC closure applied after the G_u / G_v rows (null: least-squares ABC only)
  lps::ClosureAttach *clo = nullptr;

  template <class T>
  hipError_t alloc(T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) allocs.push_back(*p);
    return e;
  }
  template <class T>
  hipError_t upload(T **p, const std::vector<T> &h) {
    hipError_t e = alloc(p, h.size());
    if (e != hipSuccess) return e;
    if (h.empty()) return hipSuccess;
    return hipMemcpyAsync(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, stream);
  }
  int32_t fail(int32_t code, const std::string &msg) {
    err = msg;
    return code;
  }
  int32_t hip_fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return LPS_ERR_HIP;
  }
  StepArgs args() const;
  StepState state(int64_t s) const;
  RingU ring(int64_t s) const {
    const Dev2 &a = ring_u(s, 0), &b = ring_u(s, 1);
    return RingU{a.x, a.y, b.x, b.y};
  }
  lps::AbcPack pack(int64_t s) const {
    return lps::AbcPack{pk_idx, pk_lvl[s & 1], pk_lvl[(s + 1) & 1], pk_th, pk_stride};
  }
};

namespace lps {
// Learned ABC closure hooks (lps_closure.hip), active when ctx->clo is set. step: records U^n, V^n
// into the history and runs the network. Without defer it runs after the ABC rows' G_u / G_v on
// their stream (at the end of the step when the context has no ABC rows) and adds each correction
// to U^{n+1}, V^{n+1} at its ABC node. With defer it only reads step n's state, so it may run from
// the step's start; it stores the corrections, and apply adds them after G_u / G_v.
hipError_t closure_step_hook(lps_ctx *c, const StepState &st, hipStream_t s, bool defer);
hipError_t closure_apply_hook(lps_ctx *c, const StepState &st, hipStream_t s);
hipError_t closure_reset_hook(lps_ctx *c);  // clears the history and the fallbacks (lps_reset)
// At each output record (count instances, de_rel set): an instance whose energy rose by more than
// the guard tolerance times its E0 since the previous record falls back to the least-squares ABC.
int32_t closure_record_hook(lps_ctx *c, const lps_diag *d, int32_t count);
void closure_free(lps_ctx *c);
// Drops the captured 30-step graphs (they hold the kernels they were captured with).
void ctx_drop_graphs(lps_ctx *c);
// Fast path: copies the interior nodes' U^{s-j}, j < depth (<= 3), from the prediction ring into
// U (s = ctx->step), on ctx->stream, where not already done since the last step.
int32_t ctx_sync_u(lps_ctx *c, int32_t depth);
}  // namespace lps

#define CTX_TRY(ctx, call)                                   \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #call); \
  } while (0)

