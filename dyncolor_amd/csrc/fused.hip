// fused.hip -- device-driven recolor of one staged batch in ONE persistent
// cooperative launch.  The phases are the same device bodies the plain driver
// launches one by one (dc_phases.h); here they are separated by grid-wide
// barriers, and the round loop's exit test is read from device counters right
// after a barrier, so every workgroup takes the same path and no host round
// trip happens per round.
//
// Counter discipline (all set to zero by the host before the launch):
//   T[k].cnt / chg[k], k < 4  : one set per pre-loop round, never reused
//   F[0] / F[1] counters       : loop frontiers; the frontier that was L in
//                                round r is C in round r+1, so block 0 zeroes
//                                it during round r's recolor phase (nobody
//                                reads L's counters after the find barrier)
//   chg[4 + (r & 1)]           : changes of loop round r; read by every block
//                                right after round r's recolor barrier, zeroed
//                                by block 0 during round r+1's recolor phase
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "dc_phases.h"

namespace cg = cooperative_groups;

namespace dc {

struct FusedResult {
  int32_t status;
  int32_t cur;
  int32_t rounds;
  int32_t phases;
  uint32_t epochs;
  uint32_t pad[3];
  unsigned long long evaluated;
  unsigned long long changed;
  unsigned long long frontier_sum;
};

struct FusedArgs {
  GraphView G;
  int32_t* color0;
  int32_t* color1;
  int32_t cur0;
  int32_t* prev;
  uint32_t* tmark;
  uint32_t* mmark;
  uint32_t epoch0;
  Frontier F0, F1, Mk;
  Frontier T0, T1, T2, T3;
  Frontier X;
  const int2* del;
  const int2* ins;
  const int2* xdel;
  const int2* xins;
  Counters* ctr;
  unsigned long long* partial;
  uint32_t* cbits[2];          // loop-round claim bitmaps (clear at launch)
  uint32_t hub_cap;
  uint32_t* chg;               // [0..3] pre-loop rounds, [4..5] loop ring
  unsigned long long* acc;     // [0] recolor slots [1] recolor live
  int64_t* trace;
  int32_t trace_cap;
  int32_t max_rounds;
  FusedResult* res;
};

// the fused driver's scratch block, cleared with one memset per batch
struct FusedScratch {
  uint32_t tcnt[4][4];
  uint32_t fcnt[3][4];
  uint32_t xcnt[4];
  uint32_t chg[8];
  unsigned long long acc[4];
  FusedResult res;
};

__device__ __forceinline__ uint32_t fcount(const Frontier& F) { return F.cnt[0] + F.cnt[1] + F.cnt[3]; }

// a recolor round: the row pass, then (after every chunk is done) the hub
// merge, then the propagation of the changed colours into ccol.  X's counters
// are cleared right after the propagation; the next phase's barrier orders
// that before the next round's pushes.
template <int MODE>
__device__ __forceinline__ void recolor_phase(cg::grid_group& grid, const RecolorArgs& R) {
  recolor_body<MODE>(R);
  grid.sync();
  hub_finalize_body<MODE>(R);
  grid.sync();
  propagate_body(R.G, R.chx, R.nxt, R.cur);
  grid.sync();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    R.chx.cnt[0] = 0;
    R.chx.cnt[1] = 0;
    R.chx.cnt[2] = 0;
    R.chx.cnt[3] = 0;
  }
}

template <int RULE>
__global__ __launch_bounds__(kBlock) void k_fused(FusedArgs A) {
  cg::grid_group grid = cg::this_grid();
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const GraphView& G = A.G;
  int32_t* color[2] = {A.color0, A.color1};
  int cur = A.cur0;
  uint32_t epoch = A.epoch0;
  const uint32_t mepoch = A.epoch0;
  int32_t nph = 0;
  unsigned long long evaluated = 0, changed_total = 0, frontier_sum = 0;
  bool change = false;
  const uint32_t nd = A.ctr->n_del, ni = A.ctr->n_ins, nxd = A.ctr->n_xdel, nxi = A.ctr->n_xins;

  auto record = [&](int64_t kind, int64_t a, int64_t b, int64_t c) {
    if (leader && nph < A.trace_cap) {
      A.trace[4 * nph + 0] = kind;
      A.trace[4 * nph + 1] = a;
      A.trace[4 * nph + 2] = b;
      A.trace[4 * nph + 3] = c;
    }
    ++nph;
  };
  auto rargs = [&](const Frontier& in, uint32_t* chg) {
    RecolorArgs R;
    R.G = G;
    R.cur = color[cur];
    R.nxt = color[cur ^ 1];
    R.prev = A.prev;
    R.in = in;
    R.out = A.Mk;
    R.mmark = A.mmark;
    R.mepoch = mepoch;
    R.pending = nullptr;
    R.partial = A.partial;
    R.chx = A.X;
    R.hub_cap = A.hub_cap;
    R.ctr = A.ctr;
    R.changed = chg;
    R.rslots = &A.acc[0];
    R.rlive = &A.acc[1];
    return R;
  };

  // ---- pre-loop rounds (HI/Color_main.cu:149-174, MG/kernel.cu:299-487,
  //      HB/Color_main.cu:160-163)
  if constexpr (RULE == 0) {
    if (nd > 0) {
      targets_del_body(A.del, &A.ctr->n_del, G, color[cur], A.tmark, ++epoch, A.T0, A.hub_cap);
      grid.sync();
      recolor_phase<M_DEL>(grid, rargs(A.T0, &A.chg[0]));
      uint32_t nt = fcount(A.T0), ch = A.chg[0];
      record(1, nt, ch, 0);
      evaluated += nt;
      changed_total += ch;
      change = change || ch > 0;
      cur ^= 1;
    }
    if (nxd > 0) {
      catchup_body(A.Mk, G, color[cur], color[cur ^ 1]);
      targets_del_body(A.xdel, &A.ctr->n_xdel, G, color[cur], A.tmark, ++epoch, A.T1, A.hub_cap);
      grid.sync();
      recolor_phase<M_DEL>(grid, rargs(A.T1, &A.chg[1]));
      uint32_t nt = fcount(A.T1), ch = A.chg[1];
      record(4, nt, ch, 0);
      evaluated += nt;
      changed_total += ch;
      change = change || ch > 0;
      cur ^= 1;
    }
    if (ni > 0) {
      catchup_body(A.Mk, G, color[cur], color[cur ^ 1]);
      targets_ins_body(A.ins, &A.ctr->n_ins, G, color[cur], nullptr, A.tmark, ++epoch, A.T2, A.hub_cap);
      grid.sync();
      recolor_phase<M_INS>(grid, rargs(A.T2, &A.chg[2]));
      uint32_t nt = fcount(A.T2), ch = A.chg[2];
      record(2, nt, ch, 0);
      evaluated += nt;
      changed_total += ch;
      change = change || nt > 0;
      cur ^= 1;
    }
    if (nxi > 0) {
      catchup_body(A.Mk, G, color[cur], color[cur ^ 1]);
      targets_ins_body(A.xins, &A.ctr->n_xins, G, color[cur], nullptr, A.tmark, ++epoch, A.T3, A.hub_cap);
      grid.sync();
      recolor_phase<M_INS>(grid, rargs(A.T3, &A.chg[3]));
      uint32_t nt = fcount(A.T3), ch = A.chg[3];
      record(5, nt, ch, 0);
      evaluated += nt;
      changed_total += ch;
      change = change || nt > 0;
      cur ^= 1;
    }
  } else {
    if (nd + ni > 0) {
      targets_ldf_body(A.del, &A.ctr->n_del, A.ins, &A.ctr->n_ins, G, A.tmark, ++epoch, A.T0,
                       A.hub_cap);
      grid.sync();
      recolor_phase<M_LDF_P>(grid, rargs(A.T0, &A.chg[0]));
      uint32_t nt = fcount(A.T0), ch = A.chg[0];
      record(3, nt, ch, 0);
      evaluated += nt;
      changed_total += ch;
      change = ch > 0;
      cur ^= 1;
    }
  }

  // ---- recolor loop (HI/Color_main.cu:181-202, HB/Color_main.cu:203-224)
  Frontier L = A.Mk;
  uint32_t nL = fcount(A.Mk);
  int ci = 0;
  int32_t r = 0;
  int32_t status = 0;
  while (change) {
    if (r >= A.max_rounds) {
      status = DC_ERR_ROUNDS;
      break;
    }
    Frontier C = ci == 0 ? A.F0 : A.F1;
    FindArgs FA{G, color[cur], color[cur ^ 1], A.prev, L, C, A.cbits[r & 1], A.cbits[(r & 1) ^ 1],
                A.hub_cap, A.ctr};
    find_body<RULE>(FA);
    grid.sync();
    const uint32_t nC = fcount(C);
    frontier_sum += nL;
    if (nC == 0) {  // nothing to recolor: the reference's last iteration
      record(10, nL, 0, 0);
      ++r;
      break;
    }
    if (leader) {
      if (r >= 1) {  // L is a loop frontier: clear it for reuse as C next round
        L.cnt[0] = 0;
        L.cnt[1] = 0;
        L.cnt[2] = 0;
        L.cnt[3] = 0;
        A.chg[4 + ((r + 1) & 1)] = 0;
      }
    }
    recolor_phase<RULE == 0 ? M_LOOP_HI : M_LOOP_LDF>(grid, rargs(C, &A.chg[4 + (r & 1)]));
    const uint32_t ch = A.chg[4 + (r & 1)];
    record(10, nL, nC, ch);
    evaluated += nC;
    changed_total += ch;
    cur ^= 1;
    change = RULE == 0 ? true : ch > 0;
    L = C;
    nL = nC;
    ci ^= 1;
    ++r;
  }
  // without a loop round nobody copied the pre-loop changes forward
  if (r == 0) catchup_body(A.Mk, G, color[cur], color[cur ^ 1]);
  if (leader) {
    A.res->status = status;
    A.res->cur = cur;
    A.res->rounds = r;
    A.res->phases = nph;
    A.res->epochs = epoch - A.epoch0 + 1;
    A.res->evaluated = evaluated;
    A.res->changed = changed_total;
    A.res->frontier_sum = frontier_sum;
  }
}

// ----------------------------------------------------------------------------
struct FusedState {
  FusedScratch* scratch = nullptr;
  int64_t* trace = nullptr;
  int32_t trace_cap = 0;
  int grid[2] = {0, 0};
};

static FusedState* fstate(Graph* g) { return reinterpret_cast<FusedState*>(g->fused); }

void fused_free(Graph* g) {
  FusedState* f = fstate(g);
  if (!f) return;
  if (f->scratch) (void)hipFree(f->scratch);
  if (f->trace) (void)hipFree(f->trace);
  delete f;
  g->fused = nullptr;
}

static int fused_grid(Graph* g, int rule, int* out) {
  FusedState* f = fstate(g);
  if (f->grid[rule] > 0) {
    *out = f->grid[rule];
    return DC_OK;
  }
  int bpm = 0;
  if (rule == 0)
    DC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpm, k_fused<0>, kBlock, 0));
  else
    DC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpm, k_fused<1>, kBlock, 0));
  hipDeviceProp_t prop;
  DC_HIP(hipGetDeviceProperties(&prop, g->device));
  int per_cu = std::min(bpm, std::max(1, g->fused_blocks_per_cu));
  if (per_cu < 1) {
    set_error("fused kernel cannot be resident on a compute unit");
    return DC_ERR_HIP;
  }
  f->grid[rule] = per_cu * prop.multiProcessorCount;
  *out = f->grid[rule];
  return DC_OK;
}

int recolor_fused(Graph* g, int rule, dc_stats* st) {
  if (!g->staged) {
    set_error("dc_recolor: no staged batch");
    return DC_ERR_STATE;
  }
  const bool partition = g->n_owned < g->n;
  if (rule != DC_RULE_INDEPENDENT && rule != DC_RULE_LDF) {
    set_error("bad rule");
    return DC_ERR_INVALID;
  }
  if (partition && rule != DC_RULE_INDEPENDENT) {
    set_error("partition share supports the independent rule only");
    return DC_ERR_INVALID;
  }
  g->staged = false;
  if (!g->fused) g->fused = new FusedState();
  FusedState* f = fstate(g);
  if (!f->scratch) DC_HIP(hipMalloc(&f->scratch, sizeof(FusedScratch)));
  const int32_t cap = (int32_t)std::min<int64_t>((int64_t)g->opt.max_rounds + 8, 1 << 16);
  if (f->trace_cap < cap) {
    if (f->trace) DC_HIP(hipFree(f->trace));
    f->trace = nullptr;
    DC_HIP(hipMalloc(&f->trace, sizeof(int64_t) * 4 * (size_t)cap));
    f->trace_cap = cap;
  }
  int rc = ensure_epochs(g, (uint32_t)std::min<int64_t>((int64_t)g->opt.max_rounds + 16, 0x7fffffff));
  if (rc) return rc;
  int grid = 0;
  if ((rc = fused_grid(g, rule, &grid))) return rc;

  DC_HIP(hipMemsetAsync(f->scratch, 0, sizeof(FusedScratch), g->stream));
  FusedScratch* s = f->scratch;
  auto frontier_with = [](Frontier base, uint32_t* cnt) {
    base.cnt = cnt;
    return base;
  };
  FusedArgs A;
  A.G = g->view();
  A.color0 = g->color[0];
  A.color1 = g->color[1];
  A.cur0 = g->cur;
  A.prev = g->prev;
  A.tmark = g->tmark;
  A.mmark = g->mmark;
  A.epoch0 = g->epoch + 1;
  A.F0 = frontier_with(g->F[0], s->fcnt[0]);
  A.F1 = frontier_with(g->F[1], s->fcnt[1]);
  A.Mk = frontier_with(g->F[2], s->fcnt[2]);
  A.T0 = frontier_with(g->T, s->tcnt[0]);
  A.T1 = frontier_with(g->T, s->tcnt[1]);
  A.T2 = frontier_with(g->T, s->tcnt[2]);
  A.T3 = frontier_with(g->T, s->tcnt[3]);
  A.X = frontier_with(g->X, s->xcnt);
  A.del = g->del;
  A.ins = g->ins;
  A.xdel = g->xdel;
  A.xins = g->xins;
  A.ctr = g->ctr;
  A.partial = g->partial;
  A.cbits[0] = g->cbits[0];
  A.cbits[1] = g->cbits[1];
  DC_HIP(hipMemsetAsync(g->cbits[0], 0, sizeof(uint32_t) * 2 * g->cbits_words, g->stream));
  A.hub_cap = g->hub_cap;
  A.chg = s->chg;
  A.acc = s->acc;
  A.trace = f->trace;
  A.trace_cap = f->trace_cap;
  A.max_rounds = g->opt.max_rounds;
  A.res = &s->res;

  DC_HIP(hipEventRecord(g->ev[0], g->stream));
  void* args[] = {&A};
  if (rule == 0)
    DC_HIP(hipLaunchCooperativeKernel((const void*)k_fused<0>, dim3(grid), dim3(kBlock), args, 0,
                                      g->stream));
  else
    DC_HIP(hipLaunchCooperativeKernel((const void*)k_fused<1>, dim3(grid), dim3(kBlock), args, 0,
                                      g->stream));
  DC_HIP(hipEventRecord(g->ev[1], g->stream));
  // the loop can end with claims still set (ldf: "no change" exit)
  DC_HIP(hipMemsetAsync(g->cbits[0], 0, sizeof(uint32_t) * 2 * g->cbits_words, g->stream));
  FusedScratch hs;
  DC_HIP(hipMemcpyAsync(&hs, s, sizeof(FusedScratch), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  float ms = 0.f;
  DC_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
  const FusedResult& R = hs.res;
  g->cur = R.cur;
  g->epoch = A.epoch0 + R.epochs;
  int32_t keep = std::min(R.phases, f->trace_cap);
  g->trace.assign((size_t)keep * 4, 0);
  if (keep > 0)
    DC_HIP(hipMemcpy(g->trace.data(), f->trace, sizeof(int64_t) * 4 * (size_t)keep,
                     hipMemcpyDeviceToHost));
  if (st) {
    st->phases = R.phases;
    st->loop_rounds = R.rounds;
    st->evaluated = (int64_t)R.evaluated;
    st->changed = (int64_t)R.changed;
    st->frontier_sum = (int64_t)R.frontier_sum;
    st->ms_recolor = ms;
    st->ms_recolor_kernel = ms;
    st->recolor_launches = 1;
    // read the scan counters accumulated over the whole launch
    Counters c;
    DC_HIP(hipMemcpy(&c, g->ctr, sizeof(Counters), hipMemcpyDeviceToHost));
    st->slots_scanned = (int64_t)(hs.acc[0] + c.slots);
    st->live_scanned = (int64_t)(hs.acc[1] + c.live);
    const uint64_t live_w = rule == DC_RULE_LDF ? 8 : 4;
    // every phase of the launch: recolor (24 B/vertex) + find (16 B/vertex of L)
    st->recolor_bytes = (int64_t)(24ull * R.evaluated + 16ull * R.frontier_sum +
                                  4ull * (hs.acc[0] + c.slots) + live_w * (hs.acc[1] + c.live));
  }
  if (R.status != 0) {
    set_error("recolor loop hit max_rounds");
    return R.status;
  }
  return DC_OK;
}

}  // namespace dc
