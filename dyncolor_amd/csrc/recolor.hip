// recolor.hip -- the recolor rounds of one update batch, GPU static coloring
// and validation.
//
// Semantics: each reference kernel launch is one round; every read in a
// round sees the colors of the round start, every write commits at round end.
// Colors live in two buffers: a round reads cur[] and writes nxt[]; the host
// swaps them between rounds, and vertices changed by the previous round are
// copied forward (cur -> nxt) by the next round's first kernel, which visits
// exactly those vertices.  Every write to vertex t is a function of (t, the
// snapshot), so results do not depend on thread order.
//
// Frontiers are sets of vertex IDs split by row length into three classes
// (8-lane groups / one wave / whole workgroups per 4096-slot chunk); a vertex
// joins a frontier at most once per round through an epoch-stamped claim and a
// wave64 ballot/mbcnt append (the role of the reference's cuCompactor,
// HI/compactor.cuh:79-132, without the O(n) flag scan).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "dc_phases.h"

namespace dc {

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_recolor(RecolorArgs A) {
  recolor_body<MODE>(A);
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_hub_finalize(RecolorArgs A) {
  hub_finalize_body<MODE>(A);
}

__global__ __launch_bounds__(kBlock) void k_propagate(GraphView G, Frontier X, const int32_t* newc,
                                                      const int32_t* oldc) {
  propagate_body(G, X, newc, oldc);
}

template <int RULE>
__global__ __launch_bounds__(kBlock) void k_find(FindArgs A) {
  find_body<RULE>(A);
}

__global__ void k_targets_del(const int2* del, const uint32_t* ndel_p, GraphView G, const int32_t* cur,
                              uint32_t* tmark, uint32_t epoch, Frontier T, uint32_t hub_cap) {
  targets_del_body(del, ndel_p, G, cur, tmark, epoch, T, hub_cap);
}

__global__ void k_targets_ldf(const int2* del, const uint32_t* ndel_p, const int2* ins,
                              const uint32_t* nins_p, GraphView G, uint32_t* tmark, uint32_t epoch,
                              Frontier T, uint32_t hub_cap) {
  targets_ldf_body(del, ndel_p, ins, nins_p, G, tmark, epoch, T, hub_cap);
}

__global__ __launch_bounds__(kBlock) void k_targets_ins(const int2* ins, const uint32_t* nins_p,
                                                        GraphView G, const int32_t* cur,
                                                        const int32_t* cmax, uint32_t* tmark,
                                                        uint32_t epoch, Frontier T,
                                                        uint32_t hub_cap) {
  targets_ins_body(ins, nins_p, G, cur, cmax, tmark, epoch, T, hub_cap);
}

__global__ __launch_bounds__(kBlock) void k_colour_max(const int32_t* cur, int32_t n, int32_t* cmax) {
  colour_max_body(cur, n, cmax);
}

__global__ void k_catchup(Frontier F, GraphView G, const int32_t* cur, int32_t* nxt) {
  catchup_body(F, G, cur, nxt);
}

// Start of a pre-loop round: the target frontier's and the changed list's
// counts, the round's counters and (insertion rounds) the colour maximum
// start at zero, in one launch.
__global__ void k_round_reset(uint32_t* t_cnt, uint32_t* x_cnt, Counters* ctr, int32_t* cmax) {
  const int i = threadIdx.x;
  if (i < 4) {
    t_cnt[i] = 0u;
    x_cnt[i] = 0u;
  }
  if (i == 0) {
    ctr->changed = 0u;
    ctr->targets = 0u;
    ctr->slots = 0ull;
    ctr->live = 0ull;
    ctr->rslots = 0ull;
    ctr->rlive = 0ull;
    if (cmax) *cmax = 0;
  }
}

// End of a pre-loop round: the marked set's colours into the other buffer,
// and the round's frontier counts and counters into its read-back slot.
__global__ void k_catchup_save(Frontier F, GraphView G, const int32_t* cur, int32_t* nxt,
                               const uint32_t* t_cnt, const Counters* ctr, uint32_t* s_cnt,
                               Counters* s_c) {
  catchup_body(F, G, cur, nxt);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int k = 0; k < 4; ++k) s_cnt[k] = t_cnt[k];
    *s_c = *ctr;
  }
}

// ----------------------------------------------------------------------------
// static coloring: pending[v] = number of predecessors (neighbours that come
// first in the greedy order); vertices with none form the first ready set.
template <int MODE>
__global__ void k_static_init(GraphView G, int32_t* pending, Frontier R, uint32_t hub_cap) {
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v0 = gwave; v0 < G.n_owned; v0 += nwaves) {
    const int32_t v = (int32_t)v0;
    const uint2 r = G.row[v];
    const int32_t pv = ModeTraits<MODE>::prop ? G.prop[v] : 0;
    uint32_t cnt = 0;
    for (uint32_t j = lane; j < r.y; j += 64) {
      int32_t w = G.col[(uint64_t)r.x + j];
      if (w < 0 || w >= G.n_owned) continue;
      int32_t pw = ModeTraits<MODE>::prop ? G.prop[w] : 0;
      if (include_nbr<MODE>(v, pv, w, pw, G.n_owned)) ++cnt;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (lane == 0) pending[v] = (int32_t)cnt;
    frontier_push(lane == 0 && cnt == 0, v, G, R, hub_cap);
  }
}

// ----------------------------------------------------------------------------
// validate: monochromatic live entries (v, w) with v < w or w a ghost, and
// the highest owned color (HI/gpuFunctions_undir.cuh:289-305,
// HI/Color_main.cu:218-226).
__global__ void k_validate(GraphView G, const int32_t* color, Counters* ctr) {
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long bad = 0;
  int32_t mc = -1;
  for (int64_t v0 = gwave; v0 < G.n_owned; v0 += nwaves) {
    const int32_t v = (int32_t)v0;
    const uint2 r = G.row[v];
    const int32_t cv = color[v];
    mc = max(mc, cv);
    for (uint32_t j = lane; j < r.y; j += 64) {
      int32_t w = G.col[(uint64_t)r.x + j];
      if (w < 0) continue;
      if ((w > v || w >= G.n_owned) && color[w] == cv) ++bad;
    }
  }
  wave_add_u64(bad, &ctr->mono);
  for (int off = 32; off > 0; off >>= 1) mc = max(mc, __shfl_xor(mc, off));
  if (lane == 0) atomicMax(&ctr->maxc, mc);
}

// ============================================================================
// host side
static int grid_waves(int64_t waves) {
  int64_t b = (waves * 64 + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}
static int grid_threads(int64_t t) {
  int64_t b = (t + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}

// Persistent-style grid for frontier kernels: enough workgroups to cover the
// device; each loop inside reads its own item counts from device memory.
static int frontier_grid(const Graph* g) {
  (void)g;
  return 2048;
}

int ensure_epochs(Graph* g, uint32_t needed) {
  if ((uint64_t)g->epoch + needed < 0xfffffff0ull) return DC_OK;
  DC_HIP(hipMemsetAsync(g->tmark, 0, sizeof(uint32_t) * ((size_t)g->n + 1), g->stream));
  DC_HIP(hipMemsetAsync(g->mmark, 0, sizeof(uint32_t) * ((size_t)g->n + 1), g->stream));
  g->epoch = 0;
  return DC_OK;
}

static int zero_frontier(Graph* g, Frontier f) {
  DC_HIP(hipMemsetAsync(f.cnt, 0, sizeof(uint32_t) * 4, g->stream));
  return DC_OK;
}

static RecolorArgs make_rargs(Graph* g, Frontier in, Frontier out, uint32_t mepoch) {
  RecolorArgs A;
  A.G = g->view();
  A.cur = g->color[g->cur];
  A.nxt = g->color[g->cur ^ 1];
  A.prev = g->prev;
  A.in = in;
  A.out = out;
  A.mmark = g->mmark;
  A.mepoch = mepoch;
  A.pending = g->pending;
  A.partial = g->partial;
  A.chx = g->X;
  A.hub_cap = g->hub_cap;
  A.ctr = g->ctr;
  A.changed = &g->ctr->changed;
  A.rslots = &g->ctr->rslots;
  A.rlive = &g->ctr->rlive;
  return A;
}

// Profiled launches record a start/stop event pair from a pool; the pairs are
// read once at the end of the batch so profiling adds no host round trip.
static int prof_begin(Graph* g) {
  if (!g->profiling) return DC_OK;
  while (g->evpool.size() < g->evused + 2) {
    hipEvent_t e;
    DC_HIP(hipEventCreate(&e));
    g->evpool.push_back(e);
  }
  DC_HIP(hipEventRecord(g->evpool[g->evused], g->stream));
  return DC_OK;
}
static int prof_end(Graph* g) {
  if (!g->profiling) return DC_OK;
  DC_HIP(hipEventRecord(g->evpool[g->evused + 1], g->stream));
  g->evused += 2;
  return DC_OK;
}
static int prof_collect(Graph* g, double* ms) {
  double t = 0;
  if (g->evused) DC_HIP(hipEventSynchronize(g->evpool[g->evused - 1]));
  for (size_t i = 0; i + 1 < g->evused; i += 2) {
    float x = 0.f;
    DC_HIP(hipEventElapsedTime(&x, g->evpool[i], g->evpool[i + 1]));
    t += x;
  }
  g->evused = 0;
  if (ms) *ms = t;
  return DC_OK;
}

// One recolor = the row pass over every class, then the hub merge pass, then
// (independent/ldf modes) the propagation of the changed colours into ccol.
template <int MODE>
static void launch_pair(Graph* g, dim3 grid, const RecolorArgs& A) {
  hipLaunchKernelGGL(k_recolor<MODE>, grid, dim3(kBlock), 0, g->stream, A);
  hipLaunchKernelGGL(k_hub_finalize<MODE>, grid, dim3(kBlock), 0, g->stream, A);
  if constexpr (!ModeTraits<MODE>::stat)
    hipLaunchKernelGGL(k_propagate, grid, dim3(kBlock), 0, g->stream, A.G, A.chx, A.nxt, A.cur);
}

static int launch_recolor(Graph* g, int mode, const RecolorArgs& A, float* ms,
                          bool clear_chx = true) {
  (void)ms;
  dim3 grid(frontier_grid(g));
  if (clear_chx) DC_HIP(hipMemsetAsync(A.chx.cnt, 0, sizeof(uint32_t) * 4, g->stream));
  int prc = prof_begin(g);
  if (prc) return prc;
  switch (mode) {
    case M_DEL: launch_pair<M_DEL>(g, grid, A); break;
    case M_INS: launch_pair<M_INS>(g, grid, A); break;
    case M_LDF_P: launch_pair<M_LDF_P>(g, grid, A); break;
    case M_LOOP_HI: launch_pair<M_LOOP_HI>(g, grid, A); break;
    case M_LOOP_LDF: launch_pair<M_LOOP_LDF>(g, grid, A); break;
    case M_STATIC_ID: launch_pair<M_STATIC_ID>(g, grid, A); break;
    case M_STATIC_LDF: launch_pair<M_STATIC_LDF>(g, grid, A); break;
    default: set_error("bad recolor mode"); return DC_ERR_INVALID;
  }
  DC_HIP(hipGetLastError());
  return prof_end(g);
}

struct RoundReadback {
  uint32_t cnt[4];
  Counters c;
};

static int readback(Graph* g, Frontier f, RoundReadback* rb) {
  DC_HIP(hipMemcpyAsync(rb->cnt, f.cnt, sizeof(uint32_t) * 4, hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipMemcpyAsync(g->hctr, g->ctr, sizeof(Counters), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  rb->c = *g->hctr;
  return DC_OK;
}

static inline uint64_t fsize(const uint32_t* cnt) { return (uint64_t)cnt[0] + cnt[1] + cnt[3]; }

// The plain per-batch driver: one launch per phase, host reads the frontier
// size after every round (the structure of HI/Color_main.cu:146-202 and
// HB/Color_main.cu:160-224, with the O(n) memsets and compactions replaced by
// epoch claims and wave-aggregated appends).
int recolor_plain(Graph* g, int rule, dc_stats* st, bool persistent_loop) {
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
  g->trace.clear();
  int rc;
  // recolor-kernel time comes from the profiling event pairs (prof_collect)
  int launches = 0;
  uint64_t evaluated = 0, changed_total = 0, frontier_sum = 0;

  rc = ensure_epochs(g, (uint32_t)std::min<int64_t>((int64_t)g->opt.max_rounds + 16, 0x7fffffff));
  if (rc) return rc;
  DC_HIP(hipEventRecord(g->ev[0], g->stream));

  // list sizes staged by the apply (kept on the host by it)
  RoundReadback rb;
  Counters staged{};
  staged.n_del = g->staged_n[0];
  staged.n_ins = g->staged_n[1];
  staged.n_xdel = g->staged_n[2];
  staged.n_xins = g->staged_n[3];
  // pre-loop rounds' results: copied on the device after each round, read
  // back together before the loop (one host wait instead of one per round)
  constexpr int kPreSlots = 5;   // up to four target rounds, then the marked set's counts
  if (!g->pre_dev) {
    DC_HIP(hipMalloc(&g->pre_dev, sizeof(RoundReadback) * kPreSlots));
    DC_HIP(hipHostMalloc(&g->pre_host, sizeof(RoundReadback) * kPreSlots, hipHostMallocDefault));
  }
  RoundReadback* pre_d = static_cast<RoundReadback*>(g->pre_dev);
  RoundReadback* pre_h = static_cast<RoundReadback*>(g->pre_host);
  struct PreRound {
    int kind, mode;
  };
  PreRound pre[kPreSlots - 1];
  int npre = 0;
  uint32_t* dctr_u32 = reinterpret_cast<uint32_t*>(g->ctr);
  const uint32_t* d_ndel = dctr_u32 + offsetof(Counters, n_del) / 4;
  const uint32_t* d_nins = dctr_u32 + offsetof(Counters, n_ins) / 4;
  const uint32_t* d_nxdel = dctr_u32 + offsetof(Counters, n_xdel) / 4;
  const uint32_t* d_nxins = dctr_u32 + offsetof(Counters, n_xins) / 4;
  // reset the per-round counters (keep the staged list sizes)
  auto reset_round_ctr = [&]() -> int {
    DC_HIP(hipMemsetAsync(&g->ctr->changed, 0, sizeof(uint32_t) * 2, g->stream));
    DC_HIP(hipMemsetAsync(&g->ctr->slots, 0, sizeof(unsigned long long) * 4, g->stream));
    return DC_OK;
  };
  uint64_t slots_total = 0, live_total = 0;

  // Algorithmic bytes of one recolor launch: per evaluated vertex its list
  // entry (4), row bounds (8), own color read (4), color write (4) and prev
  // write or own priority read (4); per row slot the neighbour ID (4); per
  // live neighbour its color (4) and, for LDF, its priority (4).
  uint64_t recolor_bytes = 0;
  auto account_recolor = [&](uint64_t nvert, const Counters& c) {
    recolor_bytes += 24ull * nvert + 4ull * c.rslots +
                     (rule == DC_RULE_LDF ? 8ull : 4ull) * c.rlive;
  };

  Frontier Mk = g->F[2];
  if ((rc = zero_frontier(g, Mk))) return rc;
  const uint32_t mepoch = next_epoch(g);
  bool change = false;
  const dim3 blk(kBlock);

  auto run_targets_round = [&](int kind, int mode, auto launch_targets) -> int {
    int r2;
    RecolorArgs A = make_rargs(g, g->T, Mk, mepoch);
    hipLaunchKernelGGL(k_round_reset, dim3(1), dim3(64), 0, g->stream, g->T.cnt, A.chx.cnt, g->ctr,
                       mode == M_INS ? &g->ctr->cmax : nullptr);
    launch_targets(next_epoch(g));
    DC_HIP(hipGetLastError());
    if ((r2 = launch_recolor(g, mode, A, nullptr, false))) return r2;
    ++launches;
    pre[npre] = PreRound{kind, mode};
    g->cur ^= 1;
    hipLaunchKernelGGL(k_catchup_save, dim3(grid_threads(g->n_owned)), blk, 0, g->stream, Mk,
                       g->view(), g->color[g->cur], g->color[g->cur ^ 1], g->T.cnt, g->ctr,
                       pre_d[npre].cnt, &pre_d[npre].c);
    DC_HIP(hipGetLastError());
    ++npre;
    return DC_OK;
  };

  const int fg = frontier_grid(g);
  // highest current colour, for the insertion targets' histogram test
  // (cmax was zeroed by the round's reset)
  auto launch_colour_max = [&]() {
    hipLaunchKernelGGL(k_colour_max, dim3(grid_threads(g->n) < 256 ? grid_threads(g->n) : 256),
                       blk, 0, g->stream, g->color[g->cur], g->n, &g->ctr->cmax);
  };
  if (rule == DC_RULE_INDEPENDENT) {
    if (staged.n_del > 0) {
      rc = run_targets_round(1, M_DEL, [&](uint32_t e) {
        hipLaunchKernelGGL(k_targets_del, dim3(grid_threads(staged.n_del)), blk, 0, g->stream,
                           g->del, d_ndel, g->view(), g->color[g->cur], g->tmark, e, g->T,
                           g->hub_cap);
      });
      if (rc) return rc;
    }
    if (staged.n_xdel > 0) {
      rc = run_targets_round(4, M_DEL, [&](uint32_t e) {
        hipLaunchKernelGGL(k_targets_del, dim3(grid_threads(staged.n_xdel)), blk, 0, g->stream,
                           g->xdel, d_nxdel, g->view(), g->color[g->cur], g->tmark, e, g->T,
                           g->hub_cap);
      });
      if (rc) return rc;
    }
    if (staged.n_ins > 0) {
      rc = run_targets_round(2, M_INS, [&](uint32_t e) {
        launch_colour_max();
        hipLaunchKernelGGL(k_targets_ins, dim3(grid_threads((int64_t)staged.n_ins * kGroup)), blk, 0, g->stream,
                           g->ins, d_nins, g->view(), g->color[g->cur], &g->ctr->cmax, g->tmark, e,
                           g->T, g->hub_cap);
      });
      if (rc) return rc;
    }
    if (staged.n_xins > 0) {
      rc = run_targets_round(5, M_INS, [&](uint32_t e) {
        launch_colour_max();
        hipLaunchKernelGGL(k_targets_ins, dim3(grid_threads((int64_t)staged.n_xins * kGroup)), blk, 0, g->stream,
                           g->xins, d_nxins, g->view(), g->color[g->cur], &g->ctr->cmax, g->tmark,
                           e, g->T, g->hub_cap);
      });
      if (rc) return rc;
    }
  } else {
    if (staged.n_del + staged.n_ins > 0) {
      rc = run_targets_round(3, M_LDF_P, [&](uint32_t e) {
        hipLaunchKernelGGL(k_targets_ldf, dim3(grid_threads(staged.n_del + staged.n_ins)), blk, 0,
                           g->stream, g->del, d_ndel, g->ins, d_nins, g->view(), g->tmark, e, g->T,
                           g->hub_cap);
      });
      if (rc) return rc;
    }
  }

  // recolor loop: L = marked set, then L <- C every round
  Frontier L = Mk;
  DC_HIP(hipMemsetAsync(g->cbits[0], 0, sizeof(uint32_t) * 2 * g->cbits_words, g->stream));
  // the pre-loop rounds' results and the marked set's size in one read-back
  DC_HIP(hipMemcpyAsync(pre_d[kPreSlots - 1].cnt, Mk.cnt, sizeof(uint32_t) * 4,
                        hipMemcpyDeviceToDevice, g->stream));
  DC_HIP(hipMemcpyAsync(pre_h, pre_d, sizeof(RoundReadback) * kPreSlots, hipMemcpyDeviceToHost,
                        g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  for (int i = 0; i < npre; ++i) {
    const RoundReadback& p = pre_h[i];
    const uint64_t nt = fsize(p.cnt);
    g->trace.insert(g->trace.end(), {(int64_t)pre[i].kind, (int64_t)nt, (int64_t)p.c.changed, 0});
    evaluated += nt;
    changed_total += p.c.changed;
    slots_total += p.c.slots + p.c.rslots;
    live_total += p.c.live + p.c.rlive;
    account_recolor(nt, p.c);
    if (pre[i].mode == M_INS) {
      if (nt > 0) change = true;
    } else if (p.c.changed > 0) {
      change = true;
    }
  }
  uint64_t nL = fsize(pre_h[kPreSlots - 1].cnt);
  // the tombstone pass's row-slot sum stays in the counters through the
  // pre-loop rounds (no round resets it)
  const int64_t del_row_slots = npre > 0 ? (int64_t)pre_h[npre - 1].c.delslots : 0;
  const bool ldf_counts = rule == DC_RULE_LDF && g->tuning.ldf_loop == 0;
  if ((rule == DC_RULE_INDEPENDENT || ldf_counts) && change) {
    // candidate counts kept incrementally (loop.hip, loop_ldf.hip); both
    // colour buffers are written there, so no catch-up follows
    LoopResult lr;
    if (rule == DC_RULE_INDEPENDENT) {
      if ((rc = loop_hi(g, Mk, (int32_t)nL, persistent_loop, &lr))) return rc;
    } else if ((rc = loop_ldf(g, Mk, &lr))) {
      return rc;
    }
    evaluated += lr.evaluated;
    changed_total += lr.changed;
    frontier_sum += lr.frontier_sum;
    slots_total += lr.slots + lr.rslots;
    live_total += lr.live + lr.rlive;
    recolor_bytes += 24ull * lr.evaluated + 4ull * lr.rslots;
    launches += (int)lr.launches;
    change = false;
    DC_HIP(hipEventRecord(g->ev[1], g->stream));
    DC_HIP(hipEventSynchronize(g->ev[1]));
    float ms = 0.f;
    DC_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
    double kernel_ms = 0.0;
    if ((rc = prof_collect(g, &kernel_ms))) return rc;
    if (st) {
      st->recolor_bytes = (int64_t)recolor_bytes;
      st->phases = (int32_t)(g->trace.size() / 4);
      st->loop_rounds = lr.rounds;
      st->loop_settled = lr.settled ? 1 : 0;
      st->pull_rounds = lr.pull_rounds;
      st->heavy_push_rounds = lr.heavy_push_rounds;
      st->heavy_push_listed = lr.heavy_push_listed;
      st->graph_captures = lr.graph_captures;
      st->lc_vertices = (int64_t)lr.lc_vertices;
      st->lc_slots = (int64_t)lr.lc_slots;
      st->del_row_slots = g->profiling ? del_row_slots : 0;
      st->evaluated = (int64_t)evaluated;
      st->changed = (int64_t)changed_total;
      st->frontier_sum = (int64_t)frontier_sum;
      st->slots_scanned = (int64_t)slots_total;
      st->live_scanned = (int64_t)live_total;
      st->ms_recolor = ms;
      st->ms_recolor_kernel = kernel_ms + lr.kernel_ms;
      st->recolor_launches = launches;
      st->ms_count_pass = lr.count_ms;
      st->ms_round_a = lr.round_a_ms;
      st->moved_slots = (int64_t)lr.moved_slots;
      st->slots_rewritten = lr.slots_rewritten ? 1 : 0;
      st->count_pass_bytes = (int64_t)lr.count_bytes;
      st->count_pass_launches = (int32_t)lr.count_launches;
    }
    return DC_OK;
  }
  int cidx = 0;
  int32_t rounds = 0;
  while (change) {
    if (rounds >= g->opt.max_rounds) {
      set_error("recolor loop hit max_rounds");
      return DC_ERR_ROUNDS;
    }
    Frontier C = g->F[cidx];
    if ((rc = zero_frontier(g, C))) return rc;
    if ((rc = reset_round_ctr())) return rc;
    FindArgs FA{g->view(), g->color[g->cur], g->color[g->cur ^ 1], g->prev, L, C,
                g->cbits[rounds & 1], g->cbits[(rounds & 1) ^ 1], g->hub_cap, g->ctr};
    if (rule == DC_RULE_INDEPENDENT)
      hipLaunchKernelGGL(k_find<0>, dim3(fg), blk, 0, g->stream, FA);
    else
      hipLaunchKernelGGL(k_find<1>, dim3(fg), blk, 0, g->stream, FA);
    DC_HIP(hipGetLastError());
    RecolorArgs A = make_rargs(g, C, Mk, 0);
    rc = launch_recolor(g, rule == DC_RULE_INDEPENDENT ? M_LOOP_HI : M_LOOP_LDF, A, nullptr);
    if (rc) return rc;
    ++launches;
    if ((rc = readback(g, C, &rb))) return rc;
    uint64_t nC = fsize(rb.cnt);
    if (rb.cnt[3] > g->hub_cap) {
      set_error("hub slot capacity exceeded");
      return DC_ERR_NOMEM;
    }
    g->trace.insert(g->trace.end(), {(int64_t)10, (int64_t)nL, (int64_t)nC, (int64_t)rb.c.changed});
    evaluated += nC;
    changed_total += rb.c.changed;
    frontier_sum += nL;
    slots_total += rb.c.slots + rb.c.rslots;
    live_total += rb.c.live + rb.c.rlive;
    account_recolor(nC, rb.c);
    g->cur ^= 1;
    change = rule == DC_RULE_INDEPENDENT ? nC > 0 : rb.c.changed > 0;
    L = C;
    nL = nC;
    cidx ^= 1;
    ++rounds;
  }
  // the write buffer still lags on the last round's changes
  if (rounds > 0) {
    hipLaunchKernelGGL(k_catchup, dim3(grid_threads(std::max<uint64_t>(nL, 1))), blk, 0, g->stream,
                       L, g->view(), g->color[g->cur], g->color[g->cur ^ 1]);
    DC_HIP(hipGetLastError());
  }
  // an ldf loop ends on "no change" with its last claims still set
  DC_HIP(hipMemsetAsync(g->cbits[0], 0, sizeof(uint32_t) * 2 * g->cbits_words, g->stream));
  DC_HIP(hipEventRecord(g->ev[1], g->stream));
  DC_HIP(hipEventSynchronize(g->ev[1]));
  float ms = 0.f;
  DC_HIP(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
  double kernel_ms = 0.0;
  if ((rc = prof_collect(g, &kernel_ms))) return rc;
  if (st) {
    st->recolor_bytes = (int64_t)recolor_bytes;
    st->phases = (int32_t)(g->trace.size() / 4);
    st->loop_rounds = rounds;
    st->evaluated = (int64_t)evaluated;
    st->changed = (int64_t)changed_total;
    st->frontier_sum = (int64_t)frontier_sum;
    st->slots_scanned = (int64_t)slots_total;
    st->live_scanned = (int64_t)live_total;
    st->ms_recolor = ms;
    st->ms_recolor_kernel = kernel_ms;
    st->recolor_launches = launches;
    st->del_row_slots = g->profiling ? del_row_slots : 0;
  }
  return DC_OK;
}

// ----------------------------------------------------------------------------
int static_color(Graph* g, int rule) {
  if (rule != DC_RULE_INDEPENDENT && rule != DC_RULE_LDF) {
    set_error("bad rule");
    return DC_ERR_INVALID;
  }
  const int mode = rule == DC_RULE_INDEPENDENT ? M_STATIC_ID : M_STATIC_LDF;
  const dim3 blk(kBlock);
  Frontier R = g->F[0], R2 = g->F[1];
  int rc;
  if ((rc = zero_frontier(g, R))) return rc;
  if (mode == M_STATIC_ID)
    hipLaunchKernelGGL(k_static_init<M_STATIC_ID>, dim3(grid_waves(g->n_owned)), blk, 0, g->stream,
                       g->view(), g->pending, R, g->hub_cap);
  else
    hipLaunchKernelGGL(k_static_init<M_STATIC_LDF>, dim3(grid_waves(g->n_owned)), blk, 0, g->stream,
                       g->view(), g->pending, R, g->hub_cap);
  DC_HIP(hipGetLastError());
  uint32_t cnt[4];
  DC_HIP(hipMemcpyAsync(cnt, R.cnt, sizeof(cnt), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  uint64_t colored = 0;
  while (fsize(cnt) > 0) {
    colored += fsize(cnt);
    if ((rc = zero_frontier(g, R2))) return rc;
    RecolorArgs A = make_rargs(g, R, R2, 0);
    A.nxt = g->color[g->cur];
    if ((rc = launch_recolor(g, mode, A, nullptr))) return rc;
    DC_HIP(hipMemcpyAsync(cnt, R2.cnt, sizeof(cnt), hipMemcpyDeviceToHost, g->stream));
    DC_HIP(hipStreamSynchronize(g->stream));
    std::swap(R, R2);
  }
  if (colored != (uint64_t)g->n_owned) {
    set_error("static coloring did not reach every vertex");
    return DC_ERR_HIP;
  }
  DC_HIP(hipMemcpyAsync(g->color[g->cur ^ 1], g->color[g->cur], sizeof(int32_t) * (size_t)g->n,
                        hipMemcpyDeviceToDevice, g->stream));
  if ((rc = rebuild_ccol(g))) return rc;
  DC_HIP(hipStreamSynchronize(g->stream));
  return DC_OK;
}

int validate(Graph* g, int64_t* mono, int32_t* maxc) {
  DC_HIP(hipMemsetAsync(&g->ctr->mono, 0, sizeof(unsigned long long), g->stream));
  int32_t neg = -1;
  DC_HIP(hipMemcpyAsync(&g->ctr->maxc, &neg, sizeof(int32_t), hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(k_validate, dim3(grid_waves(std::max(g->n_owned, 1))), dim3(kBlock), 0,
                     g->stream, g->view(), g->color[g->cur], g->ctr);
  DC_HIP(hipGetLastError());
  DC_HIP(hipMemcpyAsync(g->hctr, g->ctr, sizeof(Counters), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  if (mono) *mono = (int64_t)g->hctr->mono;
  if (maxc) *maxc = g->hctr->maxc;
  return DC_OK;
}

}  // namespace dc
