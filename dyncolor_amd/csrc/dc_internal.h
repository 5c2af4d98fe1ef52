// dc_internal.h -- device-resident graph state shared by the library's
// translation units.  Host-side only declarations plus POD structs that are
// passed to kernels by value.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dyncolor.h"

namespace dc {

// Degree classes of a frontier.  A visited row is read by
//   S: an 8-lane group          (used slots < kSmallLen, mex fits one u64),
//   M: one wave64               (kSmallLen <= slots < kMidLen, mex < 512),
//   H: whole workgroups, one per kChunk-slot chunk (slots >= kMidLen), partial
//      masks merged in a per-hub global window.
// With 4 loads in flight per lane no lane walks more than 8 slots of a row.
constexpr int kBlock = 256;
constexpr int kGroup = 8;
constexpr uint32_t kSmallLen = 64;
constexpr uint32_t kMidLen = 512;
constexpr uint32_t kChunk = 1024;
constexpr int kWinBits = 8192;               // LDS mex window (bits)
constexpr int kWinWords = kWinBits / 64;     // u64 words per window
constexpr int kHistW = 512;                  // colour histogram width (colours 0..511)

// One frontier = three class lists + counters.
struct Frontier {
  int32_t* s;      // class S vertex list
  int32_t* m;      // class M vertex list
  int4* h;         // class H chunks: {vertex, chunk index, hub slot, chunks}
  uint32_t* cnt;   // [0]=|S| [1]=|M| [2]=#H chunks [3]=#H vertices (hub slots)
};

// Per-round device counters (one cache line each group).
struct Counters {
  uint32_t changed;
  uint32_t targets;
  uint32_t overflow;      // insertion slots past row capacity
  uint32_t n_del, n_ins, n_xdel, n_xins;
  uint32_t pad0;
  unsigned long long slots;   // row slots scanned (find / target kernels)
  unsigned long long live;    // live entries scanned (find / target kernels)
  unsigned long long rslots;  // row slots scanned by the recolor kernel
  unsigned long long rlive;   // live entries scanned by the recolor kernel
  unsigned long long tomb;    // tombstones written
  unsigned long long mono;    // validate
  int32_t maxc;
  int32_t window_full;    // a hub mex window was full -> slow path used
  int32_t cmax;           // highest colour (insertion round: are histograms complete?)
  uint32_t xpart;         // insertion lines between two parts (dc_set_parts)
  unsigned long long delslots;  // deletion lines: used slots of both endpoints' rows
};

// Everything a kernel needs about one graph, passed by value.
//
// Every slot also carries its neighbour's current colour (ccol, -1 for a
// tombstone) and the position of the reverse slot inside the neighbour's row
// (twin, -1 when the neighbour is a ghost, whose row is empty).  Row visits
// then stream ccol instead of gathering colour[col[j]]; when a vertex changes
// colour, the following propagate pass rewrites ccol at its twins.
struct GraphView {
  int32_t n;          // total vertices (owned + ghosts)
  int32_t n_owned;    // vertices >= n_owned are frozen ghosts
  uint2* row;         // {start, used slots}
  uint32_t* cap;      // row capacity
  int32_t* col;       // slot -> neighbour, -1 tombstone
  int32_t* prop;      // LDF priority counter
  int32_t* ccol;      // slot -> neighbour's colour (snapshot of the current round)
  int32_t* twin;      // slot -> reverse slot position in the neighbour's row
  int32_t* mexc;      // vertex -> mex of all its live neighbours' colours, -1 = stale
  int32_t* hidx;      // vertex -> histogram row, -1 for rows with cap < kSmallLen
  uint32_t* hist;     // hub row -> live neighbours per colour 0..kHistW-1
  const int32_t* vpart;  // owned vertex -> its part (dc_set_parts), or null
};

struct Batch {
  int64_t k = 0;
  int32_t* u = nullptr;
  int32_t* v = nullptr;
  uint8_t* ins = nullptr;
};

struct Graph {
  int device = 0;
  dc_options opt{};
  dc_tuning tuning{};
  int32_t n = 0, n_owned = 0;
  uint64_t total_cap = 0;     // col[] slots (rows, then a spare tail for relocated rows)
  uint64_t rows_end = 0;      // end of the laid-out rows = start of the spare tail
  unsigned long long* tail = nullptr;  // device: next free slot of the spare tail
  int32_t staged_moves = 0;   // rows relocated into the tail by the staged batch
  hipStream_t stream = nullptr;

  // CSR
  uint2* row = nullptr;
  uint32_t* cap = nullptr;
  int32_t* col = nullptr;
  int32_t* prop = nullptr;
  int32_t* ccol = nullptr;
  int32_t* twin = nullptr;
  // Cached mex over all live neighbours (the mex of the deletion, insertion
  // and independent-rule loop rounds).  A row edit or a neighbour's colour
  // change sets it to -1; the next evaluation of the vertex rescans its row.
  int32_t* mexc = nullptr;
  // Rows with cap >= kSmallLen also keep a histogram of their live
  // neighbours' colours below kHistW, so their mex is a short read instead of
  // a row scan.
  int32_t* hidx = nullptr;
  uint32_t* hist = nullptr;
  int32_t* hrowv = nullptr;     // histogram row -> vertex
  uint32_t n_hrows = 0, hist_cap = 0;
  // every kChunk-slot chunk of every hub row's capacity: {vertex, chunk}
  int2* hch = nullptr;
  uint32_t n_hch = 0, hch_cap = 0;

  // colors: double buffer for round-synchronous snapshots
  int32_t* color[2] = {nullptr, nullptr};
  int cur = 0;
  int32_t* prev = nullptr;

  // marks (epoch stamped)
  uint32_t* tmark = nullptr;  // per-round target/candidate dedupe
  uint32_t* mmark = nullptr;  // per-batch marked-set dedupe
  uint32_t epoch = 0;
  uint32_t* cbits[2] = {nullptr, nullptr};  // loop-round claim bitmaps, 1 bit per vertex
  size_t cbits_words = 0;                   // words per bitmap (both share one allocation)

  // frontiers: F[0], F[1] alternate; M = pre-loop marked set; T = round targets;
  // X = vertices whose colour changed in the current round (propagate input)
  Frontier F[3] = {};
  Frontier T = {};
  Frontier X = {};
  int32_t* fbuf = nullptr;    // backing store of the int lists
  int4* hbuf = nullptr;       // backing store of the hub-chunk lists
  uint32_t* fcnt = nullptr;   // counters of all frontiers
  uint32_t hub_cap = 0;       // max hub slots per round
  uint32_t hchunk_cap = 0;    // max hub chunks per frontier
  unsigned long long* partial = nullptr;  // per hub chunk: partial mex window (16 u64)

  // staged batch lists: int2 pairs
  int2* del = nullptr;
  int2* ins = nullptr;
  int2* xdel = nullptr;
  int2* xins = nullptr;
  int4* ovf = nullptr;        // insertion overflow {vertex, pos, value, 0}
  int4* rch = nullptr;        // row relocation copies {old start, new start, slots, 0}
  uint64_t rch_cap = 0;
  int64_t list_cap = 0;
  // deletions grouped by scanned row (csr_apply): keys [2][list_cap], runs
  // over long rows, radix-sort storage
  uint64_t* tkeys = nullptr;
  int2* tlong = nullptr;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  bool staged = false;
  int64_t staged_k = 0, staged_ins = 0, staged_del = 0;
  int32_t staged_regrows = 0;

  Counters* ctr = nullptr;    // device
  Counters* hctr = nullptr;   // pinned host mirror
  // the staged batch's deletion, insertion, cross-deletion and cross-insertion
  // line counts (host copy from the apply)
  uint32_t staged_n[4] = {0u, 0u, 0u, 0u};
  // disjoint parts of an arena (dc_set_parts): vertex -> part, and per part
  // {changes by round parity [2], check failed, left the loop}
  int32_t nparts = 0;
  int32_t* vpart = nullptr;
  uint32_t* pstate = nullptr;
  // pre-loop rounds' results, gathered on the device and read back once:
  // kPreSlots x {frontier counts, Counters} (device and pinned host)
  void* pre_dev = nullptr;
  void* pre_host = nullptr;
  uint32_t* aux = nullptr;    // small device scratch (hub capacity count)

  // upload staging for host-array batches
  Batch scratch_batch;

  // scan scratch
  uint64_t* scan_tmp = nullptr;
  size_t scan_tmp_n = 0;

  // static coloring
  int32_t* pending = nullptr;

  // trace of the last recolor
  std::vector<int64_t> trace;

  bool profiling = false;
  hipEvent_t ev[4] = {};
  std::vector<hipEvent_t> evpool;   // start/stop pairs of profiled launches
  size_t evused = 0;

  // persistent driver state (fused.hip)
  void* fused = nullptr;
  int fused_blocks_per_cu = 4;

  // incremental independent-rule loop state (loop.hip)
  void* inc = nullptr;
  // heuristic-based loop state (loop_ldf.hip)
  void* ldf = nullptr;

  GraphView view() const {
    return GraphView{n, n_owned, row, cap, col, prop, ccol, twin, mexc, hidx, hist,
                     nparts > 0 ? vpart : nullptr};
  }
};

// error plumbing
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* where);

#define DC_HIP(expr)                                       \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return ::dc::hip_fail(_e, #expr); \
  } while (0)

// csr.hip
int csr_build(Graph* g, int64_t m, const int32_t* du, const int32_t* dv);
int csr_apply(Graph* g, const Batch& b, int rule);
int csr_compact(Graph* g);
int csr_compact_noalloc(Graph* g);   // dc_compact_noalloc: static slot map
uint64_t compact_noalloc_max_slots();
int rebuild_ccol(Graph* g);               // ccol[j] <- colour[col[j]] for every slot (+ hub histograms)
int hist_reindex(Graph* g);               // hub rows from cap[], then rebuild the histograms
int parts_set(Graph* g, int32_t nparts, const int32_t* offsets);   // nparts 0 clears
int check_consistency(Graph* g, int64_t* bad);  // ccol/twin invariants (tests)
int alloc_state(Graph* g);
void free_state(Graph* g);
int exclusive_scan_u32_to_u64(Graph* g, const uint32_t* in, uint64_t* out, int32_t n);

// recolor.hip
// persistent_loop: run the independent-rule loop as one cooperative launch
int recolor_plain(Graph* g, int rule, dc_stats* st, bool persistent_loop = false);
// loop.hip: the independent-rule loop with incrementally maintained
// candidate counts; L0 = the marked set of the pre-loop rounds
struct LoopResult {
  int32_t rounds = 0;
  uint64_t evaluated = 0, changed = 0, frontier_sum = 0;
  uint64_t slots = 0, live = 0, rslots = 0, rlive = 0;
  double kernel_ms = 0.0;
  int64_t launches = 0;
  double count_ms = 0.0;        // summed count-pass launch durations (profiling on)
  double round_a_ms = 0.0;      // summed phase-A launch durations (profiling on)
  int64_t count_launches = 0;
  uint64_t count_bytes = 0;     // algorithmic bytes of those launches
  bool settled = false;         // stopped at settled colours (see loop_hi)
  int32_t pull_rounds = 0;      // rounds recounted by pull
  int32_t heavy_push_rounds = 0;  // pushed rounds whose D rows exceeded heavy_slots
  int32_t heavy_push_listed = 0;  // ... that also listed R
  int32_t graph_captures = 0;   // round blocks captured as HIP graphs (graph driver)
  uint64_t lc_vertices = 0;     // sum over rounds of |L_r u C_r| (profiled batches)
  uint64_t lc_slots = 0;        // sum over rounds of the row slots of L_r u C_r
  uint64_t moved_slots = 0;     // ldf: row slots of the vertices the loop changed
  bool slots_rewritten = false; // ldf: every row's slot colours rewritten after the loop
};
int loop_hi(Graph* g, Frontier L0, int32_t nl0, bool persistent, LoopResult* res);
void loop_free(Graph* g);
// loop_ldf.hip: the heuristic-based loop from the marked set L0 (both colour
// buffers written; no catch-up follows)
int loop_ldf(Graph* g, Frontier L0, LoopResult* res);
void ldf_free(Graph* g);
int loop_check(Graph* g, int64_t* bad);   // candidate counts / state arrays clean at rest
// fused.hip
int recolor_fused(Graph* g, int rule, dc_stats* st);
void fused_free(Graph* g);
int static_color(Graph* g, int rule);
int validate(Graph* g, int64_t* mono, int32_t* maxc);

inline uint32_t next_epoch(Graph* g) {
  // callers reset the mark arrays when the epoch would wrap
  return ++g->epoch;
}
int ensure_epochs(Graph* g, uint32_t needed);

}  // namespace dc
