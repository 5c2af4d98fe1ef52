// loop.hip -- the independent-rule recolor loop with candidate sets kept
// incrementally.
//
// The reference loop (HI/Color_main.cu:185-202) repeats
//   C = { w : some v in L has an edge (v, w) with
//             (c[v] == c[w] && v < w) || prev[v] < c[w] }   (findEligibleNeighbors,
//                                                             HI/gpuFunctions_undir.cuh:229-260)
//   for w in C: prev[w] = c[w]; c[w] = mex(w)                 (recolorNeighbor, :264-284)
//   L = C
// until C is empty.  On power-law graphs C keeps hundreds of rounds of hub
// rows while almost nothing changes colour: re-deriving C by scanning every
// row of L, and every mex by scanning every row of C, moves the same hub rows
// through HBM each round.
//
// This loop computes the same rounds from per-vertex counts instead:
//   K[w] = #{ slots (w, v) : v in L and the marking predicate holds for (v, w) }
// so C = { w : K[w] > 0 }.  Between rounds the only vertices whose terms move
// are D = (C xor L) plus the members of both whose (colour, prev) changed.
// For x in D the count pass adds the change of x's terms to every neighbour y
// outside D (one atomic per changed pair) and recounts K[x] from x's row.
// Candidates of the next C are the old C, D and the neighbours whose count
// rose (R).  A mex comes from the per-vertex cache (mexc), a hub's colour
// histogram, or a scan of a short row.  Colours, traces and round counts are
// those of the full re-derivation.
//
// Sets are state arrays, one nibble per vertex (4 MB for both parities at
// 4.2 M vertices): round r's C_r, D_r, R_r and X_r (members whose colour
// changed) are flags in St[r & 1], so L_r = C_{r-1} and the previous round's
// D and R are in the other array, and the count pass learns a neighbour's
// C, D, R and X membership from one gather.  A phase finds its vertices by
// scanning state words (each wave compacts the vertices of its 128 words,
// spread over the whole array so that the dense low-ID region of a
// power-law graph lands on many waves, into a list in LDS) or, in rounds
// whose sets are small, from worklists: a flag set on a vertex that had none
// in the round appends the vertex to the round's touched list, and phase A
// appends D's rows to three lists by count-pass path.  The next phase A then
// takes the touched list instead of scanning, the count pass takes D's lists
// instead of the state array and the row tables, and state arrays are
// cleared entry by entry.  A scan costs the same at any set size (2 MB of
// state words and a few hundred thousand table entries at 4.2 M vertices);
// a list costs what the set holds.  Appends are wave-aggregated atomics on
// one counter per list, and on this GPU a few thousand atomics on one address
// serialise into tens of microseconds, so rounds with large sets (more than
// dc_tuning.list_c vertices in L, or more than dc_tuning.list_slots slots in
// D's rows) scan.
//
// One round = two phases:
//   A(r): candidates = L_r | D_{r-1} | R_{r-1}.  K > 0 makes a member of C_r,
//         evaluated at once (cached mex; a short row scanned by its lane; a
//         mid row by the wave; a hub from its histogram) and committed;
//         members of L_r with K == 0 leave and join D_r.
//   C(r): rows of D_r: count changes, recounts, and for vertices whose colour
//         changed the slot cache / mex cache / histograms of their neighbours.
//         Rows of 64-511 slots go by the static histogram-row table, hub rows
//         by the static hub chunk table (or all three kinds by D's lists).
//         The previous round's state array is cleared.
// The plain driver launches one kernel per phase (the host enqueues rounds in
// blocks and reads one control word per block; rounds past the last one
// return at once; settle checks are queued between rounds); the persistent
// driver runs the rounds in one cooperative launch with grid barriers between
// phases.
//
// Settled colours end the loop early: after a round without a colour change
// in which every owned vertex holds the mex of its neighbours' colours, every
// later round recolours each member to the colour it has.  With disjoint
// parts (an arena of graphs), a part that passes the same test on its own
// leaves the loop while the others go on.
//
// Outside the loop K is zero everywhere and so are the state arrays.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "dc_phases.h"

namespace dc {

constexpr int kMaxBlocks = 4096;   // largest grid of the loop kernels (per-block partials)
constexpr int kAcc = 32;           // phase A sums: workgroup b adds into slot b % kAcc
constexpr int kAccW = 8;           // words per slot (five used)
// the loop's grids and thresholds come from dc_tuning (dc_tuning_default)
constexpr int kWaveList = 64 * 16; // vertices of 128 state words (16 per lane)
// Worklists: a round lists its sets when L has at most list_c vertices (the
// smaller of dc_tuning.list_c and 1/64 of the words that the scans pass:
// below that an append's atomic costs less than the scan words it saves), and
// R too when D's rows hold at most dc_tuning.list_slots slots (and the round
// is not a pull round).
constexpr int kGraphCache = 48;   // captured round blocks kept per graph (graph driver)
// rounds 0..kPullSplitRounds-1 of the plain and graph drivers launch
// k_pull_fix after their count pass, so that a pull recount among them can
// skip about half of its neighbour-state gathers
constexpr int32_t kPullSplitRounds = 4;
constexpr uint32_t kQInvalid = 0xffffffffu;  // touched count of a round that did not list
constexpr int kQT = 0;      // qc[kQT + (r & 3)]: entries of round r's touched list
constexpr int kQD = 4;      // qc[kQD + 3 * (r & 1) + k]: D_r's short rows, mid rows, hub chunks
constexpr int kQWords = 10;

struct LoopCtl {
  int32_t done_round;   // the round whose C was empty (INT_MAX while running)
  int32_t nl0;          // |L_0|
  int32_t last_nc;      // |C| of the latest round
  int32_t last_changed; // colour changes of the latest round
  // settle checks queued on the device (plain driver)
  int32_t checks_left;  // checks still allowed
  int32_t unsettled;    // a check found a vertex off its mex
  int32_t settled;      // the loop stopped at settled colours (done_round = last round run)
  int32_t newly_left;   // parts left the loop at the latest part check
  // members of C_r whose part left after round nl_sub_round (= r): round
  // r + 1's L is C_r without them
  int32_t nl_sub;
  int32_t nl_sub_round;
  // rounds by count-pass path: pull recounts, pushes whose D rows exceeded
  // heavy_slots, and those of them that listed R_r
  int32_t pull_rounds;
  int32_t heavy_push_rounds;
  int32_t heavy_push_listed;
  int32_t split_round;  // the latest round whose pull recount left work to k_pull_fix
};

// |L_r| (uniform: every workgroup reads the same words)
__device__ __forceinline__ int32_t loop_nl(const LoopCtl* c, const int4* trace, int r) {
  if (r == 0) return c->nl0;
  return trace[r - 1].y - (c->nl_sub_round == r - 1 ? c->nl_sub : 0);
}

struct LoopState {
  GraphView G;
  int32_t* col0;       // colour buffers: both are written, so they agree at the end
  int32_t* col1;
  int32_t* prev;
  int32_t* K;
  int2* dold;          // D members: {colour, (prev + 1) | in_L << 31} before the round
  // set membership of round r in St[r & 1], one nibble per vertex (8 per
  // word): kFC C_r, kFD D_r, kFR R_r, kFX members of C_r whose colour changed
  uint32_t* St[2];
  uint32_t swords;     // words per state array
  Frontier L0;         // the marked set
  const int2* hch;     // hub chunk table {vertex, chunk}
  uint32_t n_hch;
  const int32_t* hrowv;  // rows with a colour histogram (capacity >= kSmallLen)
  uint32_t n_hrows;
  unsigned long long heavy_slots;  // D_r rows beyond this many slots: pull recount
  int32_t pull_mode;   // 0: by heavy_slots; 1: every round pulls; -1: no round pulls
  // rounds below this one split a pull recount with k_pull_fix (launched
  // after their count pass): the pull skips the slots whose neighbour has a
  // higher cached colour, k_pull_fix adds the changed members' terms on them
  // and writes the new colours into the neighbours' caches (0: no split)
  int32_t pull_split_rounds;
  // per-batch work terms of the reference's algorithm (null: not counted):
  // {sum over rounds of |L_r u C_r|, of the row slots of L_r u C_r}
  unsigned long long* terms;
  // phase A's totals by round parity, spread over kAcc slots so that no
  // address takes more than gridDim / kAcc atomics: {members, changes, slots
  // of the rows put into D, hub rows put into D, histogram rows below hub
  // length put into D}
  unsigned long long* acc;    // [2][kAcc][kAccW]
  // per workgroup: {stale-scan slots (A), pull slots (C), count slots (C),
  // row and chunk visits (C)}
  unsigned long long* bstat;  // [kMaxBlocks][4]
  int4* trace;         // per round {|L|, |C|, changed, 0}
  LoopCtl* ctl;
  // worklists: round r's vertices with a flag in St[r & 1] in tq[r & 1]
  // (count qc[kQT + (r & 3)], kQInvalid when the round did not list them
  // all); D_r's rows without a histogram, histogram rows below kMidLen and
  // hub chunks {vertex, chunk} in dq_s, dq_m, dq_h
  int32_t* tq[2];
  uint32_t tq_cap;
  int32_t* dq_s;
  int32_t* dq_m;
  int2* dq_h;
  uint32_t dq_s_cap, dq_m_cap, dq_h_cap;
  uint32_t* qc;               // [kQWords]
  // disjoint parts (dc_set_parts): vertex -> part; per part whether a colour
  // changed, by round parity [2][nparts], a failed part check [nparts], and
  // whether it left the loop [nparts] (2: at the latest check, 1: before)
  const int32_t* vpart;
  uint32_t* pchg;
  uint32_t* punset;
  uint32_t* pret;
  int32_t nparts;
  int32_t list_c;             // rounds with |L| up to this list their sets
  unsigned long long list_slots;  // ... and R too when D's rows hold at most this many slots
  unsigned long long* wstat;  // 4-byte words of state arrays and lists passed by count passes
};

// LDS of a loop workgroup, shared by both phases
// (20 KiB, so that eight workgroups fit in a CU's LDS.)  The mex windows of
// phase A share their space with the workgroup sums, which are written only
// after a barrier that follows every window use; the table-slice list of
// the count pass shares the first wave list, which it writes only after a
// barrier that follows every list use.
struct LoopLds {
  int32_t pick[kBlock / 64][64];    // a wave's compacted entries of one batch of 64
  union {
    int32_t lists[kBlock / 64][kWaveList];
    int2 chunks[kBlock];            // active table entries of one slice
  };
  union {
    unsigned long long wins[kBlock / 64][kWinWords];
    struct {
      uint32_t red[kBlock / 64];
      uint32_t red4[kBlock / 64][4];
      unsigned long long red64[kBlock / 64];
      unsigned long long tot[5];
      uint32_t nchunks[kBlock / 64];
    };
  };
};
static_assert(sizeof(LoopLds) <= 24576, "loop LDS exceeds 3/20 of a CU");

// A captured block of rounds (graph driver) and the launch state it was
// captured with.
struct GraphBlock {
  int32_t r0 = 0, nb = 0, grid = 0;
  std::vector<char> state;
  hipGraphExec_t exec = nullptr;
};

struct IncState {
  LoopState S{};
  std::vector<hipEvent_t> cev;   // profiling: start and stop of each phase-A and count-pass launch
  uint32_t* bits = nullptr;   // St[0], St[1]
  int32_t trace_cap = 0;
  LoopCtl* hctl = nullptr;
  unsigned long long* hstat = nullptr;
  int32_t* tqbuf = nullptr;     // tq[0], tq[1]
  unsigned long long* terms = nullptr;   // LoopState::terms (device, 4 words)
  std::vector<GraphBlock> graphs;        // graph driver: captured round blocks
  int64_t captures = 0;
};

// Membership flags of one vertex in a state array (one gather tells a
// neighbour's C, D, R and X membership at once).
constexpr uint32_t kFC = 1u, kFD = 2u, kFR = 4u, kFX = 8u;

__device__ __forceinline__ uint32_t st_flags(uint32_t word, int32_t v) {
  return (word >> ((v & 7) << 2)) & 15u;
}
__device__ __forceinline__ bool st_has(const uint32_t* st, int32_t v, uint32_t f) {
  return (st[v >> 3] >> ((v & 7) << 2)) & f;
}
__device__ __forceinline__ void st_set(uint32_t* st, int32_t v, uint32_t f) {
  atomicOr(&st[v >> 3], f << ((v & 7) << 2));
}
// st_set that tells whether v had no flag before
__device__ __forceinline__ bool st_set_first(uint32_t* st, int32_t v, uint32_t f) {
  const uint32_t sh = ((uint32_t)v & 7u) << 2;
  return ((atomicOr(&st[(uint32_t)v >> 3], f << sh) >> sh) & 15u) == 0u;
}

// The marking predicate of findEligibleNeighbors for a slot (v -> w).
__device__ __forceinline__ bool marks(int32_t v, int32_t cv, int32_t pv, int32_t w, int32_t cw) {
  return (cv == cw && v < w) || pv < cw;
}

// Walk the vertices of a state array with any of the flags f set.  Each
// wave takes 128 words (1024 vertices) per step, two per lane strided over
// the array, compacts the vertices into its LDS list and calls fn(list,
// count) with the whole wave.
template <class Fn>
__device__ __forceinline__ void walk_states(const uint32_t* st, uint32_t swords, uint32_t f,
                                            int32_t* list, Fn fn) {
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t fm = f * 0x11111111u;
  auto nibbles = [](uint32_t x) { return (x | (x >> 1) | (x >> 2) | (x >> 3)) & 0x11111111u; };
  for (uint64_t base = 0; base < swords; base += 128ull * nwaves) {
    const uint64_t w0 = base + (uint64_t)lane * nwaves + gwave;
    const uint64_t w1 = w0 + 64ull * nwaves;
    uint32_t b0 = w0 < swords ? nibbles(st[w0] & fm) : 0u;
    uint32_t b1 = w1 < swords ? nibbles(st[w1] & fm) : 0u;
    const uint32_t cnt = (uint32_t)(__popc(b0) + __popc(b1));
    uint32_t x = cnt;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    const uint32_t total = __shfl(x, 63);
    if (total == 0) continue;
    uint32_t pos = x - cnt;
    while (b0) {
      const int b = __ffs(b0) - 1;
      b0 &= b0 - 1u;
      list[pos++] = (int32_t)(w0 * 8 + (b >> 2));
    }
    while (b1) {
      const int b = __ffs(b1) - 1;
      b1 &= b1 - 1u;
      list[pos++] = (int32_t)(w1 * 8 + (b >> 2));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    fn(list, total);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// The entries of a worklist by the same contract: each wave takes 64 entries
// at a time (waves strided over the list), copies them into its LDS list and
// calls fn(list, count) with the whole wave.
template <class Fn>
__device__ __forceinline__ void walk_list(const int32_t* q, uint32_t n, int32_t* list, Fn fn) {
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = (uint64_t)gwave * 64; base < n; base += 64ull * nwaves) {
    const uint64_t i = base + lane;
    if (i < n) list[lane] = q[i];
    const uint32_t total = (uint32_t)min<uint64_t>(64ull, (uint64_t)n - base);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    fn(list, total);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// Append the vertices v[k] with p[k] of the executing lanes to a list of cap
// entries with one atomic for all U of them (entries past cap are counted,
// not written).
template <int U>
__device__ __forceinline__ void wave_append_many(const bool (&p)[U], const int32_t (&v)[U],
                                                 int32_t* list, uint32_t* counter, uint32_t cap) {
  unsigned long long m[U];
  uint32_t tot = 0;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    m[k] = __ballot(p[k]);
    tot += (uint32_t)__popcll(m[k]);
  }
  if (tot == 0u) return;
  const int leader = __ffsll((long long)__ballot(true)) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, tot);
  base = __shfl(base, leader);
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint32_t pos = base + lanes_below(m[k]);
    if (p[k] && pos < cap) list[pos] = v[k];
    base += (uint32_t)__popcll(m[k]);
  }
}

// A Lister appends vertices to a worklist (off when q is null).
struct Lister {
  int32_t* q;
  uint32_t* cnt;
  uint32_t cap;
};

// Sum over the workgroup (every thread gets the total).
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* scratch) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int wib = threadIdx.x >> 6;
  __syncthreads();
  if (lane_id() == 0) scratch[wib] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += scratch[i];
  __syncthreads();
  return t;
}

// A per-lane value from lane src, word by word (a shuffle every lane takes
// part in).
template <class P>
__device__ __forceinline__ P bcast(const P& p, int src) {
  static_assert(sizeof(P) % 4 == 0, "shuffled in 4-byte words");
  union W {
    P p;
    int a[sizeof(P) / 4];
  } u, o;
  u.p = p;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(P) / 4); ++i) o.a[i] = __shfl(u.a[i], src);
  return o.p;
}

// Hub chunks of a chunk table (the static one or a list) whose entry passes
// act(e): the workgroup takes the entries
// blockIdx, blockIdx + gridDim, ... (consecutive chunks of one hub go to
// different workgroups), compacts the active ones in LDS, and runs each chunk
// with all its waves: wave q of the workgroup calls fn(x, row, lo, hi, p) on
// the q-th quarter of the chunk's slots, p = pre(x).  The rows and pre() of
// up to 64 chunks are loaded at once, one per lane.  Every thread must call
// this (it synchronises the workgroup).
template <class Act, class Pre, class Fn>
__device__ __forceinline__ void each_hub_chunk(const LoopState& S, LoopLds& lds, const int2* tab,
                                               uint32_t ntab, Act act, Pre pre, Fn fn) {
  const GraphView& G = S.G;
  const int lane = lane_id();
  const int wib = threadIdx.x >> 6;
  constexpr uint32_t kQuarter = kChunk / (kBlock / 64);
  for (uint64_t base = 0; base < ntab; base += (uint64_t)kBlock * gridDim.x) {
    const uint64_t i = base + (uint64_t)threadIdx.x * gridDim.x + blockIdx.x;
    const int2 e = i < ntab ? tab[i] : make_int2(0, -1);
    const bool a = e.y >= 0 && act(e);
    const unsigned long long am = __ballot(a);
    if (lane == 0) lds.nchunks[wib] = (uint32_t)__popcll(am);
    __syncthreads();
    uint32_t off = 0, n = 0;
    for (int q = 0; q < kBlock / 64; ++q) {
      const uint32_t c = lds.nchunks[q];
      off += q < wib ? c : 0u;
      n += c;
    }
    if (a) lds.chunks[off + lanes_below(am)] = e;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
      const uint32_t m = min(64u, n - c0);
      const int2 ce = (uint32_t)lane < m ? lds.chunks[c0 + lane] : make_int2(0, 0);
      const uint2 rw = (uint32_t)lane < m ? G.row[ce.x] : make_uint2(0u, 0u);
      const auto p = pre(ce.x);
      for (uint32_t q = 0; q < m; ++q) {
        const int32_t x = __shfl(ce.x, (int)q);
        const uint32_t c = (uint32_t)__shfl(ce.y, (int)q);
        const uint2 rq = bcast(rw, (int)q);
        const auto pq = bcast(p, (int)q);
        const uint32_t lo = c * kChunk + (uint32_t)wib * kQuarter;
        // rows now shorter than kMidLen went by the short/mid path
        if (rq.y < kMidLen || lo >= rq.y) continue;
        fn(x, rq, lo, min(rq.y, lo + kQuarter), pq);
      }
    }
    __syncthreads();
  }
}

// Rows of a histogram-row table (the static one of every row of kSmallLen or
// more slots of capacity, or a list) that pass act(v) and are shorter than
// kMidLen (longer ones go by
// the hub chunk table): the workgroup takes entries blockIdx, blockIdx +
// gridDim, ..., compacts the active ones in LDS, and wave w takes entries w,
// w + 4, ...: its lanes load those rows and pre(x) at once, then the wave
// runs them in turn, fn(x, row, p).  The table spreads these rows evenly over
// the workgroups whatever their vertex IDs.  Every thread must call this (it
// synchronises the workgroup).
template <class Act, class Pre, class Fn>
__device__ __forceinline__ void each_mid_row(const LoopState& S, LoopLds& lds, const int32_t* tab,
                                             uint32_t ntab, Act act, Pre pre, Fn fn) {
  const GraphView& G = S.G;
  const int lane = lane_id();
  const int wib = threadIdx.x >> 6;
  constexpr uint32_t kWaves = kBlock / 64;
  for (uint64_t base = 0; base < ntab; base += (uint64_t)kBlock * gridDim.x) {
    const uint64_t i = base + (uint64_t)threadIdx.x * gridDim.x + blockIdx.x;
    const int32_t v = i < ntab ? tab[i] : -1;
    const bool a = v >= 0 && act(v);
    const unsigned long long am = __ballot(a);
    if (lane == 0) lds.nchunks[wib] = (uint32_t)__popcll(am);
    __syncthreads();
    uint32_t off = 0, n = 0;
    for (int q = 0; q < kBlock / 64; ++q) {
      const uint32_t c = lds.nchunks[q];
      off += q < wib ? c : 0u;
      n += c;
    }
    if (a) lds.chunks[off + lanes_below(am)] = make_int2(v, 0);
    __syncthreads();
    const uint32_t mine = n > (uint32_t)wib ? (n - (uint32_t)wib + kWaves - 1) / kWaves : 0u;
    const bool has = (uint32_t)lane < mine;
    const int32_t x = has ? lds.chunks[(uint32_t)wib + kWaves * (uint32_t)lane].x : 0;
    const uint2 rw = has ? G.row[x] : make_uint2(0u, 0u);
    const auto p = pre(x);
    for (uint32_t q = 0; q < mine; ++q) {
      const uint2 rq = bcast(rw, (int)q);
      if (rq.y >= kMidLen) continue;
      fn(__shfl(x, (int)q), rq, bcast(p, (int)q));
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// set-up: L0's bits, every marking slot of L0's rows counted into K, marked
// vertices become candidates (R_{-1}).
__global__ __launch_bounds__(kBlock) void k_loop_init(LoopState S) {
  const GraphView& G = S.G;
  const Frontier F = S.L0;
  const int lane = lane_id();
  // list counts: round 0's start at zero; round -1 (the marked set) has no
  // touched list
  if (blockIdx.x == 0 && threadIdx.x < kQWords)
    S.qc[threadIdx.x] = threadIdx.x == kQT + 3 ? kQInvalid : 0u;
  // part state (changes of both parities, checks, left) starts clear
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < 4 * S.nparts; i += blockDim.x) S.pchg[i] = 0u;
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  auto row = [&](int32_t v, uint2 r, uint32_t lo, uint32_t hi, uint32_t stride, bool head) {
    if (head) st_set(S.St[1], v, kFC);
    const int32_t cv = S.col0[v];
    const int32_t pv = S.prev[v];
    for (uint32_t j = lo; j < hi; j += stride) {
      const int32_t w = G.col[(uint64_t)r.x + j];
      const int32_t cw = G.ccol[(uint64_t)r.x + j];
      if (w >= 0 && w < G.n_owned && marks(v, cv, pv, w, cw)) {
        atomicAdd(&S.K[w], 1);
        st_set(S.St[1], w, kFR);
      }
    }
  };
  for (uint32_t ci = blockIdx.x; ci < F.cnt[2]; ci += gridDim.x) {
    const int4 it = F.h[ci];
    const uint2 r = G.row[it.x];
    const uint32_t lo = (uint32_t)it.y * kChunk;
    row(it.x, r, lo + threadIdx.x, min(r.y, lo + kChunk), kBlock, it.y == 0 && threadIdx.x == 0);
  }
  for (int64_t i = gwave; i < F.cnt[1]; i += nwaves) {
    const int32_t v = F.m[i];
    const uint2 r = G.row[v];
    row(v, r, lane, r.y, 64, lane == 0);
  }
  for (int64_t i = gwave * 64 + lane; i < F.cnt[0]; i += nwaves * 64) {
    const int32_t v = F.s[i];
    const uint2 r = G.row[v];
    row(v, r, 0, r.y, 1, true);
  }
}

// ---------------------------------------------------------------------------
// phase A

// Commit w's new colour m (one lane); w joins D when its (in_L, colour,
// prev) state changed.
__device__ __forceinline__ bool commit_member(const LoopState& S, uint32_t* sn, int32_t w,
                                              int32_t cm, int32_t so, int32_t po, bool inl,
                                              int32_t m, uint32_t& changed, uint32_t* pch) {
  if (cm < 0) S.G.mexc[w] = m;
  // both colour buffers hold so and prev holds po: a member that keeps its
  // colour (and whose prev already is its colour) stores nothing, so the
  // many stable members of a heavy round dirty no lines
  const bool chv = m != so;
  if (chv) {
    S.col0[w] = m;
    S.col1[w] = m;
  }
  if (po != so) S.prev[w] = so;
  changed += chv ? 1u : 0u;
  // the part changed this round (a flag: a counter atomic per change would
  // serialise on a large part's one address)
  if (chv && pch) {
    const int32_t p = S.vpart[w];
    if (pch[p] == 0u) pch[p] = 1u;
  }
  const bool dv = !inl || po != so || chv;
  if (dv) {
    S.dold[w] = make_int2(so, (po + 1) | (inl ? (int32_t)0x80000000 : 0));
    S.K[w] = 0;
  }
  st_set(sn, w, kFC | (chv ? kFX : 0u) | (dv ? kFD : 0u));
  return dv;
}

// mex of a row with fewer than 64 slots by an 8-lane group (lane `sub` reads
// slots sub, sub + 8, ...: the whole row in one round trip)
__device__ __forceinline__ int32_t group_small_mex(const GraphView& G, uint2 r, int sub) {
  int32_t c[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t j = (uint32_t)sub + 8u * k;
    c[k] = j < r.y ? G.ccol[(uint64_t)r.x + j] : -1;
  }
  unsigned long long mask = 0ull;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (c[k] >= 0 && c[k] < 64) mask |= 1ull << c[k];
  mask |= __shfl_xor(mask, 1);
  mask |= __shfl_xor(mask, 2);
  mask |= __shfl_xor(mask, 4);
  return __ffsll((long long)~mask) - 1;
}

// mex of a row with fewer than kSmallLen slots by one lane: its slots 16 at a
// time, all loads in flight, colours 0..63 kept as two 32-bit words (the mex
// of fewer than 64 slots is below 64)
__device__ __forceinline__ int32_t lane_small_mex(const GraphView& G, uint2 r) {
  uint32_t lo = 0u, hi = 0u;
  for (uint32_t j = 0; j < r.y; j += 16) {
    int32_t c[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) c[k] = j + k < r.y ? G.ccol[(uint64_t)r.x + j + k] : -1;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int32_t x = c[k];
      if (x >= 0 && x < 32) lo |= 1u << x;
      if (x >= 32 && x < 64) hi |= 1u << (x - 32);
    }
  }
  if (lo != 0xffffffffu) return __ffs(~lo) - 1;
  return hi != 0xffffffffu ? 32 + __ffs(~hi) - 1 : 64;
}

// mex from a colour histogram by an 8-lane group: 256 counts per step (lane
// `sub` reads 32 of them); -1 when colours 0..kHistW-1 are all present
__device__ __forceinline__ int32_t group_hist_mex(const uint32_t* hrow, int sub) {
  const uint4* p = reinterpret_cast<const uint4*>(hrow);
  for (int step = 0; step < kHistW / 256; ++step) {
    uint4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p[step * 64 + sub * 8 + k];
    int first = INT_MAX;
#pragma unroll
    for (int k = 7; k >= 0; --k) {
      const int c = step * 256 + sub * 32 + 4 * k;
      if (v[k].w == 0u) first = c + 3;
      if (v[k].z == 0u) first = c + 2;
      if (v[k].y == 0u) first = c + 1;
      if (v[k].x == 0u) first = c;
    }
    first = min(first, __shfl_xor(first, 1));
    first = min(first, __shfl_xor(first, 2));
    first = min(first, __shfl_xor(first, 4));
    if (first != INT_MAX) return first;
  }
  return -1;
}

__device__ __forceinline__ void round_a(const LoopState& S, int r, LoopLds& lds) {
  const GraphView& G = S.G;
  const uint32_t* sp = S.St[(r + 1) & 1];   // L_r = C_{r-1}, D_{r-1}, R_{r-1}
  uint32_t* sn = S.St[r & 1];               // C_r, D_r, X_r
  const int lane = lane_id();
  const int wib = threadIdx.x >> 6;
  int32_t* list = lds.lists[wib];
  unsigned long long* win = lds.wins[wib];
  RecolorArgs RA{};
  RA.G = G;
  uint32_t members = 0, changed = 0, dhubs = 0, dmids = 0;
  unsigned long long slots = 0, dsum = 0;
  // rounds with a small L list their sets (uniform: the previous round's |C|)
  const int32_t nl = loop_nl(S.ctl, S.trace, r);
  const bool listing = nl <= S.list_c;
  unsigned long long lc_n = 0, lc_deg = 0;   // terms: |L_r u C_r| and its row slots
  uint32_t* qd = S.qc + kQD + 3 * (r & 1);
  // colour changes of this round by part (parts set)
  uint32_t* pch = S.nparts ? S.pchg + (size_t)(r & 1) * S.nparts : nullptr;
  // a row put into D: its slots, and which count-pass path takes it
  auto d_row = [&](int32_t w, uint32_t len, int32_t h) {
    const bool hub = len >= kMidLen;
    const bool mid = !hub && h >= 0;
    dsum += len;
    dhubs += hub ? 1u : 0u;
    dmids += mid ? 1u : 0u;
    if (listing) {
      wave_append_capped(!hub && !mid, w, S.dq_s, qd + 0, S.dq_s_cap);
      wave_append_capped(mid, w, S.dq_m, qd + 1, S.dq_m_cap);
      if (hub) {
        const uint32_t nch = (len + kChunk - 1) / kChunk;
        const uint32_t b = atomicAdd(qd + 2, nch);
        for (uint32_t c = 0; c < nch && b + c < S.dq_h_cap; ++c)
          S.dq_h[b + c] = make_int2(w, (int32_t)c);
      }
    }
  };

  auto visit = [&](int32_t* lst, uint32_t total) {
    // members with a cached mex commit at once, and so do those with a stale
    // mex and a row without a histogram (each lane scans its own row); the
    // others (stale mex, histogram rows) are compacted to the front of the
    // list and evaluated by 8-lane groups
    uint32_t nd = 0;
    for (uint32_t base = 0; base < total; base += 64) {
      const uint32_t i = base + lane;
      const bool valid = i < total;
      const int32_t w = valid ? lst[i] : 0;
      int32_t k = 0, cm = 0, so = 0, po = 0, h = -1;
      uint2 rw = make_uint2(0u, 0u);
      bool inl = false;
      if (valid) {
        k = S.K[w];
        cm = G.mexc[w];
        so = S.col0[w];
        po = S.prev[w];
        inl = st_has(sp, w, kFC);
        h = G.hidx[w];
        rw = G.row[w];
      }
      const bool member = valid && k > 0;
      const bool leaving = valid && !member && inl;
      if (S.terms && (member || inl)) {
        lc_n += 1;
        lc_deg += rw.y;
      }
      // w gets its first flag of the round (C or D) in this phase
      if (listing)
        wave_append_capped(member || leaving, w, S.tq[r & 1], S.qc + kQT + (r & 3), S.tq_cap);
      bool dv = leaving;
      if (leaving) {
        S.dold[w] = make_int2(so, (po + 1) | (int32_t)0x80000000);
        st_set(sn, w, kFD);
      }
      int32_t mm = cm;
      if (member && cm < 0 && h < 0) {
        mm = lane_small_mex(G, rw);
        slots += rw.y;
      }
      if (member && mm >= 0) dv = commit_member(S, sn, w, cm, so, po, inl, mm, changed, pch);
      if (dv) d_row(w, rw.y, h);
      members += member ? 1u : 0u;
      const bool defer = member && mm < 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const unsigned long long dm = __ballot(defer);
      if (defer) lst[nd + lanes_below(dm)] = w;
      nd += (uint32_t)__popcll(dm);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    // stale mex from the row's colour histogram (a short row's slots if one
    // is here); a mex past the histogram goes to the front of the list again
    // (nw)
    const int grp = lane >> 3, sub = lane & 7;
    uint32_t nw = 0;
    for (uint32_t b8 = 0; b8 < nd; b8 += 8) {
      const uint32_t q = b8 + grp;
      const bool valid = q < nd;
      const int32_t w = valid ? lst[q] : 0;
      int32_t m = -1;
      if (valid) {
        const int32_t h = G.hidx[w];
        if (h >= 0) {
          m = group_hist_mex(G.hist + (uint64_t)h * kHistW, sub);
        } else {
          const uint2 rw = G.row[w];
          m = group_small_mex(G, rw, sub);
          if (sub == 0) slots += rw.y;
        }
      }
      const bool wst = valid && m < 0;
      if (valid && !wst && sub == 0 &&
          commit_member(S, sn, w, -1, S.col0[w], S.prev[w], st_has(sp, w, kFC), m, changed, pch))
        d_row(w, G.row[w].y, G.hidx[w]);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const unsigned long long wm = __ballot(wst && sub == 0);
      if (wst && sub == 0) lst[nw + lanes_below(wm)] = w;
      nw += (uint32_t)__popcll(wm);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    // mex past the histogram: one at a time by the whole wave (windowed scan)
    for (uint32_t q = 0; q < nw; ++q) {
      const int32_t w = lst[q];
      const int32_t m = wave_row_mex<M_LOOP_HI>(RA, w, win);
      if (lane == 0 &&
          commit_member(S, sn, w, -1, S.col0[w], S.prev[w], st_has(sp, w, kFC), m, changed, pch))
        d_row(w, G.row[w].y, G.hidx[w]);
    }
  };
  // the previous round's touched list when it has one, else the state array
  const uint32_t tprev = S.qc[kQT + ((r + 3) & 3)];
  if (tprev <= S.tq_cap)
    walk_list(S.tq[(r + 1) & 1], tprev, list, visit);
  else
    walk_states(sp, S.swords, kFC | kFD | kFR, list, visit);

  // the workgroup's sums in one reduction (after the barrier every wave is
  // done with its mex window, which shares LDS with red4 and red64)
  for (int off = 32; off > 0; off >>= 1) {
    members += __shfl_xor(members, off);
    changed += __shfl_xor(changed, off);
    dhubs += __shfl_xor(dhubs, off);
    dmids += __shfl_xor(dmids, off);
    dsum += __shfl_xor(dsum, off);
    slots += __shfl_xor(slots, off);
  }
  __syncthreads();
  if (lane == 0) {
    lds.red4[wib][0] = members;
    lds.red4[wib][1] = changed;
    lds.red4[wib][2] = dhubs;
    lds.red4[wib][3] = dmids;
    lds.red64[wib] = dsum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t bm = 0, bc = 0, bh = 0, bmid = 0;
    unsigned long long bd = 0;
    for (int q = 0; q < kBlock / 64; ++q) {
      bm += lds.red4[q][0];
      bc += lds.red4[q][1];
      bh += lds.red4[q][2];
      bmid += lds.red4[q][3];
      bd += lds.red64[q];
    }
    unsigned long long* a = S.acc + ((size_t)(r & 1) * kAcc + blockIdx.x % kAcc) * kAccW;
    if (bm) atomicAdd(&a[0], (unsigned long long)bm);
    if (bc) atomicAdd(&a[1], (unsigned long long)bc);
    if (bd) atomicAdd(&a[2], bd);
    if (bh) atomicAdd(&a[3], (unsigned long long)bh);
    if (bmid) atomicAdd(&a[4], (unsigned long long)bmid);
  }
  if (lane == 0 && slots) atomicAdd(&S.bstat[blockIdx.x * 4 + 0], slots);
  if (S.terms) {
    wave_add_u64(lc_n, &S.terms[0]);
    wave_add_u64(lc_deg, &S.terms[1]);
  }
}

// ---------------------------------------------------------------------------
// phase C.  For x in D_r and each slot (x, y):
//   y outside D_r: K[y] += (x's new term) - (x's old term); a rise makes y a
//                  candidate (R_r);
//   y in C_r (= L_{r+1}): y's term towards x is part of x's recount;
// and if x changed colour, y's slot cache, cached mex and histogram follow.
// A slot of a neighbour that changed colour is never read here: such
// neighbours are in D_r.
struct DState {
  int32_t x, so, po, sn, pn;
  bool inl, inc, chg, pad;
};

__device__ __forceinline__ DState load_d(const LoopState& S, const uint32_t* st, int32_t x) {
  DState d;
  d.x = x;
  const int2 o = S.dold[x];
  d.so = o.x;
  d.po = (o.y & 0x7fffffff) - 1;
  d.inl = o.y < 0;
  d.sn = S.col0[x];
  d.pn = S.prev[x];
  d.inc = st_has(st, x, kFC);
  d.chg = d.sn != d.so;
  d.pad = false;
  return d;
}

// The count and pull passes stream col and ccol with non-temporal loads, so
// that a heavy round's gigabyte of row slots does not evict the state arrays
// that every slot's neighbour lookup gathers from (measured ~1 % per batch on
// R-MAT 22).
//
// One lane's share of a count-pass visit of x's row: slots lo, lo+stride, ...
// Returns this lane's part of x's recount.  U slots at a time in four
// stages (slots; their neighbours' state words; count changes and the state
// of neighbours that changed colour; the recount), each stage's gathers all
// in flight before the next stage reads them.  x's own colour change goes
// to the neighbours separately (propagate_slots).
template <int U>
__device__ __forceinline__ uint32_t count_slots(const LoopState& S, const DState& d, uint64_t rbase,
                                                uint32_t lo, uint32_t hi, uint32_t stride,
                                                uint32_t* st, const Lister& tl,
                                                unsigned long long& slots) {
  const GraphView& G = S.G;
  uint32_t cnt = 0;
  for (uint32_t j = lo; j < hi; j += U * stride) {
    int32_t y[U], c[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t jj = j + k * stride;
      const bool in = jj < hi;
      y[k] = in ? __builtin_nontemporal_load(&G.col[rbase + jj]) : -1;
      c[k] = in ? __builtin_nontemporal_load(&G.ccol[rbase + jj]) : -1;
      slots += in ? 1u : 0u;
    }
    uint32_t fl[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      // tombstones and frozen ghosts read word 0 and are masked out below
      const bool own = y[k] >= 0 && y[k] < G.n_owned;
      fl[k] = own ? st_flags(st[(uint32_t)y[k] >> 3], y[k]) : 0u;
    }
    int32_t gc[U], gp[U];
    bool mem[U], fresh[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool own = y[k] >= 0 && y[k] < G.n_owned;
      fresh[k] = false;
      if (own && !(fl[k] & kFD)) {
        const bool to = d.inl && marks(d.x, d.so, d.po, y[k], c[k]);
        const bool tn = d.inc && marks(d.x, d.sn, d.pn, y[k], c[k]);
        if (to != tn) {
          atomicAdd(&S.K[y[k]], tn ? 1 : -1);
          // R_r only gains members in this phase: a flag seen set stays set.
          // A member of C_r is a candidate of the next round already and
          // takes no R flag (on power-law graphs thousands of rising counts
          // hit the same hubs, which are members).  A vertex that had no
          // flag goes on the touched list (one that had a flag when gathered
          // has one now: no returning atomic).
          if (tn && !(fl[k] & (kFR | kFC))) {
            if (tl.q && fl[k] == 0u)
              fresh[k] = st_set_first(st, y[k], kFR);
            else
              st_set(st, y[k], kFR);
          }
        }
      }
      mem[k] = own && (fl[k] & kFC);
      // a member that kept its colour has prev == colour == the slot value;
      // one that changed is gathered (its slots may not be updated yet)
      gc[k] = c[k];
      gp[k] = c[k];
      if (mem[k] && (fl[k] & kFX)) {
        gc[k] = S.col0[y[k]];
        gp[k] = S.prev[y[k]];
      }
    }
    if (tl.q) wave_append_many<U>(fresh, y, tl.q, tl.cnt, tl.cap);
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (mem[k] && marks(y[k], gc[k], gp[k], d.x, d.sn)) ++cnt;
  }
  return cnt;
}

__device__ __forceinline__ void clear_words(uint32_t* b, uint32_t words) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x)
    b[i] = 0u;
}

// Clear a state array: the words of a round's touched list when it has one
// (n <= cap), else every word.  Returns the 4-byte words passed.
__device__ __forceinline__ uint64_t clear_state(uint32_t* b, uint32_t words, const int32_t* q,
                                                uint32_t n, uint32_t cap) {
  if (n > cap) {
    clear_words(b, words);
    return words;
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[(uint32_t)q[i] >> 3] = 0u;
  return 2ull * n;
}

// ---- heavy rounds: every K recounted by pull, no count atomics ------------
// y's count over slots lo, lo+stride, ... of its row: members x of C_r whose
// marking predicate holds for (x, y).
// With split, a slot whose cached colour c is above y's colour is skipped
// (no state gather): while the slot caches still hold every member's colour
// from before the round (k_pull_fix writes the new ones after the pull), c is
// a member's prev and its term (c' == cy && x < y) || c < cy can hold only
// for a member that changed colour to c' == cy, which k_pull_fix adds.  About
// half of all slots have the higher colour at the other end.
template <int U>
__device__ __forceinline__ uint32_t pull_slots(const LoopState& S, int32_t y, int32_t cy,
                                               uint64_t rbase, uint32_t lo, uint32_t hi,
                                               uint32_t stride, const uint32_t* st, bool split,
                                               unsigned long long& slots) {
  const GraphView& G = S.G;
  uint32_t cnt = 0;
  for (uint32_t j = lo; j < hi; j += U * stride) {
    int32_t x[U], c[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t jj = j + k * stride;
      const bool in = jj < hi;
      x[k] = in ? __builtin_nontemporal_load(&G.col[rbase + jj]) : -1;
      c[k] = in ? __builtin_nontemporal_load(&G.ccol[rbase + jj]) : -1;
      slots += in ? 1u : 0u;
    }
    uint32_t fl[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool own = x[k] >= 0 && x[k] < G.n_owned && !(split && c[k] > cy);
      fl[k] = own ? st_flags(st[(uint32_t)x[k] >> 3], x[k]) : 0u;
    }
    int32_t gc[U], gp[U];
    bool mem[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      mem[k] = (fl[k] & kFC) != 0u;
      gc[k] = c[k];
      gp[k] = c[k];
      if (mem[k] && (fl[k] & kFX)) {
        gc[k] = S.col0[x[k]];
        gp[k] = S.prev[x[k]];
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (mem[k] && marks(x[k], gc[k], gp[k], y, cy)) ++cnt;
  }
  return cnt;
}

__device__ __forceinline__ void set_count(const LoopState& S, int32_t y, uint32_t cnt,
                                          uint32_t* st) {
  S.K[y] = (int32_t)cnt;
  if (cnt && !st_has(st, y, kFC)) st_set(st, y, kFR);
}

__device__ __forceinline__ void pull_all(const LoopState& S, uint32_t* st, LoopLds& lds,
                                         bool split, unsigned long long& slots,
                                         unsigned long long& rows) {
  const GraphView& G = S.G;
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  // rows without a histogram (fewer than kSmallLen slots) by 8-lane groups:
  // a wave loads 64 consecutive vertices' rows at once (lane l vertex l),
  // compacts the non-empty short ones and takes them 8 at a time.  An empty
  // row's count is 0 and stays 0.
  const int grp = lane >> 3, sub = lane & 7;
  int32_t* pick = lds.pick[threadIdx.x >> 6];
  const uint64_t nv = (uint64_t)G.n_owned;
  for (uint64_t base = (uint64_t)gwave * 64; base < nv; base += 64ull * nwaves) {
    const uint64_t vi = base + lane;
    const bool valid = vi < nv;
    const int32_t y = valid ? (int32_t)vi : 0;
    const int32_t h = valid ? G.hidx[y] : 0;
    const uint2 rw = valid ? G.row[y] : make_uint2(0u, 0u);
    const int32_t cy = S.col0[y];
    const bool small = valid && h < 0 && rw.y > 0;
    const unsigned long long sm = __ballot(small);
    if (small) pick[lanes_below(sm)] = lane;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t ns = (uint32_t)__popcll(sm);
    for (uint32_t t = 0; t < ns; t += 8) {
      const bool mine = t + grp < ns;
      const int src = mine ? pick[t + grp] : lane;
      const int32_t yy = __shfl(y, src);
      const int32_t cyy = __shfl(cy, src);
      const uint32_t rx = (uint32_t)__shfl((int)rw.x, src);
      const uint32_t ry_all = (uint32_t)__shfl((int)rw.y, src);   // every lane shuffles
      const uint32_t ry = mine ? ry_all : 0u;
      uint32_t cnt = pull_slots<8>(S, yy, cyy, rx, sub, ry, kGroup, st, split, slots);
      cnt += __shfl_xor(cnt, 1);
      cnt += __shfl_xor(cnt, 2);
      cnt += __shfl_xor(cnt, 4);
      if (mine && sub == 0) {
        set_count(S, yy, cnt, st);
        ++rows;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // rows with a histogram and fewer than kMidLen slots, one per wave
  each_mid_row(S, lds, S.hrowv, S.n_hrows, [](int32_t) { return true; },
               [&](int32_t x) { return S.col0[x]; },
               [&](int32_t x, uint2 rw, int32_t cx) {
    uint32_t c2 = pull_slots<8>(S, x, cx, rw.x, lane, rw.y, 64, st, split, slots);
    for (int off = 32; off > 0; off >>= 1) c2 += __shfl_xor(c2, off);
    if (lane == 0) {
      set_count(S, x, c2, st);
      ++rows;
    }
  });
  // hub rows: one workgroup per hub (the chunk table's first entries)
  for (uint32_t i = blockIdx.x; i < S.n_hch; i += gridDim.x) {
    const int2 e = S.hch[i];
    if (e.y != 0) continue;
    const int32_t y = e.x;
    const uint2 rw = G.row[y];
    if (rw.y < kMidLen) continue;  // went by the short/mid pass (uniform)
    const uint32_t cnt =
        pull_slots<8>(S, y, S.col0[y], rw.x, threadIdx.x, rw.y, kBlock, st, split, slots);
    const uint32_t tot = block_sum(cnt, lds.red);
    if (threadIdx.x == 0) {
      set_count(S, y, tot, st);
      ++rows;
    }
  }
}

// The rows of D_r's members that changed colour (X_r), in slot ranges:
// fn(x, new colour, old colour, row start, lo, hi, stride) with short rows by
// a lane each, rows below kMidLen slots by a wave, hub rows by their chunks.
// Every thread must call this (it synchronises the workgroup).
template <class Fn>
__device__ __forceinline__ void walk_changed(const LoopState& S, const uint32_t* st, LoopLds& lds,
                                             Fn fn) {
  const GraphView& G = S.G;
  const int lane = lane_id();
  int32_t* list = lds.lists[threadIdx.x >> 6];
  walk_states(st, S.swords, kFX, list, [&](int32_t* lst, uint32_t total) {
    uint32_t nw = 0;
    for (uint32_t base = 0; base < total; base += 64) {
      const uint32_t i = base + lane;
      const bool chg = i < total;
      const int32_t x = chg ? lst[i] : 0;
      int32_t so = 0, sn = 0;
      uint2 rw = make_uint2(0u, 0u);
      if (chg) {
        so = S.dold[x].x;
        sn = S.col0[x];
        rw = G.row[x];
      }
      if (chg && rw.y < kSmallLen) fn(x, sn, so, (uint64_t)rw.x, 0u, rw.y, 1u);
      const bool mid = chg && rw.y >= kSmallLen && rw.y < kMidLen;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const unsigned long long mm = __ballot(mid);
      if (mid) lst[nw + lanes_below(mm)] = x;
      nw += (uint32_t)__popcll(mm);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t q = 0; q < nw; ++q) {
      const int32_t x = lst[q];
      const uint2 rw = G.row[x];
      fn(x, S.col0[x], S.dold[x].x, (uint64_t)rw.x, (uint32_t)lane, rw.y, 64u);
    }
  });
  each_hub_chunk(S, lds, S.hch, S.n_hch, [&](int2 e) { return st_has(st, e.x, kFX); },
                 [&](int32_t x) { return make_int2(S.col0[x], S.dold[x].x); },
                 [&](int32_t x, uint2 rw, uint32_t lo, uint32_t hi, int2 cs) {
    fn(x, cs.x, cs.y, (uint64_t)rw.x, lo + (uint32_t)lane, hi, 64u);
  });
}

// The new colours of D_r's members that changed into the slot cache, mex
// cache and histograms of their neighbours.
__device__ __forceinline__ void propagate_changed(const LoopState& S, const uint32_t* st,
                                                  LoopLds& lds) {
  walk_changed(S, st, lds, [&](int32_t, int32_t sn, int32_t so, uint64_t rb, uint32_t lo,
                               uint32_t hi, uint32_t stride) {
    propagate_slots(S.G, sn, so, rb, lo, hi, stride);
  });
}

// The terms a split pull skipped, for a changed member x (colour so -> sn):
// x marks a neighbour y whose colour cy is below so iff sn == cy and x < y,
// so only a member whose colour went down has any (sn < so).
__device__ __forceinline__ void fix_slots(const LoopState& S, uint32_t* st, int32_t x, int32_t sn,
                                          int32_t so, uint64_t rbase, uint32_t lo, uint32_t hi,
                                          uint32_t stride) {
  constexpr int U = 4;
  const GraphView& G = S.G;
  for (uint32_t j = lo; j < hi; j += U * stride) {
    int32_t y[U], cy[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t jj = j + k * stride;
      y[k] = jj < hi ? G.col[rbase + jj] : -1;
    }
#pragma unroll
    for (int k = 0; k < U; ++k) cy[k] = y[k] > x && y[k] < G.n_owned ? S.col0[y[k]] : -1;
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (cy[k] != sn) continue;   // sn >= 0, so this also drops the slots without a y
      atomicAdd(&S.K[y[k]], 1);
      if (!st_has(st, y[k], kFC)) st_set(st, y[k], kFR);
    }
  }
}

__device__ __forceinline__ void round_c(const LoopState& S, int r, LoopLds& lds) {
  const GraphView& G = S.G;
  uint32_t* st = S.St[r & 1];   // C_r, D_r, X_r; R_r is filled here
  const int lane = lane_id();
  int32_t* list = lds.lists[threadIdx.x >> 6];
  uint64_t words = 0;   // state-array and list words passed (block 0's count)
  // L_r, D_{r-1}, R_{r-1} and X_{r-1} are no longer read (cleared first: its
  // memory traffic overlaps the reads below)
  words += clear_state(S.St[(r + 1) & 1], S.swords, S.tq[(r + 1) & 1], S.qc[kQT + ((r + 3) & 3)],
                       S.tq_cap);
  // D_r's list counts (used when phase A listed the round)
  const uint32_t* qd = S.qc + kQD + 3 * (r & 1);
  const uint32_t nds = qd[0], ndm = qd[1], ndh = qd[2];
  // phase A's totals (every workgroup reads the same slots: uniform choices)
  if (threadIdx.x < 64) {
    unsigned long long v[5] = {0ull, 0ull, 0ull, 0ull, 0ull};
    if (lane < kAcc) {
      const unsigned long long* a = S.acc + ((size_t)(r & 1) * kAcc + lane) * kAccW;
#pragma unroll
      for (int k = 0; k < 5; ++k) v[k] = a[k];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
    if (lane == 0)
      for (int k = 0; k < 5; ++k) lds.tot[k] = v[k];
  }
  __syncthreads();
  const unsigned long long nmem = lds.tot[0];
  const bool nonempty = nmem != 0ull;
  const bool dhubs = lds.tot[3] != 0ull;
  const bool dmids = lds.tot[4] != 0ull;
  const int32_t nl = loop_nl(S.ctl, S.trace, r);
  // pull recount when D's rows are many slots and C more than doubled (a push
  // would spend an atomic on most new marking pairs), or when D's rows are
  // half of all slots; otherwise push deltas (on R-MAT 22 the first round's
  // D is a quarter of the slots with few term changes: push 0.8 ms faster).
  // pull_mode forces either path (tests).
  const bool big = lds.tot[2] > S.heavy_slots;
  const bool heavy =
      S.pull_mode > 0 ||
      (S.pull_mode == 0 && big &&
       (nmem > 2ull * (unsigned long long)nl || lds.tot[2] > 4ull * S.heavy_slots));
  // phase A listed round r's sets (the same test as there); R_r joins the
  // touched list in light rounds with few D slots
  const bool listed = nl <= S.list_c;
  const bool listing = listed && !heavy && lds.tot[2] <= S.list_slots;
  if (blockIdx.x == 0) {
    // round r's counts into the trace; round r+1's sums start at zero
    if (threadIdx.x == 0) {
      S.trace[r].x = nl;
      S.trace[r].y = (int32_t)nmem;
      S.trace[r].z = (int32_t)lds.tot[1];
      S.trace[r].w = 0;   // parts that leave after this round (part checks)
      S.ctl->last_nc = (int32_t)nmem;
      S.ctl->last_changed = (int32_t)lds.tot[1];
      if (!nonempty) S.ctl->done_round = r;
      if (nonempty && heavy) S.ctl->pull_rounds += 1;
      if (nonempty && heavy && r < S.pull_split_rounds) S.ctl->split_round = r;
      if (nonempty && !heavy && big) {
        S.ctl->heavy_push_rounds += 1;
        if (listing) S.ctl->heavy_push_listed += 1;
      }
      // round r+1's list counts start at zero; a round that does not list
      // R_r has no usable touched list (no block reads that count here)
      S.qc[kQT + ((r + 1) & 3)] = 0u;
      for (int k = 0; k < 3; ++k) S.qc[kQD + 3 * ((r + 1) & 1) + k] = 0u;
      if (nonempty && !listing) S.qc[kQT + (r & 3)] = kQInvalid;
    }
    for (int i = threadIdx.x; i < kAcc * kAccW; i += kBlock)
      S.acc[(size_t)((r + 1) & 1) * kAcc * kAccW + i] = 0ull;
    for (int p = threadIdx.x; p < S.nparts; p += kBlock)
      S.pchg[(size_t)((r + 1) & 1) * S.nparts + p] = 0u;
  }
  if (!nonempty) {
    // D_r's leaving members
    words += clear_state(st, S.swords, S.tq[r & 1], listed ? S.qc[kQT + (r & 3)] : kQInvalid,
                         S.tq_cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(S.wstat, (unsigned long long)words);
    return;
  }
  unsigned long long slots = 0, pslots = 0, rows = 0;
  if (heavy) {
    const bool split = r < S.pull_split_rounds;
    pull_all(S, st, lds, split, pslots, rows);
    if (!split) propagate_changed(S, st, lds);
    words += S.swords;
  } else {
  // D_r's lists when phase A made them (and none overflowed)
  const bool dl = listed && nds <= S.dq_s_cap && ndm <= S.dq_m_cap && ndh <= S.dq_h_cap;
  const Lister tl = listing ? Lister{S.tq[r & 1], S.qc + kQT + (r & 3), S.tq_cap}
                            : Lister{nullptr, nullptr, 0u};
  words += dl ? (uint64_t)nds + ndm + 2ull * ndh : (uint64_t)S.swords;
  // rows of D_r without a histogram (fewer than kSmallLen slots) by 8-lane
  // groups; the longer ones go by the histogram-row table and the hub chunk
  // table below.  A row whose owner changed colour then updates its
  // neighbours' caches.
  auto short_rows = [&](int32_t* lst, uint32_t total) {
    const int grp = lane >> 3, sub = lane & 7;
    int32_t* pick = lds.pick[threadIdx.x >> 6];
    // 64 entries at a time: every lane loads its entry's row and state at
    // once, the short rows are compacted, then 8-lane groups take them
    for (uint32_t c0 = 0; c0 < total; c0 += 64) {
      const uint32_t i = c0 + lane;
      const bool valid = i < total;
      const int32_t x = valid ? lst[i] : 0;
      const int32_t h = valid ? G.hidx[x] : 0;
      const uint2 rw = valid ? G.row[x] : make_uint2(0u, 0u);
      const DState d = load_d(S, st, x);
      const bool small = valid && h < 0;
      const unsigned long long sm = __ballot(small);
      if (small) pick[lanes_below(sm)] = lane;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const uint32_t ns = (uint32_t)__popcll(sm);
      for (uint32_t t = 0; t < ns; t += 8) {
        const bool mine = t + grp < ns;
        const int src = mine ? pick[t + grp] : lane;
        const DState e = bcast(d, src);
        const uint32_t rx = (uint32_t)__shfl((int)rw.x, src);
        // every lane takes part in the shuffle (a lane reading from a lane
        // that sits out would get an undefined value)
        const uint32_t ry_all = (uint32_t)__shfl((int)rw.y, src);
        const uint32_t ry = mine ? ry_all : 0u;
        uint32_t cnt = count_slots<4>(S, e, rx, sub, ry, kGroup, st, tl, slots);
        cnt += __shfl_xor(cnt, 1);
        cnt += __shfl_xor(cnt, 2);
        cnt += __shfl_xor(cnt, 4);
        if (mine && sub == 0) {
          S.K[e.x] = (int32_t)cnt;
          ++rows;
        }
        if (mine && e.chg) propagate_slots(G, e.sn, e.so, rx, sub, ry, kGroup);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  };
  if (dl)
    walk_list(S.dq_s, nds, list, short_rows);
  else
    walk_states(st, S.swords, kFD, list, short_rows);
  if (dmids) {
    each_mid_row(S, lds, dl ? S.dq_m : S.hrowv, dl ? ndm : S.n_hrows,
                 [&](int32_t v) { return dl || st_has(st, v, kFD); },
                 [&](int32_t x) { return load_d(S, st, x); },
                 [&](int32_t x, uint2 rw, const DState& d) {
      uint32_t cnt = count_slots<8>(S, d, rw.x, lane, rw.y, 64, st, tl, slots);
      for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
      if (lane == 0) {
        S.K[x] = (int32_t)cnt;
        ++rows;
      }
      if (d.chg) propagate_slots(G, d.sn, d.so, rw.x, lane, rw.y, 64);
    });
  }
  // hub rows of D_r by their kChunk-slot chunks, a quarter chunk per wave
  // (K[x] was zeroed in A; the chunks add up)
  if (dhubs) {
    each_hub_chunk(S, lds, dl ? S.dq_h : S.hch, dl ? ndh : S.n_hch,
                   [&](int2 e) { return dl || st_has(st, e.x, kFD); },
                   [&](int32_t x) { return load_d(S, st, x); },
                   [&](int32_t x, uint2 rw, uint32_t lo, uint32_t hi, const DState& d) {
      uint32_t cnt = count_slots<4>(S, d, rw.x, lo + lane, hi, 64, st, tl, slots);
      for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
      if (lane == 0) {
        if (cnt) atomicAdd(&S.K[x], (int32_t)cnt);
        if (threadIdx.x == 0) ++rows;
      }
      if (d.chg) propagate_slots(G, d.sn, d.so, rw.x, lo + lane, hi, 64);
    });
  }
  }
  for (int off = 32; off > 0; off >>= 1) {
    slots += __shfl_xor(slots, off);
    pslots += __shfl_xor(pslots, off);
    rows += __shfl_xor(rows, off);
  }
  if (lane == 0) {
    if (pslots) atomicAdd(&S.bstat[blockIdx.x * 4 + 1], pslots);
    if (slots) atomicAdd(&S.bstat[blockIdx.x * 4 + 2], slots);
    if (rows) atomicAdd(&S.bstat[blockIdx.x * 4 + 3], rows);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(S.wstat, (unsigned long long)words);
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_round_a(LoopState S, int r) {
  __shared__ LoopLds lds;
  if (S.ctl->done_round < r) return;
  round_a(S, r, lds);
}

__global__ __launch_bounds__(kBlock) void k_round_c(LoopState S, int r) {
  __shared__ LoopLds lds;
  if (S.ctl->done_round < r) return;
  round_c(S, r, lds);
}

// After a split pull recount of round r (launched after each count pass of
// the rounds below pull_split_rounds; others return at once): the changed
// members' skipped terms, then their new colours into the neighbours' caches.
__global__ __launch_bounds__(kBlock) void k_pull_fix(LoopState S, int r) {
  __shared__ LoopLds lds;
  if (S.ctl->split_round != r) return;
  uint32_t* st = S.St[r & 1];
  walk_changed(S, st, lds, [&](int32_t x, int32_t sn, int32_t so, uint64_t rb, uint32_t lo,
                               uint32_t hi, uint32_t stride) {
    if (sn < so) fix_slots(S, st, x, sn, so, rb, lo, hi, stride);
    propagate_slots(S.G, sn, so, rb, lo, hi, stride);
  });
}

// Persistent form: rounds r0..r1-1 in one cooperative launch.  The exit test
// reads the control word written before the last grid barrier, so every
// workgroup takes the same branch.
__global__ __launch_bounds__(kBlock) void k_loop_persistent(LoopState S, int r0, int r1) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  __shared__ LoopLds lds;
  for (int r = r0; r < r1; ++r) {
    round_a(S, r, lds);
    grid.sync();
    round_c(S, r, lds);
    grid.sync();
    if (S.ctl->done_round != INT_MAX) break;
  }
}

// Settle check, between rounds: does every owned vertex hold the mex of its
// neighbours' colours?  If so, and the round just run changed no colour,
// every later round of the reference recolours each member to the colour it
// has (its mex over an unchanged snapshot), so no colour changes again and
// the loop's result is the current colouring.  A cached mex is taken as it
// is; a stale one is recomputed (histogram or short-row scan, by 8-lane
// groups) and cached.  A mex past the histogram counts as a mismatch: the
// check only errs towards "not settled".  *unsettled is set, never cleared.
// Every owned vertex not skipped: is its colour the mex of its neighbours'
// colours?  report(v) for each that is not (a mex past the histogram counts
// as not).  A stale cached mex is recomputed and cached.
template <class Skip, class Report>
__device__ __forceinline__ void settle_scan(const LoopState& S, Skip skip, Report report) {
  __shared__ int32_t stale_list[kBlock / 64][64];
  const GraphView& G = S.G;
  const int lane = lane_id();
  const int grp = lane >> 3, sub = lane & 7;
  int32_t* lst = stale_list[threadIdx.x >> 6];
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t base = gwave * 64; base < G.n_owned; base += nwaves * 64) {
    const int64_t v = base + lane;
    const bool valid = v < G.n_owned && !skip((int32_t)v);
    const int32_t m = valid ? G.mexc[v] : 0;
    if (valid && m >= 0 && m != S.col0[v]) report((int32_t)v);
    const bool stale = valid && m < 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const unsigned long long sm = __ballot(stale);
    if (stale) lst[lanes_below(sm)] = (int32_t)v;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t ns = (uint32_t)__popcll(sm);
    for (uint32_t b8 = 0; b8 < ns; b8 += 8) {
      const uint32_t q = b8 + grp;
      const bool ok = q < ns;
      const int32_t w = ok ? lst[q] : 0;
      int32_t mw = -1;
      if (ok) {
        const int32_t h = G.hidx[w];
        mw = h >= 0 ? group_hist_mex(G.hist + (uint64_t)h * kHistW, sub)
                    : group_small_mex(G, G.row[w], sub);
      }
      if (ok && sub == 0) {
        if (mw < 0 || mw != S.col0[w]) report(w);
        if (mw >= 0) G.mexc[w] = mw;
      }
    }
  }
}

__device__ __forceinline__ void settle_check_body(const LoopState& S, int32_t* unsettled) {
  bool bad = false;
  settle_scan(S, [](int32_t) { return false; }, [&](int32_t) { bad = true; });
  if (__ballot(bad) != 0ull && lane_id() == 0) *unsettled = 1;
}

__global__ __launch_bounds__(kBlock) void k_settle_check(LoopState S, int32_t* unsettled) {
  settle_check_body(S, unsettled);
}

// Settle checks queued in the plain driver's stream after a round: the check
// runs only if the loop is still going, that round changed no colour and
// checks are left (every workgroup reads the same control words, which no
// kernel changes between the check and its decision); k_settle_decide then
// takes its result: settled colours end the loop after that round.
__device__ __forceinline__ bool settle_check_due(const LoopCtl* c) {
  return c->done_round == INT_MAX && c->last_changed == 0 && c->checks_left > 0;
}

__global__ __launch_bounds__(kBlock) void k_settle_check_if(LoopState S) {
  if (!settle_check_due(S.ctl)) return;
  settle_check_body(S, &S.ctl->unsettled);
}

// Part checks (parts set), queued after the check rounds: a part that
// changed no colour in round r and holds every vertex at its mex leaves the
// loop: its later rounds in the reference recolour each member to the colour
// it has, and no edge reaches another part.  k_part_check_if marks the parts
// that fail, k_part_decide (one workgroup) marks the parts that leave (pret
// 2) and counts them into round r's trace entry, k_part_clear takes their
// vertices out of the loop (count 0, no state flags).
__global__ __launch_bounds__(kBlock) void k_part_check_if(LoopState S, int32_t r) {
  if (S.ctl->done_round != INT_MAX) return;
  const uint32_t* pc = S.pchg + (size_t)(r & 1) * S.nparts;
  settle_scan(S,
              [&](int32_t v) {
                const int32_t p = S.vpart[v];
                return S.pret[p] != 0u || pc[p] != 0u;
              },
              [&](int32_t v) { S.punset[S.vpart[v]] = 1u; });
}

__global__ __launch_bounds__(1024) void k_part_decide(LoopState S, int32_t r) {
  if (S.ctl->done_round != INT_MAX) return;
  if (threadIdx.x == 0) {
    S.ctl->newly_left = 0;
    S.ctl->nl_sub = 0;
    S.ctl->nl_sub_round = r;
  }
  __syncthreads();
  const uint32_t* pc = S.pchg + (size_t)(r & 1) * S.nparts;
  uint32_t left = 0;
  for (int32_t p = threadIdx.x; p < S.nparts; p += blockDim.x) {
    uint32_t ret = S.pret[p];
    if (ret == 2u) {
      S.pret[p] = 1u;
      ret = 1u;
    }
    if (ret == 0u && pc[p] == 0u && S.punset[p] == 0u) {
      S.pret[p] = 2u;
      ++left;
    }
    S.punset[p] = 0u;
  }
  if (left) {
    atomicAdd(&S.trace[r].w, (int32_t)left);
    S.ctl->newly_left = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_part_clear(LoopState S, int32_t r) {
  if (S.ctl->newly_left == 0) return;
  // members of C_r taken out: round r + 1's L does not hold them
  const uint32_t* cr = S.St[r & 1];
  uint32_t gone = 0;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < S.swords; w += gridDim.x * blockDim.x) {
    const uint32_t a = S.St[0][w], b = S.St[1][w];
    const uint32_t c = cr[w];
    uint32_t clear = 0u;
    for (int i = 0; i < 8; ++i) {
      const uint32_t f = ((a | b) >> (4 * i)) & 15u;
      const int32_t v = (int32_t)(w * 8 + i);
      if (f && v < S.G.n_owned && S.pret[S.vpart[v]] == 2u) {
        S.K[v] = 0;
        clear |= 15u << (4 * i);
        gone += (c >> (4 * i)) & kFC;
      }
    }
    if (clear) {
      S.St[0][w] = a & ~clear;
      S.St[1][w] = b & ~clear;
    }
  }
  wave_add_u32(gone, reinterpret_cast<uint32_t*>(&S.ctl->nl_sub));
}

__global__ void k_settle_decide(LoopState S, int32_t r) {
  LoopCtl* c = S.ctl;
  if (threadIdx.x != 0 || !settle_check_due(c)) return;
  c->checks_left -= 1;
  if (c->unsettled == 0) {
    c->done_round = r;
    c->settled = 1;
  }
  c->unsettled = 0;
}

// K and the loop's state arrays are zero between batches.
__global__ void k_zero_check(const int32_t* K, size_t nk, const uint32_t* bits, size_t words,
                             unsigned long long* bad) {
  unsigned long long cnt = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (K)
    for (size_t i = t0; i < nk; i += stride) cnt += K[i] != 0;
  for (size_t i = t0; i < words; i += stride) cnt += bits[i] != 0u;
  wave_add_u64(cnt, bad);
}

// ============================================================================
// host side
static IncState* inc_state(Graph* g) { return static_cast<IncState*>(g->inc); }

static int inc_alloc(Graph* g, IncState* s) {
  const size_t n = (size_t)g->n + 1;
  LoopState& S = s->S;
  S.swords = (uint32_t)(((size_t)g->n + 8) / 8);
  DC_HIP(hipMalloc(&S.K, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.dold, sizeof(int2) * n));
  DC_HIP(hipMalloc(&s->bits, sizeof(uint32_t) * 2 * (size_t)S.swords));
  S.St[0] = s->bits;
  S.St[1] = s->bits + S.swords;
  DC_HIP(hipMalloc(&S.acc, sizeof(unsigned long long) * 2 * kAcc * kAccW));
  // per-workgroup stats, then the count passes' word count (wstat)
  DC_HIP(hipMalloc(&S.bstat, sizeof(unsigned long long) * (4 * kMaxBlocks + 1)));
  S.wstat = S.bstat + 4 * kMaxBlocks;
  DC_HIP(hipMalloc(&S.ctl, sizeof(LoopCtl)));
  DC_HIP(hipMalloc(&s->tqbuf, sizeof(int32_t) * 2 * n));
  S.tq[0] = s->tqbuf;
  S.tq[1] = s->tqbuf + n;
  S.tq_cap = (uint32_t)n;
  DC_HIP(hipMalloc(&S.dq_s, sizeof(int32_t) * n));
  S.dq_s_cap = (uint32_t)n;
  S.dq_m = nullptr;
  S.dq_h = nullptr;
  S.dq_m_cap = S.dq_h_cap = 0;
  DC_HIP(hipMalloc(&S.qc, sizeof(uint32_t) * kQWords));
  DC_HIP(hipMalloc(&s->terms, sizeof(unsigned long long) * 4));
  DC_HIP(hipHostMalloc(&s->hctl, sizeof(LoopCtl), hipHostMallocDefault));
  s->hstat = new unsigned long long[4 * kMaxBlocks + 1];
  DC_HIP(hipMemsetAsync(S.K, 0, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(s->bits, 0, sizeof(uint32_t) * 2 * (size_t)S.swords, g->stream));

  return DC_OK;
}

static int inc_trace(Graph* g, IncState* s, int32_t need) {
  if (need <= s->trace_cap) return DC_OK;
  int32_t cap = std::max<int32_t>(1024, s->trace_cap);
  while (cap < need) cap *= 2;
  int4* t = nullptr;
  DC_HIP(hipMalloc(&t, sizeof(int4) * (size_t)cap));
  if (s->S.trace) {
    DC_HIP(hipMemcpyAsync(t, s->S.trace, sizeof(int4) * (size_t)s->trace_cap,
                          hipMemcpyDeviceToDevice, g->stream));
    DC_HIP(hipStreamSynchronize(g->stream));
    DC_HIP(hipFree(s->S.trace));
  }
  s->S.trace = t;
  s->trace_cap = cap;
  return DC_OK;
}

void loop_free(Graph* g) {
  IncState* s = inc_state(g);
  if (!s) return;
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(s->S.K); f(s->S.dold); f(s->bits); f(s->S.acc); f(s->S.bstat); f(s->S.trace);
  f(s->S.ctl); f(s->tqbuf); f(s->S.dq_s); f(s->S.dq_m); f(s->S.dq_h); f(s->S.qc); f(s->terms);
  if (s->hctl) (void)hipHostFree(s->hctl);
  for (hipEvent_t e : s->cev) (void)hipEventDestroy(e);
  for (auto& e : s->graphs) (void)hipGraphExecDestroy(e.exec);
  delete[] s->hstat;
  delete s;
  g->inc = nullptr;
}

int loop_check(Graph* g, int64_t* bad) {
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(g->aux);
  DC_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), g->stream));
  hipLaunchKernelGGL(k_zero_check, dim3(256), dim3(kBlock), 0, g->stream, (const int32_t*)nullptr,
                     (size_t)0, g->cbits[0], 2 * g->cbits_words, cnt);
  IncState* s = inc_state(g);
  if (s)
    hipLaunchKernelGGL(k_zero_check, dim3(1024), dim3(kBlock), 0, g->stream, s->S.K,
                       (size_t)g->n, s->bits, 2 * (size_t)s->S.swords, cnt);
  DC_HIP(hipGetLastError());
  unsigned long long h = 0;
  DC_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  *bad = (int64_t)h;
  return DC_OK;
}

static int loop_reset(Graph* g, IncState* s) {
  DC_HIP(hipMemsetAsync(s->S.K, 0, sizeof(int32_t) * ((size_t)g->n + 1), g->stream));
  DC_HIP(hipMemsetAsync(s->bits, 0, sizeof(uint32_t) * 2 * (size_t)s->S.swords, g->stream));
  return DC_OK;
}

static int persistent_grid(Graph* g, int* grid) {
  int dev = 0, cus = 0, bpm = 0;
  DC_HIP(hipGetDevice(&dev));
  DC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  DC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpm, (const void*)k_loop_persistent, kBlock, 0));
  if (bpm < 1) {
    set_error("persistent loop kernel does not fit on a CU");
    return DC_ERR_HIP;
  }
  *grid = std::min(kMaxBlocks, std::max(1, std::min(bpm, g->fused_blocks_per_cu)) * cus);
  return DC_OK;
}

// Host blocks of rounds of the persistent driver (the host reads the control
// word after each and may run a settle check): rounds 0-2, then one at a time
// up to 8 (a batch's changes usually end by round 4, and the settle check
// wants the first round without a change), 8-11, then blocks ending at
// 2^k - 4 (28, 60, 124, ...).
static int32_t block_end_persistent(int32_t r) {
  if (r < 3) return 3;
  if (r < 8) return r + 1;
  if (r < 12) return 12;
  int32_t e = 28;
  while (e <= r) e = 2 * e + 4;
  return e;
}

// Host blocks of the plain driver, whose settle checks are queued on the
// device between rounds: 0-2, 3-8, 9-12, then blocks ending at 2^k - 3 (29,
// 61, ...), at most 32 rounds long.
static int32_t block_end_plain(int32_t r) {
  if (r < 3) return 3;
  if (r < 9) return 9;
  if (r < 13) return 13;
  int32_t e = 29;
  while (e <= r) e = 2 * e + 3;
  return std::min(e, r + 32);
}

// rounds after which a settle check may run (the persistent driver's block
// ends; the plain driver queues its checks after the same rounds)
static bool settle_point(int32_t r) {
  return (r >= 3 && r <= 8) || r == 12 || (r >= 28 && ((r + 4) & (r + 3)) == 0);
}

int loop_hi(Graph* g, Frontier L0, int32_t nl0, bool persistent, LoopResult* res) {
  IncState* s = inc_state(g);
  if (!s) {
    s = new IncState();
    g->inc = s;
    int rc = inc_alloc(g, s);
    if (rc) return rc;
  }
  int rc = inc_trace(g, s, 256);
  if (rc) return rc;
  const int32_t max_rounds = g->opt.max_rounds;
  const dc_tuning tun = g->tuning;

  LoopState& S = s->S;
  S.G = g->view();
  S.col0 = g->color[g->cur];
  S.col1 = g->color[g->cur ^ 1];
  S.prev = g->prev;
  S.L0 = L0;
  S.hch = g->hch;
  S.n_hch = g->n_hch;
  S.hrowv = g->hrowv;
  S.n_hrows = g->n_hrows;
  // disjoint parts: checked on their own in the plain driver
  S.nparts = (!persistent && !g->opt.exact_rounds) ? g->nparts : 0;
  S.vpart = S.nparts ? g->vpart : nullptr;
  S.pchg = S.nparts ? g->pstate : nullptr;
  S.punset = S.nparts ? g->pstate + 2 * (size_t)S.nparts : nullptr;
  S.pret = S.nparts ? g->pstate + 3 * (size_t)S.nparts : nullptr;
  // D's mid-row and hub-chunk lists hold at most the static tables' entries
  if (S.dq_m_cap < g->n_hrows) {
    if (S.dq_m) DC_HIP(hipFree(S.dq_m));
    S.dq_m = nullptr;
    S.dq_m_cap = 0;
    DC_HIP(hipMalloc(&S.dq_m, sizeof(int32_t) * ((size_t)g->hist_cap + 1)));
    S.dq_m_cap = g->hist_cap + 1;
  }
  if (S.dq_h_cap < g->n_hch) {
    if (S.dq_h) DC_HIP(hipFree(S.dq_h));
    S.dq_h = nullptr;
    S.dq_h_cap = 0;
    DC_HIP(hipMalloc(&S.dq_h, sizeof(int2) * ((size_t)g->hch_cap + 1)));
    S.dq_h_cap = g->hch_cap + 1;
  }
  // rounds whose D rows hold more than 1/8 of all slots (measured on R-MAT
  // 22 against 1/2 and 1/32) recount by pull when C also more than doubled
  S.heavy_slots = g->opt.heavy_slots > 0
                      ? (unsigned long long)g->opt.heavy_slots
                      : std::max<unsigned long long>(g->rows_end / (uint64_t)tun.heavy_div, 1ull << 16);
  S.pull_mode = g->opt.pull_mode > 0 ? 1 : g->opt.pull_mode < 0 ? -1 : 0;
  // the reference's work terms are counted in profiled batches only (one
  // more row-bounds gather per candidate)
  S.terms = g->profiling ? s->terms : nullptr;
  S.list_c = (int32_t)std::min<uint64_t>(
      (uint64_t)tun.list_c, ((uint64_t)S.swords + g->n_hrows + g->n_hch) / 64);
  S.list_slots = (unsigned long long)tun.list_slots;
  if (g->opt.list_limit != 0) {
    S.list_c = g->opt.list_limit < 0 ? -1 : g->opt.list_limit;
    S.list_slots = ~0ull;
  }

  *s->hctl = LoopCtl{};
  s->hctl->done_round = INT_MAX;
  s->hctl->nl0 = nl0;
  s->hctl->checks_left = g->opt.exact_rounds ? 0 : 3;
  s->hctl->nl_sub_round = -1;
  s->hctl->split_round = -1;
  DC_HIP(hipMemcpyAsync(S.ctl, s->hctl, sizeof(LoopCtl), hipMemcpyHostToDevice, g->stream));
  DC_HIP(hipMemsetAsync(S.bstat, 0, sizeof(unsigned long long) * (4 * kMaxBlocks + 1), g->stream));
  DC_HIP(hipMemsetAsync(S.acc, 0, sizeof(unsigned long long) * 2 * kAcc * kAccW, g->stream));
  if (S.terms) DC_HIP(hipMemsetAsync(S.terms, 0, sizeof(unsigned long long) * 4, g->stream));
  const dim3 blk(kBlock);
  hipLaunchKernelGGL(k_loop_init, dim3(tun.loop_blocks), blk, 0, g->stream, S);
  DC_HIP(hipGetLastError());
  int64_t launches = 1;
  int grid = tun.loop_blocks;
  if (persistent && (rc = persistent_grid(g, &grid))) return rc;
  // the first rounds' pull recounts (round 1 on R-MAT 22, where C grows
  // eightfold) are split with k_pull_fix; the persistent kernel has no
  // launch between rounds to split them with
  S.pull_split_rounds = persistent ? 0 : kPullSplitRounds;
  // captured blocks replace the launches in unprofiled batches of the graph
  // driver (per-launch events cannot be recorded inside a capture)
  const bool use_graphs = !persistent && g->opt.driver == DC_DRIVER_GRAPH && !g->profiling;
  // light rounds (small C) launch fewer workgroups: their work is a few
  // thousand vertices, and dispatching 2048 workgroups twice a round costs
  // more than it brings
  int light_grid = tun.light_blocks;
  if (light_grid <= 0) {
    int dev = 0, cus = 0;
    DC_HIP(hipGetDevice(&dev));
    DC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    light_grid = std::min(kMaxBlocks, std::max(32, 2 * cus));
  }
  int round_grid = grid;
  double kms = 0.0, cms = 0.0, ams = 0.0;
  int64_t claunches = 0;
  hipEvent_t e0 = g->ev[2], e1 = g->ev[3];
  // settle checks: at most three per batch, after a round without a change,
  // at the check rounds of settle_point (block ends of both drivers, so both
  // stop at the same round)
  int settle_left = g->opt.exact_rounds ? 0 : 3;
  bool settled = false;
  int32_t rounds = -1, r = 0;
  const int64_t captures0 = s->captures;
  while (r < max_rounds && rounds < 0) {
    const int32_t nb =
        std::min(persistent ? block_end_persistent(r) : block_end_plain(r), max_rounds) - r;
    if ((rc = inc_trace(g, s, r + nb))) return rc;
    if (g->profiling) {
      while (s->cev.size() < 6 * (size_t)nb) {
        hipEvent_t e;
        DC_HIP(hipEventCreate(&e));
        s->cev.push_back(e);
      }
      DC_HIP(hipEventRecord(e0, g->stream));
    }
    // rounds r .. r + nb - 1 of the plain driver: two launches per round, and
    // after the check rounds the queued settle (and part) checks
    auto enqueue_block = [&](bool events) -> int {
      for (int32_t i = 0; i < nb; ++i) {
        const int32_t ri = r + i;
        if (events) {
          // events taken at the launches' own start and end (an event
          // recorded before a launch would also time the host's submission
          // at the start of a block)
          hipExtLaunchKernelGGL(k_round_a, dim3(round_grid), blk, 0, g->stream, s->cev[6 * i],
                                s->cev[6 * i + 1], 0u, S, ri);
          hipExtLaunchKernelGGL(k_round_c, dim3(round_grid), blk, 0, g->stream, s->cev[6 * i + 2],
                                s->cev[6 * i + 3], 0u, S, ri);
          if (ri < S.pull_split_rounds)
            hipExtLaunchKernelGGL(k_pull_fix, dim3(round_grid), blk, 0, g->stream,
                                  s->cev[6 * i + 4], s->cev[6 * i + 5], 0u, S, ri);
        } else {
          hipLaunchKernelGGL(k_round_a, dim3(round_grid), blk, 0, g->stream, S, ri);
          hipLaunchKernelGGL(k_round_c, dim3(round_grid), blk, 0, g->stream, S, ri);
          if (ri < S.pull_split_rounds)
            hipLaunchKernelGGL(k_pull_fix, dim3(round_grid), blk, 0, g->stream, S, ri);
        }
        if (ri < S.pull_split_rounds) ++launches;
        if (settle_point(r + i + 1) && !g->opt.exact_rounds) {
          hipLaunchKernelGGL(k_settle_check_if, dim3(tun.loop_blocks), blk, 0, g->stream, S);
          hipLaunchKernelGGL(k_settle_decide, dim3(1), dim3(64), 0, g->stream, S, r + i);
          launches += 2;
          if (S.nparts) {
            hipLaunchKernelGGL(k_part_check_if, dim3(tun.loop_blocks), blk, 0, g->stream, S, r + i);
            hipLaunchKernelGGL(k_part_decide, dim3(1), dim3(1024), 0, g->stream, S, r + i);
            hipLaunchKernelGGL(k_part_clear,
                               dim3((uint32_t)std::min<uint64_t>(2048, (S.swords + kBlock - 1) / kBlock)),
                               blk, 0, g->stream, S, r + i);
            launches += 3;
          }
        }
      }
      DC_HIP(hipGetLastError());
      launches += 2 * nb;
      return DC_OK;
    };
    if (persistent) {
      int32_t r0 = r, r1 = r + nb;
      void* args[] = {&S, &r0, &r1};
      DC_HIP(hipLaunchCooperativeKernel((const void*)k_loop_persistent, dim3(grid), blk, args, 0,
                                        g->stream));
      ++launches;
    } else if (use_graphs) {
      // the block as a captured HIP graph, replayed while the loop state, the
      // round range and the grid are the same (kernel arguments are baked in)
      GraphBlock* gb = nullptr;
      for (auto& e : s->graphs)
        if (e.r0 == r && e.nb == nb && e.grid == round_grid &&
            std::memcmp(e.state.data(), &S, sizeof(S)) == 0) {
          gb = &e;
          break;
        }
      if (!gb) {
        if (s->graphs.size() >= kGraphCache) {
          for (auto& e : s->graphs) (void)hipGraphExecDestroy(e.exec);
          s->graphs.clear();
        }
        hipGraph_t graph = nullptr;
        DC_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeRelaxed));
        const int64_t l0 = launches;
        const int erc = enqueue_block(false);
        const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
        if (erc) return erc;
        if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture");
        launches = l0;
        GraphBlock nbk;
        nbk.r0 = r;
        nbk.nb = nb;
        nbk.grid = round_grid;
        nbk.state.resize(sizeof(S));
        std::memcpy(nbk.state.data(), &S, sizeof(S));
        const hipError_t ei = hipGraphInstantiate(&nbk.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
        s->graphs.push_back(std::move(nbk));
        gb = &s->graphs.back();
        ++s->captures;
      }
      DC_HIP(hipGraphLaunch(gb->exec, g->stream));
      ++launches;
    } else {
      if ((rc = enqueue_block(g->profiling))) return rc;
    }
    if (g->profiling) DC_HIP(hipEventRecord(e1, g->stream));
    DC_HIP(hipMemcpyAsync(s->hctl, S.ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost, g->stream));
    DC_HIP(hipStreamSynchronize(g->stream));
    const int32_t done = s->hctl->done_round;
    if (g->profiling) {
      float x = 0.f;
      DC_HIP(hipEventElapsedTime(&x, e0, e1));
      kms += x;
      if (!persistent) {
        // phase-A and count-pass launches of executed rounds (later ones
        // return at once)
        for (int32_t i = 0; i < nb && r + i <= done; ++i) {
          DC_HIP(hipEventElapsedTime(&x, s->cev[6 * i], s->cev[6 * i + 1]));
          ams += x;
          DC_HIP(hipEventElapsedTime(&x, s->cev[6 * i + 2], s->cev[6 * i + 3]));
          cms += x;
          if (r + i < S.pull_split_rounds) {   // the count pass's split-off part
            DC_HIP(hipEventElapsedTime(&x, s->cev[6 * i + 4], s->cev[6 * i + 5]));
            cms += x;
          }
          ++claunches;
        }
      }
    }
    r += nb;
    if (done != INT_MAX) {
      rounds = done + 1;
      if (s->hctl->settled) {
        // a queued settle check ended the loop after round `done`
        settled = true;
        if ((rc = loop_reset(g, s))) return rc;
      }
    } else if (persistent && settle_left > 0 && settle_point(r) && s->hctl->last_changed == 0) {
      // rounds 0..r-1 ran in full; the last one changed nothing
      --settle_left;
      int32_t* flag = reinterpret_cast<int32_t*>(g->aux);
      DC_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), g->stream));
      hipLaunchKernelGGL(k_settle_check, dim3(tun.loop_blocks), blk, 0, g->stream, S, flag);
      DC_HIP(hipGetLastError());
      ++launches;
      int32_t h = 1;
      DC_HIP(hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, g->stream));
      DC_HIP(hipStreamSynchronize(g->stream));
      if (h == 0) {
        settled = true;
        rounds = r;
        if ((rc = loop_reset(g, s))) return rc;
      }
    }
    round_grid = s->hctl->last_nc < tun.light_c ? light_grid : grid;
  }
  if (rounds < 0) {
    int rc2 = loop_reset(g, s);
    if (rc2) return rc2;
    set_error("recolor loop hit max_rounds");
    return DC_ERR_ROUNDS;
  }
  std::vector<int4> tr((size_t)rounds);
  DC_HIP(hipMemcpyAsync(tr.data(), S.trace, sizeof(int4) * (size_t)rounds, hipMemcpyDeviceToHost,
                        g->stream));
  DC_HIP(hipMemcpyAsync(s->hstat, S.bstat, sizeof(unsigned long long) * (4 * kMaxBlocks + 1),
                        hipMemcpyDeviceToHost, g->stream));
  unsigned long long terms[4] = {0ull, 0ull, 0ull, 0ull};
  if (S.terms)
    DC_HIP(hipMemcpyAsync(terms, S.terms, sizeof(terms), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  res->pull_rounds = s->hctl->pull_rounds;
  res->heavy_push_rounds = s->hctl->heavy_push_rounds;
  res->heavy_push_listed = s->hctl->heavy_push_listed;
  res->graph_captures = (int32_t)(s->captures - captures0);
  res->lc_vertices = terms[0];
  res->lc_slots = terms[1];
  res->rounds = rounds;
  res->settled = settled;
  for (int32_t i = 0; i < rounds; ++i) {
    g->trace.insert(g->trace.end(),
                    {(int64_t)10, (int64_t)tr[i].x, (int64_t)tr[i].y, (int64_t)tr[i].z});
    // parts found settled after this round left the loop
    if (tr[i].w > 0)
      g->trace.insert(g->trace.end(), {(int64_t)12, (int64_t)i + 1, (int64_t)tr[i].w, (int64_t)0});
    res->frontier_sum += (uint64_t)tr[i].x;
    res->evaluated += (uint64_t)tr[i].y;
    res->changed += (uint64_t)tr[i].z;
  }
  if (settled) g->trace.insert(g->trace.end(), {(int64_t)11, (int64_t)rounds, (int64_t)0, (int64_t)0});
  uint64_t pull = 0, cnt = 0, rowv = 0;
  for (int b = 0; b < kMaxBlocks; ++b) {
    res->rslots += s->hstat[4 * b + 0];
    pull += s->hstat[4 * b + 1];
    cnt += s->hstat[4 * b + 2];
    rowv += s->hstat[4 * b + 3];
  }
  res->slots = pull + cnt;
  res->kernel_ms = kms;
  res->launches = launches;
  // Algorithmic bytes of the count passes: the state-array words they clear
  // and walk and the list entries they read (4 B each, counted on the
  // device); per count-pass slot col + ccol (8 B); per pulled slot col (4 B);
  // per row or chunk visit its bounds and count or saved state (16 B).
  res->count_ms = cms;
  res->round_a_ms = ams;
  res->count_launches = claunches;
  res->count_bytes = 4ull * s->hstat[4 * kMaxBlocks] + 8ull * cnt + 4ull * pull + 16ull * rowv;
  return DC_OK;
}

}  // namespace dc
