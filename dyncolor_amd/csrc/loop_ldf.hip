// loop_ldf.hip -- the heuristic-based (largest-degree-first) recolor loop
// with its sets kept incrementally.
//
// The reference loop (HB/Color_main.cu:203-224) repeats
//   C = { w : some v in L has an edge (v, w) with v > w }          (v precedes w)
//     u { v in L : some neighbour w > v has v's colour }           (findEligibleNeighbors,
//                                                                   HB/gpuFunctions_undir.cuh:259-291)
//   for w in C: w's colour = mex of its predecessors' colours, if it differs
//                                                                  (recolorNeighbor, :295-314)
//   L = C
// while some colour changed.  With prop fixed during a batch the predecessor
// relation is fixed, so the loop can be kept in counts:
//   K[w]  = predecessor slots of w whose neighbour is in L,
//   cc[v] = predecessor slots of v whose neighbour has v's colour,
//   pm[v] = mex of v's predecessors' colours (cached; -1 stale),
// and C = { w : K[w] > 0 } u { v in L : cc[v] > 0 }.  Between rounds only
// D = C xor L moves K (the rows of D add +-1 to their successors), and only
// the members whose colour changed (X) move cc and the cached mex of their
// successors: a row is read when its vertex enters or leaves L or changes
// colour, not every round it stays in L.  Colours, traces and round counts
// are those of the re-derivation (tests compare both with the oracle).
//
// One round = four launches over per-vertex flag bytes (no worklists, so no
// appends on shared counters): members and their mex (short rows by their
// own thread), stale mex of long rows (a wave per row of the long-row
// table), the row passes of D and X over their successors (K, cc, cached
// mexes), commit.  A vertex's colour, its new colour in the round and its
// cached mex sit together in one 16-byte record, so the row pass gathers one
// record per successor; cc has an array of its own, so that the row pass's
// count atomics do not hit the lines it gathers (on MI355X that costs a
// third of the pass).  The batch's set-up records which slots hold a
// predecessor (pdir, and for long rows a list of them), so the scans and row
// passes read no priorities; the scans read a predecessor's colour from the
// colour array, so a colour change writes into no slot during the loop: the
// changed vertices' slots in their neighbours' rows, and their neighbours'
// histograms and all-neighbour mex caches (the independent rule's), take the
// final colours once after the loop (k_ldf_settle).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "dc_phases.h"

namespace dc {

constexpr int kLdfBlocks = 2048;
constexpr int kLdfAcc = 32;        // round sums: workgroup b adds into slot b % kLdfAcc
constexpr int kWin32 = kWinBits / 32;   // 32-bit words of a wave's mex window
constexpr uint8_t kFlD = 1, kFlX = 2, kFlPm = 4;   // per-vertex flags of a round
// pdir per slot: the neighbour does not precede the row's vertex (kPdNone) or
// precedes it (kPdPred).  During the loop a predecessor's colour is read from
// the colour array, not from the slot colour cache, so a colour change writes
// into none of its successors' rows (measured on R-MAT 22: 5 % faster than
// reading predecessors below 512 slots from their slots, which costs the row
// pass a slot write per successor).
constexpr uint8_t kPdNone = 0, kPdPred = 1;

// Per-vertex state of the loop, gathered as one 16-byte record: the colour
// (a copy of the colour array, kept by the commit), the new colour of a
// vertex that changes in the round (-1 otherwise, rewritten every round) and
// the cached predecessor mex (-1 stale).
struct alignas(16) LdfV {
  int32_t c, nc, pm, pad;
};

// after the loop, every row's slot colours are rewritten once the changed
// vertices' rows hold more than 1/kSettleDiv of the laid-out slots (and more
// than kSettleMin): a row pass costs ~6x a rewrite per slot on MI355X
constexpr uint64_t kSettleDiv = 6;
constexpr uint64_t kSettleMin = 1ull << 16;

struct LdfCtl {
  int32_t done_round;   // the last round run (INT_MAX while running)
  int32_t pad;
  unsigned long long moved_slots;   // row slots of the vertices changed so far
};

// Rows by length: short (cap < kSmallLen) by their own thread; longer rows
// (the histogram-row table) by a wave for their mex, and for their row
// passes by a wave (cap < kMidLen) or a workgroup per kChunk-slot chunk of
// the hub chunk table (cap >= kMidLen).
struct LdfState {
  GraphView G;
  int32_t* c0;          // colour buffers (both written at commit: they agree)
  int32_t* c1;
  int32_t* K;
  LdfV* V;
  int32_t* cc;          // predecessors with the vertex's colour (atomics of the row pass)
  int32_t* corig;       // colour at the loop's start of every vertex that changed (-1: none)
  uint8_t* pdir;        // slot -> its neighbour precedes the row's vertex (set up per batch)
  uint32_t* pslot;      // rows with cap >= kSmallLen: their predecessors' vertex IDs, listed
                        // from the row's start (set up per batch)
  int32_t* npred;       // ... and how many
  uint8_t* F[2];        // F[r & 1] = L_r, F[(r + 1) & 1] = C_r = L_{r+1}
  uint8_t* fl;          // kFlD / kFlX / kFlPm of the current round
  const int32_t* hrowv; // rows with cap >= kSmallLen
  uint32_t n_hrows;
  const int2* hch;      // {hub, chunk} for every kChunk slots of every hub's capacity
  uint32_t n_hch;
  unsigned long long* acc;   // [2][kLdfAcc][4]: |L|, |C|, changed
  int4* trace;
  LdfCtl* ctl;
  // profiling: the reference's work terms, summed over rounds: |L_r u C_r| and
  // the row slots of L_r u C_r (the rows it rescans every round)
  unsigned long long* terms;
  int32_t prof;
};

// ---------------------------------------------------------------------------
// Predecessor colour bits below 64 are kept as two 32-bit words.
__device__ __forceinline__ void add_colour_bit(uint32_t c, uint32_t& lo, uint32_t& hi) {
  lo |= c < 32u ? 1u << c : 0u;
  hi |= c - 32u < 32u ? 1u << (c - 32u) : 0u;
}
// mex of colour bits 0..63, -1 when all are present
__device__ __forceinline__ int32_t mex64(uint32_t lo, uint32_t hi) {
  return lo != ~0u ? __ffs(~lo) - 1 : hi != ~0u ? 32 + __ffs(~hi) - 1 : -1;
}

// Set-up of a short row (used < kSmallLen) of v by its own thread: classifies
// the slots by the priorities into pdir (fixed for the batch: no priority
// moves inside the loop; the loop's scans read pdir), and returns the mex of
// the predecessors' colours and the predecessor slots with colour cv, from
// the slot colour cache (current at the batch's start).  At most 63
// predecessors, so the mex is below 64.
__device__ __forceinline__ void short_pred_init(const GraphView& G, uint8_t* pdir, int32_t v,
                                                int32_t cv, int32_t& pm, int32_t& cc) {
  const uint2 r = G.row[v];
  const int32_t pv = G.prop[v];
  uint32_t lo = 0u, hi = 0u;
  int32_t k = 0;
  for (uint32_t j0 = 0; j0 < r.y; j0 += 4) {
    int32_t c[4];
    uint8_t pd[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool in = j0 + q < r.y;
      const uint64_t j = (uint64_t)r.x + j0 + q;
      c[q] = in ? G.ccol[j] : -1;
      const int32_t w = in ? G.col[j] : -1;
      const bool pr = w >= 0 && precedes(nullptr, w, G.prop[w], v, pv);
      pd[q] = pr ? kPdPred : kPdNone;
      if (in) pdir[j] = pd[q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (pd[q] == kPdNone) continue;
      add_colour_bit((uint32_t)c[q], lo, hi);
      k += c[q] == cv ? 1 : 0;
    }
  }
  pm = mex64(lo, hi);
  cc = k;
}

// A whole row of v by one wave: the mex over LDS windows of kWinBits colours
// (a window past the first only when it is full) and the predecessor slots
// with colour cv.  Every lane returns both.  INIT (the batch's set-up, colours
// from the slot colour cache) classifies every slot by the priorities (pdir,
// as for short_pred_init) and lists the predecessors' vertex IDs in
// pslot[row start ...] (npred[v] of them: the row's own span of a slot-sized
// array); the loop's scans walk that list and read the colour array, so a
// hub row with few predecessors costs a few loads however long it is.
template <bool INIT>
__device__ __forceinline__ void wave_pred_scan(const GraphView& G, uint8_t* pdir, uint32_t* pslot,
                                               int32_t* npred, const int32_t* col0, int32_t v,
                                               int32_t cv, uint32_t* ww, int32_t& pm, int32_t& cc) {
  const int lane = lane_id();
  const uint2 r = G.row[v];
  int32_t k = 0;
  uint32_t np;
  for (int i = lane; i < kWin32; i += 64) ww[i] = 0u;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (INIT) {
    // classify and list the slots; window 0's bits from the same pass
    const int32_t pv = G.prop[v];
    uint32_t base = 0;
    for (uint32_t j0 = 0; j0 < r.y; j0 += 64) {
      const uint32_t j = j0 + (uint32_t)lane;
      const uint64_t s = (uint64_t)r.x + j;
      const int32_t w = j < r.y ? G.col[s] : -1;
      const bool pr = w >= 0 && precedes(nullptr, w, G.prop[w], v, pv);
      if (j < r.y) pdir[s] = pr ? kPdPred : kPdNone;
      const unsigned long long b = __ballot(pr);
      if (pr) {
        pslot[(uint64_t)r.x + base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull))] = (uint32_t)w;
        const int32_t c = G.ccol[s];
        if (c == cv) ++k;
        if ((uint32_t)c < (uint32_t)kWinBits) atomicOr(&ww[(uint32_t)c >> 5], 1u << ((uint32_t)c & 31u));
      }
      base += (uint32_t)__popcll(b);
    }
    if (lane == 0) npred[v] = (int32_t)base;
    np = base;
    // the list is read back (by other lanes) only past a full first window
    __threadfence_block();
  } else {
    np = (uint32_t)npred[v];
  }
  for (int32_t base = 0;; base += kWinBits) {
    if (!INIT || base > 0) {
      if (base > 0) {
        for (int i = lane; i < kWin32; i += 64) ww[i] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      for (uint32_t t = lane; t < np; t += 64) {
        const int32_t c = col0[pslot[(uint64_t)r.x + t]];
        if (base == 0 && c == cv) ++k;
        const uint32_t cw = (uint32_t)(c - base);
        if (cw < (uint32_t)kWinBits) atomicOr(&ww[cw >> 5], 1u << (cw & 31u));
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int32_t found = INT_MAX;
    for (int i = lane; i < kWin32; i += 64)
      if (ww[i] != ~0u && found == INT_MAX) found = i * 32 + __ffs(~ww[i]) - 1;
    for (int off = 32; off > 0; off >>= 1) found = min(found, __shfl_xor(found, off));
    if (base == 0)
      for (int off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (found != INT_MAX) {
      pm = base + found;
      cc = k;
      return;
    }
  }
}

// Rows of the histogram-row table (cap >= kSmallLen; with MID only those
// below kMidLen) for which pick(v) holds, each by a whole wave: the lanes
// test 64 table entries at once, then the wave takes the picked rows one
// after another.
template <bool MID, class Pick, class Body>
__device__ __forceinline__ void each_long_row(const LdfState& S, Pick pick, Body body) {
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  // entries strided over the waves: the table lists the largest rows
  // together, which one wave would otherwise take in a row
  for (uint32_t base = 0; base < S.n_hrows; base += nwaves * 64u) {
    const uint32_t i = base + (uint32_t)lane * nwaves + gwave;
    int32_t v = -1;
    if (i < S.n_hrows) {
      v = S.hrowv[i];
      if ((MID && S.G.cap[v] >= kMidLen) || !pick(v)) v = -1;
    }
    unsigned long long act = __ballot(v >= 0);
    while (act) {
      const int src = __ffsll((long long)act) - 1;
      act &= act - 1ull;
      body(__shfl(v, src));
    }
  }
}

// Hub chunks whose hub satisfies pick(v), each by the whole workgroup: the
// threads test kBlock table entries at once, then the workgroup takes the
// picked chunks one after another.
template <class Pick, class Body>
__device__ __forceinline__ void each_hub_chunk(const LdfState& S, Pick pick, Body body) {
  __shared__ int2 list[kBlock];
  __shared__ uint32_t nlist;
  // entries strided over the workgroups: a hub's chunks are listed together
  for (uint32_t base = 0; base < S.n_hch; base += gridDim.x * (uint32_t)kBlock) {
    __syncthreads();
    if (threadIdx.x == 0) nlist = 0u;
    __syncthreads();
    const uint32_t i = base + threadIdx.x * gridDim.x + blockIdx.x;
    if (i < S.n_hch) {
      const int2 e = S.hch[i];
      if (pick(e.x)) list[atomicAdd(&nlist, 1u)] = e;
    }
    __syncthreads();
    const uint32_t m = nlist;
    for (uint32_t q = 0; q < m; ++q) body(list[q]);
  }
}

// Adds the workgroup's sums of (a, b, c) into round slot set `slot`.
__device__ __forceinline__ void ldf_sums(const LdfState& S, int slot, uint32_t a, uint32_t b,
                                         uint32_t c) {
  __shared__ uint32_t red[kBlock / 64][3];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
    c += __shfl_xor(c, off);
  }
  const int wib = threadIdx.x >> 6;
  __syncthreads();
  if (lane_id() == 0) {
    red[wib][0] = a;
    red[wib][1] = b;
    red[wib][2] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s[3] = {0u, 0u, 0u};
    for (int q = 0; q < kBlock / 64; ++q)
      for (int i = 0; i < 3; ++i) s[i] += red[q][i];
    unsigned long long* acc = S.acc + ((size_t)slot * kLdfAcc + blockIdx.x % kLdfAcc) * 4;
    for (int i = 0; i < 3; ++i)
      if (s[i]) atomicAdd(&acc[i], (unsigned long long)s[i]);
  }
}

// ---------------------------------------------------------------------------
// batch set-up: every owned vertex's cc and cached mex
__global__ __launch_bounds__(kBlock) void k_ldf_init_cc(LdfState S) {
  __shared__ uint32_t win[kBlock / 64][kWin32];
  const GraphView& G = S.G;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < G.n_owned;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (G.hidx[v] >= 0) continue;
    int32_t pm, cc;
    const int32_t cv = S.c0[v];
    short_pred_init(G, S.pdir, (int32_t)v, cv, pm, cc);
    S.V[v] = LdfV{cv, -1, pm, 0};
    S.cc[v] = cc;
  }
  uint32_t* ww = win[threadIdx.x >> 6];
  each_long_row<false>(S, [](int32_t) { return true; }, [&](int32_t v) {
    int32_t pm, cc;
    const int32_t cv = S.c0[v];
    wave_pred_scan<true>(G, S.pdir, S.pslot, S.npred, S.c0, v, cv, ww, pm, cc);
    if (lane_id() == 0) {
      S.V[v] = LdfV{cv, -1, pm, 0};
      S.cc[v] = cc;
    }
  });
}

// L_0 = the marked set (round -1's C): its flags, and its rows as D of round
// -1 (the row pass then counts them into K).
__global__ void k_ldf_seed(LdfState S, Frontier L0) {
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = t0; i < L0.cnt[0]; i += stride) {
    S.F[0][L0.s[i]] = 1;
    S.fl[L0.s[i]] = kFlD;
  }
  for (int64_t i = t0; i < L0.cnt[1]; i += stride) {
    S.F[0][L0.m[i]] = 1;
    S.fl[L0.m[i]] = kFlD;
  }
  for (int64_t i = t0; i < L0.cnt[2]; i += stride) {
    const int4 h = L0.h[i];
    if (h.y != 0) continue;
    S.F[0][h.x] = 1;
    S.fl[h.x] = kFlD;
  }
}

// ---------------------------------------------------------------------------
// The short rows picked among 64 vertices (one per lane; `pick` set on the
// lanes whose row is taken), each by an 8-lane group: the lanes list the
// picked lanes in LDS, then group g takes entries g, g + 8, ... and its lane
// `sub` the slots sub, sub + 8, ... of the row.  body(src lane, sub) runs on
// all 64 lanes for every step of 8 entries (src = -1 on lanes without one),
// so it may shuffle.
template <class Body>
__device__ __forceinline__ void short_rows_by_groups(bool pick, int32_t* lst, Body body) {
  const int lane = lane_id();
  const unsigned long long b = __ballot(pick);
  if (pick) lst[__popcll(b & ((1ull << lane) - 1ull))] = lane;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const uint32_t n = (uint32_t)__popcll(b);
  const int grp = lane >> 3, sub = lane & 7;
  for (uint32_t t = 0; t < n; t += 8) {
    const bool mine = t + (uint32_t)grp < n;
    body(mine ? lst[t + grp] : -1, sub);
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// round r, launch 1: members, their mex (short rows), D and X flags.  A wave
// takes 64 vertices at a time; the stale mexes of their short member rows are
// taken by 8-lane groups (a lane per vertex would leave most lanes idle).
__global__ __launch_bounds__(kBlock) void k_ldf_members(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  __shared__ int32_t lists[kBlock / 64][64];
  __shared__ int32_t results[kBlock / 64][64];
  const GraphView& G = S.G;
  const uint8_t* Lf = S.F[r & 1];
  uint8_t* Cf = S.F[(r + 1) & 1];
  const int lane = lane_id();
  int32_t* lst = lists[threadIdx.x >> 6];
  int32_t* res = results[threadIdx.x >> 6];
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t nl = 0, nc = 0, nx = 0;
  for (int64_t base = gwave * 64; base < G.n_owned; base += nwaves * 64) {
    const int64_t v = base + lane;
    const bool valid = v < G.n_owned;
    bool member = false, stale_short = false;
    uint8_t f = 0;
    int32_t m = -1, cv = 0;
    LdfV rec{0, -1, -1, 0};
    int32_t nc0 = -1;
    if (valid) {
      rec = S.V[v];
      nc0 = rec.nc;
      const bool inl = Lf[v] != 0;
      member = S.K[v] > 0 || (inl && S.cc[v] > 0);
      Cf[v] = member ? 1 : 0;
      f = inl != member ? kFlD : 0;
      nl += inl ? 1u : 0u;
      nc += member ? 1u : 0u;
      if (member) {
        m = rec.pm;
        cv = rec.c;
        if (m < 0) {
          if (G.hidx[v] < 0) stale_short = true;
          else f |= kFlPm;   // long row: the next launch
        }
      }
    }
    short_rows_by_groups(stale_short, lst, [&](int src, int sub) {
      const int32_t x = __shfl((int32_t)v, src < 0 ? lane : src);
      uint32_t lo = 0u, hi = 0u;
      if (src >= 0) {
        // 4 slots per lane per step: their predecessor bytes and slot colours
        // in one round trip, a hub predecessor's colour in a second
        const uint2 rw = G.row[x];
        for (uint32_t j0 = (uint32_t)sub; j0 < rw.y; j0 += 32) {
          uint8_t pd[4];
          int32_t c[4], w[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t j = j0 + 8u * q;
            const uint64_t sl = (uint64_t)rw.x + j;
            pd[q] = j < rw.y ? S.pdir[sl] : kPdNone;
            w[q] = j < rw.y ? G.col[sl] : 0;
          }
          // a predecessor's colour from the colour array
#pragma unroll
          for (int q = 0; q < 4; ++q) c[q] = pd[q] != kPdNone ? S.c0[w[q]] : 0;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (pd[q] != kPdNone) add_colour_bit((uint32_t)c[q], lo, hi);
        }
      }
      for (int off = 1; off < 8; off <<= 1) {
        lo |= __shfl_xor(lo, off);
        hi |= __shfl_xor(hi, off);
      }
      if (src >= 0 && sub == 0) res[src] = mex64(lo, hi);
    });
    if (stale_short) {
      m = res[lane];   // fewer than 64 predecessors: below 64
      rec.pm = m;
    }
    rec.nc = -1;
    if (member && m >= 0 && m != cv) {
      rec.nc = m;
      S.cc[v] = 0;   // its changed predecessors re-count it in the row pass
      f |= kFlX;
      ++nx;
    }
    if (valid) {
      // the record is rewritten only when it moved: a new mex, or a new
      // colour (set, or cleared after last round's)
      if (stale_short || rec.nc != nc0) S.V[v] = rec;
      S.fl[v] = f;
    }
  }
  ldf_sums(S, r & 1, nl, nc, nx);
}

// launch 2: stale mex of mid rows (a wave per row) and hub rows (a
// workgroup per chunk), and their colour changes
__global__ __launch_bounds__(kBlock) void k_ldf_long_mex(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  __shared__ uint32_t win[kBlock / 64][kWin32];
  const GraphView& G = S.G;
  uint32_t* ww = win[threadIdx.x >> 6];
  uint32_t nx = 0;
  auto stale = [&](int32_t v) { return (S.fl[v] & kFlPm) != 0; };
  each_long_row<false>(S, stale, [&](int32_t v) {
    int32_t m, cc;
    const int32_t cv = S.c0[v];
    wave_pred_scan<false>(G, S.pdir, S.pslot, S.npred, S.c0, v, cv, ww, m, cc);
    if (lane_id() == 0) {
      S.V[v].pm = m;
      if (m != cv) {
        S.V[v].nc = m;
        S.cc[v] = 0;
        S.fl[v] = S.fl[v] | kFlX;
        ++nx;
      }
    }
  });
  ldf_sums(S, r & 1, 0u, 0u, nx);
}

// x's colour c (was old) as owned neighbour w keeps it: the slot of x in w's
// row (position t), and w's histogram and cached mex over all its neighbours,
// as propagate_slots keeps them.
__device__ __forceinline__ void nbr_colour(const GraphView& G, int32_t w, int32_t t, int32_t c,
                                           int32_t old) {
  if (w >= G.n_owned) return;
  const int32_t h = G.hidx[w];
  const int32_t m = G.mexc[w];
  G.ccol[(uint64_t)G.row[w].x + (uint32_t)t] = c;
  uint32_t before = 2u;
  if (h >= 0) {
    uint32_t* hr = G.hist + (uint64_t)h * kHistW;
    if (old >= 0 && old < kHistW) before = atomicSub(&hr[old], 1u);
    if (c < kHistW) atomicAdd(&hr[c], 1u);
  }
  const bool stale = h >= 0 ? (old < m && (old >= kHistW || before == 1u)) || c == m
                            : old < m || c == m;
  if (m >= 0 && stale) G.mexc[w] = -1;
}

// One lane's share of the row pass of x (slots lo, lo + stride, ...), over
// x's successors y only: D adds x's +-1 to K[y]; X (colour a -> b) moves
// cc[y] and drops y's predecessor mex when the change can move it (y reads
// b from the colour array).  A cached mex m of y's predecessors' colours
// moves only if a < m (a may have been the only predecessor holding a colour
// below m) or b == m; a > m leaves every colour below m present, and b != m
// adds no colour that was missing at m.  When several predecessors of y
// change in one round, each tests the same round-start m, so m is dropped if
// any of their changes can move it.
// The loop reads no slot colour, histogram or all-neighbour mex: x's slots
// in its neighbours' rows, their histograms and all-neighbour mex caches take
// x's final colour once, after the loop (k_ldf_settle).
__device__ __forceinline__ void ldf_row_part(const LdfState& S, int32_t x, uint8_t f, int32_t ksign,
                                             int32_t a, int32_t b, uint64_t rbase, uint32_t lo,
                                             uint32_t hi, uint32_t stride) {
  const GraphView& G = S.G;
  const bool xc = (f & kFlX) != 0;
  for (uint32_t j0 = lo; j0 < hi; j0 += 4 * stride) {
    int32_t y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t j = j0 + q * stride;
      // successors only: live slots whose neighbour does not precede x
      y[q] = j < hi && S.pdir[rbase + j] == kPdNone ? G.col[rbase + j] : -1;
    }
    // the successors' records, all in flight at once
    LdfV vy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      vy[q] = LdfV{-1, -1, -1, 0};
      if (xc && y[q] >= 0) vy[q] = S.V[y[q]];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (y[q] < 0) continue;
      if (f & kFlD) atomicAdd(&S.K[y[q]], ksign);
      if (xc) {
        // y's cc counts predecessors with y's colour: its new one when y
        // changes in this round too (nc >= 0)
        const int32_t dlt = vy[q].nc >= 0 ? (b == vy[q].nc ? 1 : 0)
                                          : (b == vy[q].c ? 1 : 0) - (a == vy[q].c ? 1 : 0);
        if (dlt) atomicAdd(&S.cc[y[q]], dlt);
        const int32_t my = vy[q].pm;
        if (my >= 0 && (a < my || b == my)) S.V[y[q]].pm = -1;
      }
    }
  }
}

// launch 3: the row passes of D and X
__global__ __launch_bounds__(kBlock) void k_ldf_rows(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  const GraphView& G = S.G;
  const uint8_t* Cf = S.F[(r + 1) & 1];
  auto pass = [&](int32_t x, uint32_t lo, uint32_t hi_cap, uint32_t stride) {
    const uint8_t f = S.fl[x];
    const uint2 rw = G.row[x];
    const uint32_t hi = min(rw.y, hi_cap);
    const int32_t a = S.c0[x];
    const int32_t b = (f & kFlX) ? S.V[x].nc : a;
    ldf_row_part(S, x, f, Cf[x] ? 1 : -1, a, b, rw.x, lo, hi, stride);
  };
  {
    // short rows of D and X: per wave 64 vertices at a time, the moved ones
    // taken by 8-lane groups
    __shared__ int32_t lists[kBlock / 64][64];
    int32_t* lst = lists[threadIdx.x >> 6];
    const int lane = lane_id();
    const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = gwave * 64; base < G.n_owned; base += nwaves * 64) {
      const int64_t v = base + lane;
      const bool pick = v < G.n_owned && (S.fl[v] & (kFlD | kFlX)) && G.hidx[v] < 0;
      short_rows_by_groups(pick, lst, [&](int src, int sub) {
        if (src >= 0) pass((int32_t)(base + src), (uint32_t)sub, 0xffffffffu, 8u);
      });
    }
  }
  auto moved = [&](int32_t v) { return (S.fl[v] & (kFlD | kFlX)) != 0; };
  each_long_row<true>(S, moved, [&](int32_t v) { pass(v, (uint32_t)lane_id(), 0xffffffffu, 64u); });
  each_hub_chunk(S, moved, [&](int2 e) {
    const uint32_t lo = (uint32_t)e.y * kChunk;
    pass(e.x, lo + threadIdx.x, lo + kChunk, kBlock);
  });
}

// after the loop: every vertex whose colour differs from its colour at the
// loop's start writes its final colour into its slots in all its neighbours'
// rows and moves their histograms and all-neighbour mex caches from the
// start colour to it
__global__ __launch_bounds__(kBlock) void k_ldf_settle(LdfState S) {
  const GraphView& G = S.G;
  auto pass = [&](int32_t x, uint32_t lo, uint32_t hi_cap, uint32_t stride) {
    const uint2 rw = G.row[x];
    const uint32_t hi = min(rw.y, hi_cap);
    const int32_t a = S.corig[x], b = S.c0[x];
    for (uint32_t j = lo; j < hi; j += stride) {
      const uint64_t sl = (uint64_t)rw.x + j;
      const int32_t y = G.col[sl];
      if (y >= 0) nbr_colour(G, y, G.twin[sl], b, a);
    }
  };
  auto changed = [&](int32_t v) { return S.corig[v] >= 0 && S.corig[v] != S.c0[v]; };
  {
    // short rows: per wave 64 vertices at a time, the changed ones taken by
    // 8-lane groups
    __shared__ int32_t lists[kBlock / 64][64];
    int32_t* lst = lists[threadIdx.x >> 6];
    const int lane = lane_id();
    const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = gwave * 64; base < G.n_owned; base += nwaves * 64) {
      const int64_t v = base + lane;
      const bool pick = v < G.n_owned && G.hidx[v] < 0 && changed((int32_t)v);
      short_rows_by_groups(pick, lst, [&](int src, int sub) {
        if (src >= 0) pass((int32_t)(base + src), (uint32_t)sub, 0xffffffffu, 8u);
      });
    }
  }
  each_long_row<true>(S, changed, [&](int32_t v) { pass(v, (uint32_t)lane_id(), 0xffffffffu, 64u); });
  each_hub_chunk(S, changed, [&](int2 e) {
    const uint32_t lo = (uint32_t)e.y * kChunk;
    pass(e.x, lo + threadIdx.x, lo + kChunk, kBlock);
  });
}

// launch 4: commit the new colours; round r's counts into the trace; the
// loop ends after a round without a colour change
__global__ __launch_bounds__(kBlock) void k_ldf_commit(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  unsigned long long lcv = 0ull, lcs = 0ull, ms = 0ull;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < S.G.n_owned;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (S.fl[v] & kFlX) {
      const int32_t nc = S.V[v].nc;
      if (S.corig[v] < 0) {   // v's first change in this loop
        S.corig[v] = S.c0[v];
        ms += S.G.row[v].y;
      }
      S.c0[v] = nc;
      S.c1[v] = nc;
      S.V[v].c = nc;
    }
    if (S.prof && (S.F[0][v] | S.F[1][v])) {   // v in L_r u C_r
      ++lcv;
      lcs += S.G.row[v].y;
    }
  }
  for (int off = 32; off > 0; off >>= 1) ms += __shfl_xor(ms, off);
  if (lane_id() == 0 && ms) atomicAdd(&S.ctl->moved_slots, ms);
  if (S.prof) {
    for (int off = 32; off > 0; off >>= 1) {
      lcv += __shfl_xor(lcv, off);
      lcs += __shfl_xor(lcs, off);
    }
    if (lane_id() == 0 && lcv) {
      atomicAdd(&S.terms[0], lcv);
      atomicAdd(&S.terms[1], lcs);
    }
  }
  if (blockIdx.x == 0) {
    unsigned long long* nxt = S.acc + (size_t)((r + 1) & 1) * kLdfAcc * 4;
    if (threadIdx.x < 64) {
      unsigned long long t[3] = {0ull, 0ull, 0ull};
      if (threadIdx.x < kLdfAcc) {
        const unsigned long long* a = S.acc + ((size_t)(r & 1) * kLdfAcc + threadIdx.x) * 4;
        for (int i = 0; i < 3; ++i) t[i] = a[i];
      }
      for (int i = 0; i < 3; ++i)
        for (int off = 32; off > 0; off >>= 1) t[i] += __shfl_xor(t[i], off);
      if (threadIdx.x == 0) {
        S.trace[r] = make_int4((int32_t)t[0], (int32_t)t[1], (int32_t)t[2], 0);
        if (t[2] == 0ull) S.ctl->done_round = r;
      }
    }
    for (int i = threadIdx.x; i < kLdfAcc * 4; i += blockDim.x) nxt[i] = 0ull;
  }
}

// ============================================================================
// host side
struct LdfHost {
  LdfState S{};
  int32_t n = 0;
  int32_t trace_cap = 0;
  LdfCtl* hctl = nullptr;
  uint8_t* flags = nullptr;   // F[0], F[1], fl
  uint64_t pdir_cap = 0;      // slots pdir and pslot hold
  hipEvent_t ev[2] = {nullptr, nullptr};   // around a block of rounds (profiling)
};

static LdfHost* ldf_state(Graph* g) { return static_cast<LdfHost*>(g->ldf); }

void ldf_free(Graph* g) {
  LdfHost* s = ldf_state(g);
  if (!s) return;
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(s->S.K); f(s->S.V); f(s->S.cc); f(s->S.corig); f(s->flags); f(s->S.acc); f(s->S.trace);
  f(s->S.ctl); f(s->S.pdir); f(s->S.pslot); f(s->S.npred); f(s->S.terms);
  if (s->hctl) (void)hipHostFree(s->hctl);
  for (hipEvent_t e : s->ev)
    if (e) (void)hipEventDestroy(e);
  delete s;
  g->ldf = nullptr;
}

static int ldf_alloc(Graph* g, LdfHost* s) {
  const size_t n = (size_t)g->n + 1;
  LdfState& S = s->S;
  DC_HIP(hipMalloc(&S.K, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.V, sizeof(LdfV) * n));
  DC_HIP(hipMalloc(&S.cc, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.corig, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.npred, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&s->flags, 3 * n));
  S.F[0] = s->flags;
  S.F[1] = s->flags + n;
  S.fl = s->flags + 2 * n;
  DC_HIP(hipMalloc(&S.acc, sizeof(unsigned long long) * 2 * kLdfAcc * 4));
  DC_HIP(hipMalloc(&S.ctl, sizeof(LdfCtl)));
  DC_HIP(hipMalloc(&S.terms, sizeof(unsigned long long) * 2));
  DC_HIP(hipHostMalloc(&s->hctl, sizeof(LdfCtl), hipHostMallocDefault));
  DC_HIP(hipEventCreate(&s->ev[0]));
  DC_HIP(hipEventCreate(&s->ev[1]));
  s->n = g->n;
  return DC_OK;
}

static int ldf_trace(Graph* g, LdfHost* s, int32_t need) {
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

// host blocks of rounds: 0-7, then 32 at a time (rounds after the last one
// return at once)
static int32_t ldf_block_end(int32_t r) { return r < 8 ? 8 : r + 32; }

int loop_ldf(Graph* g, Frontier L0, LoopResult* res) {
  LdfHost* s = ldf_state(g);
  if (!s) {
    s = new LdfHost();
    g->ldf = s;
    int rc = ldf_alloc(g, s);
    if (rc) {
      ldf_free(g);   // a later call starts over instead of using a partial state
      return rc;
    }
  }
  int rc = ldf_trace(g, s, 256);
  if (rc) return rc;
  LdfState& S = s->S;
  S.G = g->view();
  S.c0 = g->color[g->cur];
  S.c1 = g->color[g->cur ^ 1];
  S.hrowv = g->hrowv;
  S.n_hrows = g->n_hrows;
  S.hch = g->hch;
  S.n_hch = g->n_hch;
  if (s->pdir_cap < g->total_cap) {
    if (S.pdir) DC_HIP(hipFree(S.pdir));
    if (S.pslot) DC_HIP(hipFree(S.pslot));
    S.pdir = nullptr;
    S.pslot = nullptr;
    s->pdir_cap = 0;
    DC_HIP(hipMalloc(&S.pdir, g->total_cap));
    DC_HIP(hipMalloc(&S.pslot, sizeof(uint32_t) * g->total_cap));
    s->pdir_cap = g->total_cap;
  }
  const size_t n = (size_t)g->n + 1;
  DC_HIP(hipMemsetAsync(S.K, 0, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(S.corig, 0xff, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(s->flags, 0, 3 * n, g->stream));
  DC_HIP(hipMemsetAsync(S.acc, 0, sizeof(unsigned long long) * 2 * kLdfAcc * 4, g->stream));
  S.prof = g->profiling ? 1 : 0;
  if (S.prof) DC_HIP(hipMemsetAsync(S.terms, 0, sizeof(unsigned long long) * 2, g->stream));
  *s->hctl = LdfCtl{};
  s->hctl->done_round = INT_MAX;
  DC_HIP(hipMemcpyAsync(S.ctl, s->hctl, sizeof(LdfCtl), hipMemcpyHostToDevice, g->stream));
  const dim3 blk(kBlock), grid(kLdfBlocks);
  hipLaunchKernelGGL(k_ldf_init_cc, grid, blk, 0, g->stream, S);
  hipLaunchKernelGGL(k_ldf_seed, dim3(256), blk, 0, g->stream, S, L0);
  // round -1: C = L_0, so the row pass counts L_0's rows into K (+1)
  hipLaunchKernelGGL(k_ldf_rows, grid, blk, 0, g->stream, S, -1);
  DC_HIP(hipGetLastError());
  int64_t launches = 3;
  const int32_t max_rounds = g->opt.max_rounds;
  int32_t rounds = -1, r = 0;
  double loop_ms = 0.0;
  while (r < max_rounds && rounds < 0) {
    const int32_t nb = std::min(ldf_block_end(r), max_rounds) - r;
    if ((rc = ldf_trace(g, s, r + nb))) return rc;
    if (S.prof) DC_HIP(hipEventRecord(s->ev[0], g->stream));
    for (int32_t i = 0; i < nb; ++i) {
      hipLaunchKernelGGL(k_ldf_members, grid, blk, 0, g->stream, S, r + i);
      hipLaunchKernelGGL(k_ldf_long_mex, grid, blk, 0, g->stream, S, r + i);
      hipLaunchKernelGGL(k_ldf_rows, grid, blk, 0, g->stream, S, r + i);
      hipLaunchKernelGGL(k_ldf_commit, grid, blk, 0, g->stream, S, r + i);
    }
    DC_HIP(hipGetLastError());
    launches += 4 * nb;
    if (S.prof) DC_HIP(hipEventRecord(s->ev[1], g->stream));
    DC_HIP(hipMemcpyAsync(s->hctl, S.ctl, sizeof(LdfCtl), hipMemcpyDeviceToHost, g->stream));
    DC_HIP(hipStreamSynchronize(g->stream));
    if (S.prof) {
      float ms = 0.f;
      DC_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
      loop_ms += ms;
    }
    r += nb;
    if (s->hctl->done_round != INT_MAX) rounds = s->hctl->done_round + 1;
  }
  // the final colours into the neighbours' slots, histograms and
  // all-neighbour mexes: by the changed vertices' rows, or by a rewrite of
  // every row's slot colours once those rows hold more than 1/kSettleDiv of
  // the laid-out slots (a row pass costs several scattered accesses per slot,
  // the rewrite one gather per slot)
  res->moved_slots = s->hctl->moved_slots;
  res->slots_rewritten = res->moved_slots > std::max<uint64_t>(g->rows_end / kSettleDiv, kSettleMin);
  if (res->slots_rewritten) {
    if ((rc = rebuild_ccol(g))) return rc;
    launches += 3;
  } else {
    hipLaunchKernelGGL(k_ldf_settle, grid, blk, 0, g->stream, S);
    DC_HIP(hipGetLastError());
    ++launches;
  }
  if (rounds < 0) {
    DC_HIP(hipStreamSynchronize(g->stream));
    set_error("recolor loop hit max_rounds");
    return DC_ERR_ROUNDS;
  }
  std::vector<int4> tr((size_t)rounds);
  DC_HIP(hipMemcpyAsync(tr.data(), S.trace, sizeof(int4) * (size_t)rounds, hipMemcpyDeviceToHost,
                        g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  res->rounds = rounds;
  for (int32_t i = 0; i < rounds; ++i) {
    g->trace.insert(g->trace.end(),
                    {(int64_t)10, (int64_t)tr[i].x, (int64_t)tr[i].y, (int64_t)tr[i].z});
    res->frontier_sum += (uint64_t)tr[i].x;
    res->evaluated += (uint64_t)tr[i].y;
    res->changed += (uint64_t)tr[i].z;
  }
  res->launches = launches;
  if (S.prof) {
    // the rounds' kernel time, and the reference's bytes for them: per
    // vertex of L_r u C_r its list entry, row bounds and own colour (16),
    // per row slot the neighbour ID, its colour and its priority (12), per
    // round a compaction bitmap over the vertices (n/8) and 4 per member of
    // C_r
    unsigned long long t[2];
    DC_HIP(hipMemcpy(t, S.terms, sizeof(t), hipMemcpyDeviceToHost));
    res->lc_vertices = t[0];
    res->lc_slots = t[1];
    res->kernel_ms = loop_ms;
    res->count_ms = loop_ms;
    res->count_launches = rounds;
    res->count_bytes = 16ull * t[0] + 12ull * t[1] +
                       (uint64_t)rounds * ((uint64_t)g->n_owned / 8) + 4ull * res->evaluated;
  }
  return DC_OK;
}

}  // namespace dc
