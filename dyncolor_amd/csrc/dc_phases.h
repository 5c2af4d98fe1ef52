// dc_phases.h -- device bodies of every recolor phase, shared by the
// one-launch-per-phase kernels (recolor.hip) and the persistent cooperative
// kernel (fused.hip).  Each body walks its whole input with grid-stride loops
// over the launch's workgroups and reads its item counts from device memory.
//
// Row visits are split by slot count so that no lane walks more than 8 slots
// of one row per phase: S rows (< 64 slots) by an 8-lane group, M rows (64..
// 511) by one wave64, H rows (>= 512) by one workgroup per 1024-slot chunk.
// Every lane issues 4 independent slot loads (and then 4 independent color
// gathers) per step to keep enough memory requests in flight.
#pragma once

#include <hip/hip_runtime.h>

#include "dc_device.h"
#include "dc_internal.h"

namespace dc {

enum Mode {
  M_DEL = 0,       // deleteEdge: mex of all live nbrs, write if changed (HI/gpuFunctions_undir.cuh:147-177)
  M_INS = 1,       // insEdge: unconditional (HI/gpuFunctions_undir.cuh:182-223)
  M_LDF_P = 2,     // processChangeEdge: predecessor mex, write if changed (HB/gpuFunctions_undir.cuh:222-249)
  M_LOOP_HI = 3,   // recolorNeighbor, unconditional (HI/gpuFunctions_undir.cuh:264-284)
  M_LOOP_LDF = 4,  // recolorNeighbor, predecessor mex if changed (HB/gpuFunctions_undir.cuh:295-314)
  M_STATIC_ID = 5, // greedy in ID order (dependency-counting schedule)
  M_STATIC_LDF = 6 // greedy in LDF priority order
};

template <int MODE>
struct ModeTraits {
  static constexpr bool prop = MODE == M_LDF_P || MODE == M_LOOP_LDF || MODE == M_STATIC_LDF;
  static constexpr bool stat = MODE == M_STATIC_ID || MODE == M_STATIC_LDF;
  static constexpr bool marks = MODE == M_DEL || MODE == M_INS || MODE == M_LDF_P;
  // the mex of every live neighbour: may use and refresh the per-vertex cache
  static constexpr bool cached = MODE == M_DEL || MODE == M_INS || MODE == M_LOOP_HI;
};

constexpr int kUnroll = 4;

// A hub chunk's partial window: colors 0..511 of the chunk's included
// neighbours (8 words) and whether any included colour was >= 512.
constexpr int kPartWords = 8;
constexpr int kPartStride = 16;  // u64 per chunk entry (word 8 = overflow flag)

struct RecolorArgs {
  GraphView G;
  const int32_t* cur;
  int32_t* nxt;
  int32_t* prev;
  Frontier in;        // vertices to evaluate
  Frontier out;       // marked set (pre-loop rounds) or next ready set (static)
  uint32_t* mmark;
  uint32_t mepoch;
  int32_t* pending;
  unsigned long long* partial;   // kPartStride u64 per hub chunk of `in`
  Frontier chx;                  // vertices whose colour this launch changed
  uint32_t hub_cap;
  Counters* ctr;                 // window_full flag
  uint32_t* changed;             // color changes of this launch
  unsigned long long* rslots;    // row slots read
  unsigned long long* rlive;     // live entries read
};

// Is neighbour w (priority pw) part of v's mex set under MODE?
template <int MODE>
__device__ __forceinline__ bool include_nbr(int32_t v, int32_t pv, int32_t w, int32_t pw,
                                            int32_t n_owned) {
  if constexpr (MODE == M_STATIC_ID) return w >= n_owned || w < v;
  else if constexpr (MODE == M_STATIC_LDF) return w >= n_owned || precedes(nullptr, w, pw, v, pv);
  else if constexpr (ModeTraits<MODE>::prop) return precedes(nullptr, w, pw, v, pv);
  else return true;
}

// Commit rule of each mode for vertex v with snapshot color cv and mex m.
// Returns whether v joins the output marked set.
template <int MODE>
__device__ __forceinline__ bool finish_vertex(const RecolorArgs& A, int32_t v, int32_t cv,
                                              int32_t m, uint32_t& changed, bool& ch) {
  ch = m != cv;
  changed += ch ? 1u : 0u;
  if constexpr (MODE == M_DEL) {
    if (ch) {
      A.nxt[v] = m;
      A.prev[v] = cv;
      return claim(A.mmark, v, A.mepoch);
    }
    return false;
  } else if constexpr (MODE == M_INS) {
    A.nxt[v] = m;
    A.prev[v] = cv;
    return claim(A.mmark, v, A.mepoch);
  } else if constexpr (MODE == M_LDF_P) {
    if (ch) {
      A.nxt[v] = m;
      return claim(A.mmark, v, A.mepoch);
    }
    return false;
  } else if constexpr (MODE == M_LOOP_HI) {
    A.nxt[v] = m;
    A.prev[v] = cv;
    return false;
  } else if constexpr (MODE == M_LOOP_LDF) {
    if (ch) A.nxt[v] = m;
    return false;
  } else {
    A.nxt[v] = m;  // static: nxt aliases cur, predecessors are final
    return false;
  }
}

// Static modes: a neighbour that v precedes loses one pending predecessor;
// the last one to drop it to zero schedules it.
template <int MODE>
__device__ __forceinline__ bool release_nbr(const RecolorArgs& A, int32_t v, int32_t pv, int32_t w,
                                            int32_t pw) {
  if constexpr (ModeTraits<MODE>::stat) {
    if (w >= 0 && w < A.G.n_owned && !include_nbr<MODE>(v, pv, w, pw, A.G.n_owned))
      return atomicSub(&A.pending[w], 1) == 1;
  }
  return false;
}

// Colors 0..255 of a row visit are collected in four per-lane register words
// (combined across lanes with xor-shuffles); colors in [256, lim) go straight
// to an LDS window `hiw` indexed by word (c >> 6).
struct RegMask {
  unsigned long long m0, m1, m2, m3;
  bool over;  // a colour >= lim was seen
  __device__ __forceinline__ void clear() {
    m0 = m1 = m2 = m3 = 0ull;
    over = false;
  }
  __device__ __forceinline__ void set(int32_t c, unsigned long long* hiw, int32_t lim) {
    if (c < 0) return;
    if (c < 256) {
      const unsigned long long b = 1ull << (c & 63);
      const int wi = c >> 6;
      m0 |= wi == 0 ? b : 0ull;
      m1 |= wi == 1 ? b : 0ull;
      m2 |= wi == 2 ? b : 0ull;
      m3 |= wi == 3 ? b : 0ull;
    } else if (c < lim) {
      atomicOr(&hiw[c >> 6], 1ull << (c & 63));
    } else {
      over = true;
    }
  }
  // OR across aligned groups of `width` lanes (width a power of two <= 64)
  __device__ __forceinline__ void reduce(int width) {
    for (int off = width >> 1; off > 0; off >>= 1) {
      m0 |= __shfl_xor(m0, off);
      m1 |= __shfl_xor(m1, off);
      m2 |= __shfl_xor(m2, off);
      m3 |= __shfl_xor(m3, off);
    }
  }
  // least color < 256 absent, or -1 if 0..255 are all present
  __device__ __forceinline__ int32_t first_zero() const {
    if (m0 != ~0ull) return __ffsll((long long)~m0) - 1;
    if (m1 != ~0ull) return 64 + __ffsll((long long)~m1) - 1;
    if (m2 != ~0ull) return 128 + __ffsll((long long)~m2) - 1;
    if (m3 != ~0ull) return 192 + __ffsll((long long)~m3) - 1;
    return -1;
  }
};

// One lane's share of a row visit: slots lo, lo+stride, ... below hi, four
// at a time.  Collects the mex set into mk; in the static modes also releases
// successors and schedules the ones that become ready.
template <int MODE>
__device__ __forceinline__ void mex_slots(const RecolorArgs& A, int32_t v, int32_t pv, uint64_t rbase,
                                          uint32_t lo, uint32_t hi, uint32_t stride, RegMask& mk,
                                          unsigned long long* hiw, int32_t lim,
                                          unsigned long long& slots, unsigned long long& live) {
  const GraphView& G = A.G;
  // independent-rule modes stream ccol only; ldf modes also need the
  // neighbour's ID for its priority; the static modes run before ccol is
  // valid and gather colours.
  constexpr bool kCcolOnly = !ModeTraits<MODE>::prop && !ModeTraits<MODE>::stat;
  for (uint32_t j = lo; j < hi; j += kUnroll * stride) {
    int32_t w[kUnroll], c[kUnroll], p[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      const uint32_t jj = j + k * stride;
      const bool in = jj < hi;
      if constexpr (kCcolOnly) {
        c[k] = in ? G.ccol[rbase + jj] : -2;
        w[k] = c[k] >= 0 ? 0 : c[k];  // live marker only
      } else if constexpr (ModeTraits<MODE>::stat) {
        w[k] = in ? G.col[rbase + jj] : -2;
      } else {
        w[k] = in ? G.col[rbase + jj] : -2;
        c[k] = in ? G.ccol[rbase + jj] : -1;
      }
    }
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      p[k] = 0;
      if constexpr (ModeTraits<MODE>::stat) {
        c[k] = w[k] >= 0 ? A.cur[w[k]] : -1;
      }
      if constexpr (ModeTraits<MODE>::prop) {
        if (w[k] >= 0) p[k] = G.prop[w[k]];
      }
    }
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      slots += w[k] != -2 ? 1ull : 0ull;
      if (w[k] >= 0) {
        ++live;
        if (include_nbr<MODE>(v, pv, w[k], p[k], G.n_owned)) mk.set(c[k], hiw, lim);
      }
    }
    if constexpr (ModeTraits<MODE>::stat) {
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) {
        const bool ready = release_nbr<MODE>(A, v, pv, w[k], p[k]);
        frontier_push(ready, w[k], G, A.out, A.hub_cap);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// recolor: evaluate every vertex of A.in
template <int MODE>
__device__ __forceinline__ void recolor_body(const RecolorArgs& A) {
  __shared__ unsigned long long wwin[kBlock / 64][8];   // M rows: colors 256..511
  __shared__ unsigned long long bwin[kPartWords + 1];   // H chunk: colors 0..511 + overflow
  const int lane = lane_id();
  const int wib = threadIdx.x >> 6;
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const GraphView& G = A.G;
  const uint32_t nS = A.in.cnt[0], nM = A.in.cnt[1], nH = A.in.cnt[2];
  uint32_t changed = 0;
  unsigned long long slots = 0, live = 0;

  // ---- H rows: one workgroup per kChunk-slot chunk.  The chunk stores its
  //      partial window; hub_finalize_body (a later launch or phase) merges a
  //      hub's chunks and takes the mex, so no workgroup waits on another.
  for (uint32_t ci = blockIdx.x; ci < nH; ci += gridDim.x) {
    const int4 it = A.in.h[ci];
    const int32_t v = it.x;
    const uint32_t c = (uint32_t)it.y;
    // cached modes take a hub's mex from its cache or histogram (finalize)
    if (ModeTraits<MODE>::cached) break;
    const uint2 r = G.row[v];
    const int32_t pv = ModeTraits<MODE>::prop ? G.prop[v] : 0;
    if (threadIdx.x <= kPartWords) bwin[threadIdx.x] = 0ull;
    __syncthreads();
    const uint32_t lo = c * kChunk;
    const uint32_t hi = min(r.y, lo + kChunk);
    RegMask mk;
    mk.clear();
    // colors >= 512 land in bwin[8] (any bit there = overflow)
    mex_slots<MODE>(A, v, pv, r.x, lo + threadIdx.x, hi, kBlock, mk, bwin, 512, slots, live);
    const bool over = __ballot(mk.over) != 0ull;
    mk.reduce(64);
    if (lane == 0) {
      if (mk.m0) atomicOr(&bwin[0], mk.m0);
      if (mk.m1) atomicOr(&bwin[1], mk.m1);
      if (mk.m2) atomicOr(&bwin[2], mk.m2);
      if (mk.m3) atomicOr(&bwin[3], mk.m3);
      if (over) atomicOr(&bwin[kPartWords], 1ull);
    }
    __syncthreads();
    if (threadIdx.x <= kPartWords) A.partial[(uint64_t)ci * kPartStride + threadIdx.x] = bwin[threadIdx.x];
    __syncthreads();
  }

  // ---- M rows: one wave per vertex (mex <= slots <= 511)
  unsigned long long* hw = wwin[wib];
  for (int64_t i = gwave; i < nM; i += nwaves) {
    const int32_t v = A.in.m[i];
    const uint2 r = G.row[v];
    const int32_t pv = ModeTraits<MODE>::prop ? G.prop[v] : 0;
    int32_t m = ModeTraits<MODE>::cached ? G.mexc[v] : -1;
    if (m < 0) {
      if (lane < 8) hw[lane] = 0ull;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      RegMask mk;
      mk.clear();
      mex_slots<MODE>(A, v, pv, r.x, lane, r.y, 64, mk, hw, 512, slots, live);
      mk.reduce(64);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      m = mk.first_zero();
      if (m < 0) {
        const unsigned long long x = lane >= 4 && lane < 8 ? hw[lane] : 0ull;
        const unsigned long long b = __ballot(lane >= 4 && lane < 8 && x != ~0ull);
        const int wi = __ffsll((long long)b) - 1;
        const unsigned long long word = __shfl(x, wi < 0 ? 0 : wi);
        m = wi * 64 + __ffsll((long long)~word) - 1;
      }
      if (ModeTraits<MODE>::cached && lane == 0) G.mexc[v] = m;
    }
    bool push = false, chv = false;
    if (lane == 0) push = finish_vertex<MODE>(A, v, A.cur[v], m, changed, chv);
    frontier_push(push, v, G, A.out, A.hub_cap);
    if constexpr (!ModeTraits<MODE>::stat) frontier_push(chv, v, G, A.chx, A.hub_cap);
  }

  // ---- S rows: 8-lane groups, one u64 mask (mex <= slots < 64)
  {
    const int grp = lane >> 3, sub = lane & 7;
    for (int64_t b = gwave * 8; b < nS; b += nwaves * 8) {
      const int64_t idx = b + grp;
      const bool act = idx < nS;
      const int32_t v = act ? A.in.s[idx] : 0;
      const uint2 r = act ? G.row[v] : make_uint2(0u, 0u);
      const int32_t pv = (act && ModeTraits<MODE>::prop) ? G.prop[v] : 0;
      const int32_t cm = (act && ModeTraits<MODE>::cached) ? G.mexc[v] : -1;
      RegMask mk;
      mk.clear();
      mex_slots<MODE>(A, v, pv, r.x, sub, cm >= 0 ? 0u : r.y, kGroup, mk, nullptr, 0, slots, live);
      unsigned long long mask = mk.m0;
      mask |= __shfl_xor(mask, 1);
      mask |= __shfl_xor(mask, 2);
      mask |= __shfl_xor(mask, 4);
      bool push = false, chv = false;
      if (act && sub == 0) {
        const int32_t m = cm >= 0 ? cm : __ffsll((long long)~mask) - 1;
        if (ModeTraits<MODE>::cached && cm < 0) G.mexc[v] = m;
        push = finish_vertex<MODE>(A, v, A.cur[v], m, changed, chv);
      }
      frontier_push(push, v, G, A.out, A.hub_cap);
      if constexpr (!ModeTraits<MODE>::stat) frontier_push(chv, v, G, A.chx, A.hub_cap);
    }
  }
  wave_add_u32(changed, A.changed);
  wave_add_u64(slots, A.rslots);
  wave_add_u64(live, A.rlive);
}

// Mex of a hub from its colour histogram, by one wave (lane l looks at colours
// 8l..8l+7); -1 when colours 0..kHistW-1 are all present.
__device__ __forceinline__ int32_t hist_mex(const uint32_t* hrow) {
  const int lane = lane_id();
  const uint4* p = reinterpret_cast<const uint4*>(hrow) + 2 * lane;
  const uint4 a = p[0], b = p[1];
  const int idx = a.x == 0u ? 0 : a.y == 0u ? 1 : a.z == 0u ? 2 : a.w == 0u ? 3
                : b.x == 0u ? 4 : b.y == 0u ? 5 : b.z == 0u ? 6 : b.w == 0u ? 7 : 8;
  const unsigned long long z = __ballot(idx < 8);
  if (z == 0ull) return -1;
  const int l = __ffsll((long long)z) - 1;
  return l * 8 + __shfl(idx, l);
}

// Exact mex of v's included neighbours by one wave, scanning the whole row in
// windows of kWinBits colours (LDS); used when a hub's colours 0..511 are all
// taken and some neighbour colour is larger.
template <int MODE>
__device__ inline int32_t wave_row_mex(const RecolorArgs& A, int32_t v, unsigned long long* ww) {
  const int lane = lane_id();
  const GraphView& G = A.G;
  const uint2 r = G.row[v];
  const int32_t pv = ModeTraits<MODE>::prop ? G.prop[v] : 0;
  for (int32_t base = 0;; base += kWinBits) {
    ww[lane] = 0ull;
    ww[lane + 64] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t j = lane; j < r.y; j += 64) {
      const int32_t w = G.col[(uint64_t)r.x + j];
      if (w < 0) continue;
      const int32_t pw = ModeTraits<MODE>::prop ? G.prop[w] : 0;
      if (!include_nbr<MODE>(v, pv, w, pw, G.n_owned)) continue;
      // the static modes run before ccol is valid; every other mode reads
      // the round-start colour from the slot
      const int32_t cw = (ModeTraits<MODE>::stat ? A.cur[w] : G.ccol[(uint64_t)r.x + j]) - base;
      if (cw >= 0 && cw < kWinBits) atomicOr(&ww[cw >> 6], 1ull << (cw & 63));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const unsigned long long x0 = ww[lane], x1 = ww[lane + 64];
    const unsigned long long b0 = __ballot(x0 != ~0ull);
    if (b0) {
      const int wi = __ffsll((long long)b0) - 1;
      return base + wi * 64 + __ffsll((long long)~__shfl(x0, wi)) - 1;
    }
    const unsigned long long b1 = __ballot(x1 != ~0ull);
    if (b1) {
      const int wi = __ffsll((long long)b1) - 1;
      return base + (64 + wi) * 64 + __ffsll((long long)~__shfl(x1, wi)) - 1;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Second half of a recolor launch for H rows: one wave per hub ORs the
// partial windows of the hub's chunks (stored contiguously in A.in.h) and
// commits the mex.  Runs after every chunk of the launch finished (next kernel
// or after a grid barrier), so it needs no cross-workgroup signalling.
template <int MODE>
__device__ __forceinline__ void hub_finalize_body(const RecolorArgs& A) {
  __shared__ unsigned long long fwin[kBlock / 64][kWinWords];
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t nH = A.in.cnt[2];
  uint32_t changed = 0;
  for (int64_t ci = gwave; ci < nH; ci += nwaves) {
    const int4 it = A.in.h[ci];
    if (it.y != 0) continue;  // one wave per hub: its first chunk entry
    const int32_t v = it.x;
    const uint32_t nch = (uint32_t)it.w;
    if constexpr (ModeTraits<MODE>::cached) {
      int32_t m = A.G.mexc[v];
      if (m < 0) {
        m = hist_mex(A.G.hist + (uint64_t)A.G.hidx[v] * kHistW);
        if (m < 0) m = wave_row_mex<MODE>(A, v, fwin[threadIdx.x >> 6]);
        if (lane == 0) A.G.mexc[v] = m;
      }
      bool push = false, chv = false;
      if (lane == 0) push = finish_vertex<MODE>(A, v, A.cur[v], m, changed, chv);
      frontier_push(push, v, A.G, A.out, A.hub_cap);
      frontier_push(chv, v, A.G, A.chx, A.hub_cap);
      continue;
    }
    unsigned long long w[kPartWords], ov = 0ull;
#pragma unroll
    for (int i = 0; i < kPartWords; ++i) w[i] = 0ull;
    for (uint32_t k = lane; k < nch; k += 64) {
      const unsigned long long* p = A.partial + (uint64_t)(ci + k) * kPartStride;
#pragma unroll
      for (int i = 0; i < kPartWords; ++i) w[i] |= p[i];
      ov |= p[kPartWords];
    }
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
      for (int i = 0; i < kPartWords; ++i) w[i] |= __shfl_xor(w[i], off);
      ov |= __shfl_xor(ov, off);
    }
    int32_t m = -1;
#pragma unroll
    for (int i = kPartWords - 1; i >= 0; --i)
      if (w[i] != ~0ull) m = i * 64 + __ffsll((long long)~w[i]) - 1;
    if (m < 0) m = ov ? wave_row_mex<MODE>(A, v, fwin[threadIdx.x >> 6]) : 64 * kPartWords;
    if (ModeTraits<MODE>::cached && lane == 0) A.G.mexc[v] = m;
    bool push = false, chv = false;
    if (lane == 0) push = finish_vertex<MODE>(A, v, A.cur[v], m, changed, chv);
    frontier_push(push, v, A.G, A.out, A.hub_cap);
    if constexpr (!ModeTraits<MODE>::stat) frontier_push(chv, v, A.G, A.chx, A.hub_cap);
  }
  wave_add_u32(changed, A.changed);
}

// ----------------------------------------------------------------------------
// propagate: every vertex v of X changed colour this round; write its new
// colour newc[v] into the ccol slot that points back at v in each owned
// neighbour's row (found through v's twins).  Runs between rounds, when no
// row visit reads ccol.
// A neighbour w of a vertex that changed colour `old` -> `c`: its colour
// histogram (if it keeps one) follows, and its cached mex m becomes stale only
// if the change can move it: `old` < m lost its last copy (known from the
// histogram's count; without one, any `old` < m), or `c` == m.  When several
// neighbours of w change in the same round this stays safe: a colour below m
// that ends at count 0 had a decrement that left it at 0, and colour m can
// only be gained by a change to m.
// (v changed from colour `old` to `c`: each owned neighbour's slot cache,
// histogram and cached mex follow.)  8 slots per step in three stages, each
// stage's gathers in flight together: the slots and their twins; the
// neighbours' row start, histogram index and cached mex; the histogram
// updates, after which a cached mex that the change may have moved is
// dropped.
__device__ __forceinline__ void propagate_slots(const GraphView& G, int32_t c, int32_t old,
                                                uint64_t rbase, uint32_t lo, uint32_t hi,
                                                uint32_t stride) {
  constexpr int U = 8;
  for (uint32_t j = lo; j < hi; j += U * stride) {
    int32_t w[U], t[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t jj = j + k * stride;
      const bool in = jj < hi;
      w[k] = in ? G.col[rbase + jj] : -1;
      t[k] = in ? G.twin[rbase + jj] : 0;
    }
    uint32_t rx[U];
    int32_t h[U], m[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool own = w[k] >= 0 && w[k] < G.n_owned;
      rx[k] = 0u;
      h[k] = -1;
      m[k] = -1;
      if (own) {
        rx[k] = G.row[w[k]].x;
        h[k] = G.hidx[w[k]];
        m[k] = G.mexc[w[k]];
      }
    }
    uint32_t before[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool own = w[k] >= 0 && w[k] < G.n_owned;
      before[k] = 2u;
      if (own) {
        G.ccol[(uint64_t)rx[k] + (uint32_t)t[k]] = c;
        if (h[k] >= 0) {
          uint32_t* hr = G.hist + (uint64_t)h[k] * kHistW;
          if (old >= 0 && old < kHistW) before[k] = atomicSub(&hr[old], 1u);
          if (c < kHistW) atomicAdd(&hr[c], 1u);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const bool own = w[k] >= 0 && w[k] < G.n_owned;
      const bool stale = h[k] >= 0 ? (old < m[k] && (old >= kHistW || before[k] == 1u)) || c == m[k]
                                   : old < m[k] || c == m[k];
      if (own && m[k] >= 0 && stale) G.mexc[w[k]] = -1;
    }
  }
}

__device__ __forceinline__ void propagate_body(const GraphView& G, const Frontier& X,
                                               const int32_t* newc, const int32_t* oldc) {
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t nS = X.cnt[0], nM = X.cnt[1], nH = X.cnt[2];
  for (uint32_t ci = blockIdx.x; ci < nH; ci += gridDim.x) {
    const int4 it = X.h[ci];
    const uint2 r = G.row[it.x];
    const uint32_t lo = (uint32_t)it.y * kChunk;
    propagate_slots(G, newc[it.x], oldc[it.x], r.x, lo + threadIdx.x, min(r.y, lo + kChunk), kBlock);
  }
  for (int64_t i = gwave; i < nM; i += nwaves) {
    const int32_t v = X.m[i];
    const uint2 r = G.row[v];
    propagate_slots(G, newc[v], oldc[v], r.x, lane, r.y, 64);
  }
  const int grp = lane >> 3, sub = lane & 7;
  for (int64_t b = gwave * 8; b < nS; b += nwaves * 8) {
    const int64_t idx = b + grp;
    if (idx < nS) {
      const int32_t v = X.s[idx];
      const uint2 r = G.row[v];
      propagate_slots(G, newc[v], oldc[v], r.x, sub, r.y, kGroup);
    }
  }
}

// ----------------------------------------------------------------------------
// find: for v in L, mark the candidates of the next recolor
// (findEligibleNeighbors: HI/gpuFunctions_undir.cuh:229-260 for RULE 0,
// HB/gpuFunctions_undir.cuh:259-291 for RULE 1), and copy v's color forward
// into the write buffer.
// Candidates are claimed in this round's bitmap; the other bitmap holds the
// claims of the previous round, i.e. exactly the bits of L's vertices, and each
// visit of v in L zeroes v's word there, so it is clean again for the next
// round (every set bit in a zeroed word belongs to L as well).
struct FindArgs {
  GraphView G;
  const int32_t* cur;
  int32_t* nxt;
  const int32_t* prev;
  Frontier in;
  Frontier out;
  uint32_t* claim_bits;
  uint32_t* clear_bits;
  uint32_t hub_cap;
  Counters* ctr;
};

// One lane's share of a find visit of row v (slots lo, lo+stride, ... < hi).
template <int RULE>
__device__ __forceinline__ void find_slots(const FindArgs& A, int32_t v, int32_t cv, int32_t prv,
                                           int32_t pv, uint64_t rbase, uint32_t lo, uint32_t hi,
                                           uint32_t stride, unsigned long long& slots,
                                           unsigned long long& live) {
  const GraphView& G = A.G;
  for (uint32_t j = lo; j < hi; j += kUnroll * stride) {
    int32_t w[kUnroll], c[kUnroll], p[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      const uint32_t jj = j + k * stride;
      const bool in = jj < hi;
      w[k] = in ? G.col[rbase + jj] : -2;
      c[k] = in ? G.ccol[rbase + jj] : -1;
    }
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      p[k] = 0;
      if constexpr (RULE == 1) {
        if (w[k] >= 0) p[k] = G.prop[w[k]];
      }
    }
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
      slots += w[k] != -2 ? 1ull : 0ull;
      live += w[k] >= 0 ? 1ull : 0ull;
      bool mw = false, mv = false;
      if constexpr (RULE == 0) {
        // conflict with the larger-ID endpoint, or v's previous color below w's
        if (w[k] >= 0 && w[k] < G.n_owned) mw = (cv == c[k] && v < w[k]) || prv < c[k];
      } else {
        if (w[k] >= 0) {
          const bool vfirst = precedes(nullptr, v, pv, w[k], p[k]);
          mw = vfirst;                       // v precedes w: mark w (covers the conflict case)
          mv = (cv == c[k]) && !vfirst;      // conflict with a predecessor: mark v
        }
      }
      const bool pushw = mw && claim_bit(A.claim_bits, w[k]);
      frontier_push(pushw, w[k], G, A.out, A.hub_cap);
      if constexpr (RULE == 1) {
        const bool pushv = mv && claim_bit(A.claim_bits, v);
        frontier_push(pushv, v, G, A.out, A.hub_cap);
      }
    }
  }
}

template <int RULE>
__device__ __forceinline__ void find_body(const FindArgs& A) {
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const GraphView& G = A.G;
  const uint32_t nS = A.in.cnt[0], nM = A.in.cnt[1], nH = A.in.cnt[2];
  unsigned long long slots = 0, live = 0;

  // H chunks: the whole workgroup walks the chunk
  for (uint32_t ci = blockIdx.x; ci < nH; ci += gridDim.x) {
    const int4 it = A.in.h[ci];
    const int32_t v = it.x;
    const uint32_t c = (uint32_t)it.y;
    const uint2 r = G.row[v];
    const int32_t cv = A.cur[v];
    const int32_t prv = RULE == 0 ? A.prev[v] : 0;
    const int32_t pv = RULE == 1 ? G.prop[v] : 0;
    if (c == 0 && threadIdx.x == 0) {
      A.nxt[v] = cv;
      A.clear_bits[v >> 5] = 0u;
    }
    const uint32_t lo = c * kChunk;
    const uint32_t hi = min(r.y, lo + kChunk);
    find_slots<RULE>(A, v, cv, prv, pv, r.x, lo + threadIdx.x, hi, kBlock, slots, live);
  }

  // M rows: one wave per vertex
  for (int64_t i = gwave; i < nM; i += nwaves) {
    const int32_t v = A.in.m[i];
    const uint2 r = G.row[v];
    const int32_t cv = A.cur[v];
    const int32_t prv = RULE == 0 ? A.prev[v] : 0;
    const int32_t pv = RULE == 1 ? G.prop[v] : 0;
    if (lane == 0) {
      A.nxt[v] = cv;
      A.clear_bits[v >> 5] = 0u;
    }
    find_slots<RULE>(A, v, cv, prv, pv, r.x, lane, r.y, 64, slots, live);
  }

  // S rows: 8-lane groups
  {
    const int grp = lane >> 3, sub = lane & 7;
    for (int64_t b = gwave * 8; b < nS; b += nwaves * 8) {
      const int64_t idx = b + grp;
      const bool act = idx < nS;
      const int32_t v = act ? A.in.s[idx] : 0;
      const uint2 r = act ? G.row[v] : make_uint2(0u, 0u);
      const int32_t cv = act ? A.cur[v] : 0;
      const int32_t prv = (act && RULE == 0) ? A.prev[v] : 0;
      const int32_t pv = (act && RULE == 1) ? G.prop[v] : 0;
      if (act && sub == 0) {
        A.nxt[v] = cv;
        A.clear_bits[v >> 5] = 0u;
      }
      find_slots<RULE>(A, v, cv, prv, pv, r.x, sub, r.y, kGroup, slots, live);
    }
  }
  wave_add_u64(slots, &A.ctr->slots);
  wave_add_u64(live, &A.ctr->live);
}

// ----------------------------------------------------------------------------
// target bodies of the pre-loop rounds

// deletion round: target = endpoint with the larger snapshot color, tie ->
// node2 (HI/gpuFunctions_undir.cuh:153); a frozen ghost target is skipped
// (MG/gpuFunctions_undir.cuh:181).
__device__ __forceinline__ void targets_del_body(const int2* del, const uint32_t* ndel_p, GraphView G,
                                                 const int32_t* cur, uint32_t* tmark, uint32_t epoch,
                                                 Frontier T, uint32_t hub_cap) {
  const uint32_t nd = *ndel_p;
  for (uint32_t base = blockIdx.x * blockDim.x; base < nd; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    bool p = false;
    int32_t t = 0;
    if (i < nd) {
      const int2 e = del[i];
      t = cur[e.x] > cur[e.y] ? e.x : e.y;
      p = t < G.n_owned && claim(tmark, t, epoch);
    }
    frontier_push(p, t, G, T, hub_cap);
  }
}

// LDF change round: both endpoints of every deletion and of every insertion
// kept by the batch-start color filter (HB/gpuFunctions_undir.cuh:227-247).
__device__ __forceinline__ void targets_ldf_body(const int2* del, const uint32_t* ndel_p,
                                                 const int2* ins, const uint32_t* nins_p, GraphView G,
                                                 uint32_t* tmark, uint32_t epoch, Frontier T,
                                                 uint32_t hub_cap) {
  const uint32_t nd = *ndel_p, ni = *nins_p, tot = nd + ni;
  for (uint32_t base = blockIdx.x * blockDim.x; base < tot; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    int2 e = make_int2(0, 0);
    const bool ok = i < tot;
    if (ok) e = i < nd ? del[i] : ins[i - nd];
    const bool pa = ok && claim(tmark, e.x, epoch);
    frontier_push(pa, e.x, G, T, hub_cap);
    const bool pb = ok && claim(tmark, e.y, epoch);
    frontier_push(pb, e.y, G, T, hub_cap);
  }
}

// Number of distinct colors among a row's live neighbours (saturation size,
// the popcount of computeSC's mask, HI/gpuFunctions_undir.cuh:203-207) by an
// 8-lane group: lane `sub` reads slots sub, sub + 8, ...; windows of kHistW
// colours in the group's 8 LDS words (one pass per window up to the row's
// highest colour).
__device__ __forceinline__ uint32_t group_distinct_colors(const GraphView& G, uint2 r, int sub,
                                                          unsigned long long* gw) {
  uint32_t total = 0;
  int32_t rowmax = 0;
  for (int32_t base = 0; base <= rowmax; base += kHistW) {
    gw[sub] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int32_t lmax = 0;
    for (uint32_t j0 = sub; j0 < r.y; j0 += 8 * kGroup) {
      int32_t cc[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t jj = j0 + k * kGroup;
        cc[k] = jj < r.y ? G.ccol[(uint64_t)r.x + jj] : -1;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int32_t c = cc[k];
        if (c < 0) continue;
        lmax = max(lmax, c);
        const int32_t cw = c - base;
        if (cw >= 0 && cw < kHistW) atomicOr(&gw[cw >> 6], 1ull << (cw & 63));
      }
    }
    lmax = max(lmax, __shfl_xor(lmax, 1));
    lmax = max(lmax, __shfl_xor(lmax, 2));
    lmax = max(lmax, __shfl_xor(lmax, 4));
    if (base == 0) rowmax = lmax;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t pc = (uint32_t)__popcll(gw[sub]);
    pc += __shfl_xor(pc, 1);
    pc += __shfl_xor(pc, 2);
    pc += __shfl_xor(pc, 4);
    total += pc;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  return total;
}

// The same count from a row's colour histogram (nonzero bins) by an 8-lane
// group.  Exact when no colour reaches kHistW.
__device__ __forceinline__ uint32_t group_hist_distinct(const uint32_t* hrow, int sub) {
  const uint4* h4 = reinterpret_cast<const uint4*>(hrow);
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kHistW / 4 / kGroup; ++k) {
    const uint4 b = h4[k * kGroup + sub];
    c += (b.x != 0u) + (b.y != 0u) + (b.z != 0u) + (b.w != 0u);
  }
  c += __shfl_xor(c, 1);
  c += __shfl_xor(c, 2);
  c += __shfl_xor(c, 4);
  return c;
}

// highest colour of cur[0..n) into *cmax (zeroed beforehand): 16-byte loads,
// four in flight per lane, one atomic per workgroup
__device__ __forceinline__ void colour_max_body(const int32_t* cur, int32_t n, int32_t* cmax) {
  __shared__ int32_t wmax[kBlock / 64];
  int32_t m = 0;
  const int4* c4 = reinterpret_cast<const int4*>(cur);
  const uint32_t n4 = (uint32_t)n / 4;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += 4 * stride) {
    int4 x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = i + k * stride < n4 ? c4[i + k * stride] : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) m = max(m, max(max(x[k].x, x[k].y), max(x[k].z, x[k].w)));
  }
  for (uint32_t i = 4 * n4 + blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)n; i += stride)
    m = max(m, cur[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  if (lane_id() == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kBlock / 64; ++q) m = max(m, wmax[q]);
    if (m > 0) atomicMax(cmax, m);
  }
}

// frontier_push for a whole workgroup (every thread calls it): one atomic
// per list and workgroup instead of one per wave.
__device__ __forceinline__ void block_frontier_push(bool p, int32_t v, const GraphView& G,
                                                    Frontier F, uint32_t hub_cap,
                                                    uint32_t (*cnt)[2], uint32_t* base) {
  const int lane = lane_id();
  const int wib = threadIdx.x >> 6;
  const uint32_t len = p ? G.row[v].y : 0u;
  const bool is_s = p && len < kSmallLen;
  const bool is_m = p && len >= kSmallLen && len < kMidLen;
  const bool is_h = p && len >= kMidLen;
  const unsigned long long ms = __ballot(is_s), mm = __ballot(is_m);
  if (lane == 0) {
    cnt[wib][0] = (uint32_t)__popcll(ms);
    cnt[wib][1] = (uint32_t)__popcll(mm);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t tot = 0;
    for (int q = 0; q < kBlock / 64; ++q) tot += cnt[q][threadIdx.x];
    base[threadIdx.x] = tot ? atomicAdd(&F.cnt[threadIdx.x], tot) : 0u;
  }
  __syncthreads();
  uint32_t os = base[0], om = base[1];
  for (int q = 0; q < wib; ++q) {
    os += cnt[q][0];
    om += cnt[q][1];
  }
  if (is_s) F.s[os + lanes_below(ms)] = v;
  if (is_m) F.m[om + lanes_below(mm)] = v;
  if (is_h) {
    const uint32_t nch = (len + kChunk - 1) / kChunk;
    const uint32_t slot = atomicAdd(&F.cnt[3], 1u);
    if (slot < hub_cap) {
      const uint32_t b = atomicAdd(&F.cnt[2], nch);
      for (uint32_t c = 0; c < nch; ++c) F.h[b + c] = make_int4(v, (int)c, (int)slot, (int)nch);
    }
  }
  __syncthreads();
}

// insertion round: for an edge whose endpoints share a snapshot color, the
// endpoint with fewer distinct neighbour colors is the target (tie -> node2,
// HI/gpuFunctions_undir.cuh:186-209); a ghost endpoint counts as +inf (its
// owner's size is not available on this GPU, MG/gpuFunctions_undir.cuh:
// 260-263).  One 8-lane group per edge (most endpoints have a handful of
// neighbours); when no colour reaches kHistW (*cmax; no test when cmax is
// null), a row with a colour histogram counts its nonzero bins instead of
// scanning its slots.  Targets go into the frontier a workgroup at a time.
__device__ __forceinline__ void targets_ins_body(const int2* ins, const uint32_t* nins_p, GraphView G,
                                                 const int32_t* cur, const int32_t* cmax,
                                                 uint32_t* tmark, uint32_t epoch, Frontier T,
                                                 uint32_t hub_cap) {
  constexpr int kGroups = kBlock / kGroup;
  __shared__ unsigned long long gwin[kGroups][kHistW / 64];
  __shared__ uint32_t pcnt[kBlock / 64][2];
  __shared__ uint32_t pbase[2];
  const int sub = lane_id() & (kGroup - 1);
  const int gib = threadIdx.x / kGroup;
  unsigned long long* gw = gwin[gib];
  const uint32_t ni = *nins_p;
  const bool hist_ok = cmax != nullptr && *cmax < kHistW;
  auto distinct = [&](int32_t v) -> uint32_t {
    if (v >= G.n_owned) return 0xffffffffu;
    const int32_t h = hist_ok ? G.hidx[v] : -1;
    return h >= 0 ? group_hist_distinct(G.hist + (uint64_t)h * kHistW, sub)
                  : group_distinct_colors(G, G.row[v], sub, gw);
  };
  for (uint64_t b0 = (uint64_t)blockIdx.x * kGroups; b0 < ni; b0 += (uint64_t)gridDim.x * kGroups) {
    const uint64_t i = b0 + gib;
    bool push = false;
    int32_t t = 0;
    if (i < ni) {
      const int2 e = ins[i];
      if (cur[e.x] == cur[e.y]) {
        const uint32_t ka = distinct(e.x);
        const uint32_t kb = distinct(e.y);
        t = ka < kb ? e.x : e.y;
        if (sub == 0) push = t < G.n_owned && claim(tmark, t, epoch);
      }
    }
    block_frontier_push(push, t, G, T, hub_cap, pcnt, pbase);
  }
}

// copy the marked set's colors forward into the write buffer
__device__ __forceinline__ void catchup_body(Frontier F, GraphView G, const int32_t* cur, int32_t* nxt) {
  (void)G;
  const uint32_t nS = F.cnt[0], nM = F.cnt[1], nH = F.cnt[2];
  const uint32_t tot = nS + nM + nH;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += gridDim.x * blockDim.x) {
    int32_t v;
    if (i < nS) v = F.s[i];
    else if (i < nS + nM) v = F.m[i - nS];
    else v = F.h[i - nS - nM].x;
    nxt[v] = cur[v];
  }
}

}  // namespace dc
