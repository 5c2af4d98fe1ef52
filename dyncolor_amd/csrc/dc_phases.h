// dc_phases.h -- device bodies of every recolor phase, shared by the
// one-launch-per-phase kernels (recolor.hip) and the persistent cooperative
// kernel (fused.hip).  Each body walks its whole input with grid-stride loops
// over the launch's workgroups and reads its item counts from device memory.
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
};

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
  unsigned long long* hubmask;
  uint32_t* hubdone;
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
                                              int32_t m, uint32_t& changed) {
  bool ch = m != cv;
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
    if (w < A.G.n_owned && !include_nbr<MODE>(v, pv, w, pw, A.G.n_owned))
      return atomicSub(&A.pending[w], 1) == 1;
  }
  return false;
}

// ----------------------------------------------------------------------------
// recolor kernel: evaluate every vertex of A.in
template <int MODE>
__device__ __forceinline__ void recolor_body(const RecolorArgs& A) {
  __shared__ unsigned long long win[kBlock / 64][kWinWords];
  __shared__ unsigned long long bwin[kWinWords];
  __shared__ int sh_last;
  const int lane = lane_id();
  const int wib = threadIdx.x >> 6;
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const GraphView& G = A.G;
  const uint32_t nS = A.in.cnt[0], nM = A.in.cnt[1], nH = A.in.cnt[2];
  uint32_t changed = 0;
  unsigned long long slots = 0, live = 0;

  // ---- class H: one workgroup per 4096-slot chunk, partial masks merged in a
  //      per-hub global window; the last chunk to finish computes the mex.
  for (uint32_t ci = blockIdx.x; ci < nH; ci += gridDim.x) {
    int4 it = A.in.h[ci];
    const int32_t v = it.x;
    const uint32_t c = (uint32_t)it.y, slot = (uint32_t)it.z, nch = (uint32_t)it.w;
    const uint2 r = G.row[v];
    const int32_t pv = ModeTraits<MODE>::prop ? G.prop[v] : 0;
    for (int i = threadIdx.x; i < kWinWords; i += kBlock) bwin[i] = 0ull;
    __syncthreads();
    const uint32_t lo = c * kChunk;
    const uint32_t hi = min(r.y, lo + kChunk);
    for (uint32_t j = lo + threadIdx.x; j < hi; j += kBlock) {
      int32_t w = G.col[(uint64_t)r.x + j];
      ++slots;
      if (w < 0) continue;
      ++live;
      int32_t pw = ModeTraits<MODE>::prop ? G.prop[w] : 0;
      if (include_nbr<MODE>(v, pv, w, pw, G.n_owned)) {
        int32_t cw = A.cur[w];
        if (cw >= 0 && cw < kWinBits) atomicOr(&bwin[cw >> 6], 1ull << (cw & 63));
      }
      if constexpr (ModeTraits<MODE>::stat) {
        bool ready = release_nbr<MODE>(A, v, pv, w, pw);
        if (ready) {
          // scheduled from inside a block loop: append lane by lane
          frontier_push(true, w, G, A.out, A.hub_cap);
        }
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kWinWords; i += kBlock) {
      unsigned long long x = bwin[i];
      if (x) atomicOr(&A.hubmask[(uint64_t)slot * kWinWords + i], x);
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) sh_last = atomicAdd(&A.hubdone[slot], 1u) == nch - 1;
    __syncthreads();
    if (sh_last) {
      __threadfence();
      // read-and-clear the merged window
      for (int i = threadIdx.x; i < kWinWords; i += kBlock)
        bwin[i] = atomicExch(&A.hubmask[(uint64_t)slot * kWinWords + i], 0ull);
      __syncthreads();
      __shared__ int sh_mex;
      if (threadIdx.x == 0) {
        int m = -1;
        for (int i = 0; i < kWinWords; ++i)
          if (bwin[i] != ~0ull) {
            m = i * 64 + __ffsll((long long)~bwin[i]) - 1;
            break;
          }
        sh_mex = m;
        A.hubdone[slot] = 0;
      }
      __syncthreads();
      // window full: rescan the whole row, one 8192-color window at a time
      int32_t base = kWinBits;
      while (sh_mex < 0) {
        if (threadIdx.x == 0) A.ctr->window_full = 1;
        for (int i = threadIdx.x; i < kWinWords; i += kBlock) bwin[i] = 0ull;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < r.y; j += kBlock) {
          int32_t w = G.col[(uint64_t)r.x + j];
          if (w < 0) continue;
          int32_t pw = ModeTraits<MODE>::prop ? G.prop[w] : 0;
          if (!include_nbr<MODE>(v, pv, w, pw, G.n_owned)) continue;
          int32_t cw = A.cur[w] - base;
          if (cw >= 0 && cw < kWinBits) atomicOr(&bwin[cw >> 6], 1ull << (cw & 63));
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          for (int i = 0; i < kWinWords; ++i)
            if (bwin[i] != ~0ull) {
              sh_mex = base + i * 64 + __ffsll((long long)~bwin[i]) - 1;
              break;
            }
        }
        __syncthreads();
        base += kWinBits;
      }
      bool push = false;
      if (threadIdx.x == 0) push = finish_vertex<MODE>(A, v, A.cur[v], sh_mex, changed);
      if (wib == 0) frontier_push(push, v, G, A.out, A.hub_cap);
    }
    __syncthreads();
  }

  // ---- class M: one wave per vertex, 8192-bit LDS window (mex <= slots <= 4096)
  unsigned long long* ww = win[wib];
  for (int64_t i = gwave; i < nM; i += nwaves) {
    const int32_t v = A.in.m[i];
    const uint2 r = G.row[v];
    const int32_t pv = ModeTraits<MODE>::prop ? G.prop[v] : 0;
    ww[lane] = 0ull;
    ww[lane + 64] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t j = lane; j < r.y; j += 64) {
      int32_t w = G.col[(uint64_t)r.x + j];
      ++slots;
      bool ready = false;
      if (w >= 0) {
        ++live;
        int32_t pw = ModeTraits<MODE>::prop ? G.prop[w] : 0;
        if (include_nbr<MODE>(v, pv, w, pw, G.n_owned)) {
          int32_t cw = A.cur[w];
          if (cw >= 0 && cw < kWinBits) atomicOr(&ww[cw >> 6], 1ull << (cw & 63));
        }
        ready = release_nbr<MODE>(A, v, pv, w, pw);
      }
      if constexpr (ModeTraits<MODE>::stat) frontier_push(ready, w, G, A.out, A.hub_cap);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    unsigned long long w0 = ww[lane], w1 = ww[lane + 64];
    unsigned long long b0 = __ballot(w0 != ~0ull);
    int32_t m;
    if (b0) {
      int wi = __ffsll((long long)b0) - 1;
      unsigned long long word = __shfl(w0, wi);
      m = wi * 64 + __ffsll((long long)~word) - 1;
    } else {
      unsigned long long b1 = __ballot(w1 != ~0ull);
      int wi = __ffsll((long long)b1) - 1;
      unsigned long long word = __shfl(w1, wi);
      m = (64 + wi) * 64 + __ffsll((long long)~word) - 1;
    }
    bool push = false;
    if (lane == 0) push = finish_vertex<MODE>(A, v, A.cur[v], m, changed);
    frontier_push(push, v, G, A.out, A.hub_cap);
  }

  // ---- class S: 8-lane groups, one u64 mask (mex <= slots < 64)
  {
    const int grp = lane >> 3, sub = lane & 7;
    for (int64_t b = gwave * 8; b < nS; b += nwaves * 8) {
      const int64_t idx = b + grp;
      const bool act = idx < nS;
      const int32_t v = act ? A.in.s[idx] : 0;
      const uint2 r = act ? G.row[v] : make_uint2(0u, 0u);
      const int32_t pv = (act && ModeTraits<MODE>::prop) ? G.prop[v] : 0;
      unsigned long long mask = 0ull;
      for (uint32_t j = sub; j < r.y; j += kGroup) {
        int32_t w = G.col[(uint64_t)r.x + j];
        ++slots;
        bool ready = false;
        if (w >= 0) {
          ++live;
          int32_t pw = ModeTraits<MODE>::prop ? G.prop[w] : 0;
          if (include_nbr<MODE>(v, pv, w, pw, G.n_owned)) {
            int32_t cw = A.cur[w];
            if (cw >= 0 && cw < 64) mask |= 1ull << cw;
          }
          ready = release_nbr<MODE>(A, v, pv, w, pw);
        }
        if constexpr (ModeTraits<MODE>::stat) frontier_push(ready, w, G, A.out, A.hub_cap);
      }
      mask |= __shfl_xor(mask, 1);
      mask |= __shfl_xor(mask, 2);
      mask |= __shfl_xor(mask, 4);
      bool push = false;
      if (act && sub == 0) {
        int32_t m = __ffsll((long long)~mask) - 1;
        push = finish_vertex<MODE>(A, v, A.cur[v], m, changed);
      }
      frontier_push(push, v, G, A.out, A.hub_cap);
    }
  }
  wave_add_u32(changed, A.changed);
  wave_add_u64(slots, A.rslots);
  wave_add_u64(live, A.rlive);
}

// ----------------------------------------------------------------------------
// find kernel: for v in L, mark the candidates of the next recolor
// (findEligibleNeighbors: HI/gpuFunctions_undir.cuh:229-260 for RULE 0,
// HB/gpuFunctions_undir.cuh:259-291 for RULE 1), and copy v's color forward
// into the write buffer.
struct FindArgs {
  GraphView G;
  const int32_t* cur;
  int32_t* nxt;
  const int32_t* prev;
  Frontier in;
  Frontier out;
  uint32_t* tmark;
  uint32_t epoch;
  uint32_t hub_cap;
  Counters* ctr;
};

template <int RULE>
__device__ __forceinline__ void find_visit(const FindArgs& A, int32_t v, int32_t cv, int32_t prv,
                                           int32_t pv, int32_t w, bool& pw_out, bool& pv_out) {
  // pw_out: w becomes a candidate; pv_out: v becomes a candidate
  const GraphView& G = A.G;
  if constexpr (RULE == 0) {
    if (w < G.n_owned) {
      int32_t cw = A.cur[w];
      pw_out = (cv == cw && v < w) || prv < cw;
    }
  } else {
    int32_t cw = A.cur[w];
    int32_t pw = G.prop[w];
    bool vfirst = precedes(nullptr, v, pv, w, pw);
    pw_out = vfirst;                     // v precedes w: mark w (covers the conflict case too)
    pv_out = (cv == cw) && !vfirst;      // conflict with a predecessor: mark v
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

  // class H chunks: the whole block walks the chunk
  for (uint32_t ci = blockIdx.x; ci < nH; ci += gridDim.x) {
    int4 it = A.in.h[ci];
    const int32_t v = it.x;
    const uint32_t c = (uint32_t)it.y;
    const uint2 r = G.row[v];
    const int32_t cv = A.cur[v];
    const int32_t prv = RULE == 0 ? A.prev[v] : 0;
    const int32_t pv = RULE == 1 ? G.prop[v] : 0;
    if (c == 0 && threadIdx.x == 0) A.nxt[v] = cv;
    const uint32_t lo = c * kChunk;
    const uint32_t hi = min(r.y, lo + kChunk);
    for (uint32_t j0 = lo; j0 < hi; j0 += kBlock) {
      uint32_t j = j0 + threadIdx.x;
      bool mw = false, mv = false;
      int32_t w = -1;
      if (j < hi) {
        w = G.col[(uint64_t)r.x + j];
        ++slots;
        if (w >= 0) {
          ++live;
          find_visit<RULE>(A, v, cv, prv, pv, w, mw, mv);
        }
      }
      bool pushw = mw && claim(A.tmark, w, A.epoch);
      frontier_push(pushw, w, G, A.out, A.hub_cap);
      if constexpr (RULE == 1) {
        bool pushv = mv && claim(A.tmark, v, A.epoch);
        frontier_push(pushv, v, G, A.out, A.hub_cap);
      }
    }
  }

  // class M: one wave per vertex
  for (int64_t i = gwave; i < nM; i += nwaves) {
    const int32_t v = A.in.m[i];
    const uint2 r = G.row[v];
    const int32_t cv = A.cur[v];
    const int32_t prv = RULE == 0 ? A.prev[v] : 0;
    const int32_t pv = RULE == 1 ? G.prop[v] : 0;
    if (lane == 0) A.nxt[v] = cv;
    for (uint32_t j = lane; j < r.y; j += 64) {
      int32_t w = G.col[(uint64_t)r.x + j];
      ++slots;
      bool mw = false, mv = false;
      if (w >= 0) {
        ++live;
        find_visit<RULE>(A, v, cv, prv, pv, w, mw, mv);
      }
      bool pushw = mw && claim(A.tmark, w, A.epoch);
      frontier_push(pushw, w, G, A.out, A.hub_cap);
      if constexpr (RULE == 1) {
        bool pushv = mv && claim(A.tmark, v, A.epoch);
        frontier_push(pushv, v, G, A.out, A.hub_cap);
      }
    }
  }

  // class S: 8-lane groups
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
      if (act && sub == 0) A.nxt[v] = cv;
      for (uint32_t j = sub; j < r.y; j += kGroup) {
        int32_t w = G.col[(uint64_t)r.x + j];
        ++slots;
        bool mw = false, mv = false;
        if (w >= 0) {
          ++live;
          find_visit<RULE>(A, v, cv, prv, pv, w, mw, mv);
        }
        bool pushw = mw && claim(A.tmark, w, A.epoch);
        frontier_push(pushw, w, G, A.out, A.hub_cap);
        if constexpr (RULE == 1) {
          bool pushv = mv && claim(A.tmark, v, A.epoch);
          frontier_push(pushv, v, G, A.out, A.hub_cap);
        }
      }
    }
  }
  wave_add_u64(slots, &A.ctr->slots);
  wave_add_u64(live, &A.ctr->live);
}

// ----------------------------------------------------------------------------
// target kernels of the pre-loop rounds

// deletion round: target = endpoint with the larger snapshot color, tie ->
// node2 (HI/gpuFunctions_undir.cuh:153); a frozen ghost target is skipped
// (MG/gpuFunctions_undir.cuh:181).
__device__ __forceinline__ void targets_del_body(const int2* del, const uint32_t* ndel_p, GraphView G, const int32_t* cur,
                              uint32_t* tmark, uint32_t epoch, Frontier T, uint32_t hub_cap) {
  const uint32_t nd = *ndel_p;
  for (uint32_t base = blockIdx.x * blockDim.x; base < nd; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    bool p = false;
    int32_t t = 0;
    if (i < nd) {
      int2 e = del[i];
      t = cur[e.x] > cur[e.y] ? e.x : e.y;
      p = t < G.n_owned && claim(tmark, t, epoch);
    }
    frontier_push(p, t, G, T, hub_cap);
  }
}

// LDF change round: both endpoints of every deletion and of every insertion
// kept by the batch-start color filter (HB/gpuFunctions_undir.cuh:227-247).
__device__ __forceinline__ void targets_ldf_body(const int2* del, const uint32_t* ndel_p, const int2* ins,
                              const uint32_t* nins_p, GraphView G, uint32_t* tmark, uint32_t epoch,
                              Frontier T, uint32_t hub_cap) {
  const uint32_t nd = *ndel_p, ni = *nins_p, tot = nd + ni;
  for (uint32_t base = blockIdx.x * blockDim.x; base < tot; base += gridDim.x * blockDim.x) {
    uint32_t i = base + threadIdx.x;
    int2 e = make_int2(0, 0);
    bool ok = i < tot;
    if (ok) e = i < nd ? del[i] : ins[i - nd];
    bool pa = ok && claim(tmark, e.x, epoch);
    frontier_push(pa, e.x, G, T, hub_cap);
    bool pb = ok && claim(tmark, e.y, epoch);
    frontier_push(pb, e.y, G, T, hub_cap);
  }
}

// Number of distinct colors among v's live neighbours (saturation size, the
// popcount of computeSC's mask, HI/gpuFunctions_undir.cuh:203-207), one wave,
// windows of 8192 colors in LDS.
__device__ inline uint32_t wave_distinct_colors(const GraphView& G, const int32_t* cur, int32_t v,
                                         unsigned long long* ww) {
  const int lane = lane_id();
  const uint2 r = G.row[v];
  uint32_t total = 0;
  int32_t rowmax = 0;
  for (int32_t base = 0; base <= rowmax; base += kWinBits) {
    ww[lane] = 0ull;
    ww[lane + 64] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int32_t lmax = 0;
    for (uint32_t j = lane; j < r.y; j += 64) {
      int32_t w = G.col[(uint64_t)r.x + j];
      if (w < 0) continue;
      int32_t c = cur[w];
      lmax = max(lmax, c);
      int32_t cw = c - base;
      if (cw >= 0 && cw < kWinBits) atomicOr(&ww[cw >> 6], 1ull << (cw & 63));
    }
    for (int off = 32; off > 0; off >>= 1) lmax = max(lmax, __shfl_xor(lmax, off));
    if (base == 0) rowmax = lmax;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t pc = (uint32_t)(__popcll(ww[lane]) + __popcll(ww[lane + 64]));
    for (int off = 32; off > 0; off >>= 1) pc += __shfl_xor(pc, off);
    total += pc;
    __builtin_amdgcn_wave_barrier();
  }
  return total;
}

// insertion round: for an edge whose endpoints share a snapshot color, the
// endpoint with fewer distinct neighbour colors is the target (tie -> node2,
// HI/gpuFunctions_undir.cuh:186-209); a ghost endpoint counts as +inf (its
// owner's size is not available on this GPU, MG/gpuFunctions_undir.cuh:
// 260-263).  One wave per edge.
__device__ __forceinline__ void targets_ins_body(const int2* ins, const uint32_t* nins_p,
                                                        GraphView G, const int32_t* cur,
                                                        uint32_t* tmark, uint32_t epoch, Frontier T,
                                                        uint32_t hub_cap) {
  __shared__ unsigned long long win[kBlock / 64][kWinWords];
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long* ww = win[threadIdx.x >> 6];
  const uint32_t ni = *nins_p;
  for (int64_t i = gwave; i < ni; i += nwaves) {
    int2 e = ins[i];
    bool push = false;
    int32_t t = 0;
    if (cur[e.x] == cur[e.y]) {
      uint32_t ka = e.x >= G.n_owned ? 0xffffffffu : wave_distinct_colors(G, cur, e.x, ww);
      uint32_t kb = e.y >= G.n_owned ? 0xffffffffu : wave_distinct_colors(G, cur, e.y, ww);
      t = ka < kb ? e.x : e.y;
      if (lane == 0) push = t < G.n_owned && claim(tmark, t, epoch);
    }
    frontier_push(push, t, G, T, hub_cap);
  }
}

// copy the marked set's colors forward into the write buffer
__device__ __forceinline__ void catchup_body(Frontier F, GraphView G, const int32_t* cur, int32_t* nxt) {
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
