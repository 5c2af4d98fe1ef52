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
// table), the row passes of D and X (K, cc, mex caches, and the slot colour
// cache / histograms / all-neighbour mex cache of the other loop), commit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "dc_phases.h"

namespace dc {

constexpr int kLdfBlocks = 2048;
constexpr int kLdfAcc = 32;        // round sums: workgroup b adds into slot b % kLdfAcc
constexpr uint8_t kFlD = 1, kFlX = 2, kFlPm = 4;   // per-vertex flags of a round

struct LdfCtl {
  int32_t done_round;   // the last round run (INT_MAX while running)
  int32_t pad[3];
};

struct LdfState {
  GraphView G;
  int32_t* c0;          // colour buffers (both written at commit: they agree)
  int32_t* c1;
  int32_t* K;
  int32_t* cc;
  int32_t* pm;
  int32_t* newc;
  uint8_t* F[2];        // F[r & 1] = L_r, F[(r + 1) & 1] = C_r = L_{r+1}
  uint8_t* fl;          // kFlD / kFlX / kFlPm of the current round
  const int32_t* hrowv; // rows with cap >= kSmallLen (the long-row table)
  uint32_t n_hrows;
  unsigned long long* acc;   // [2][kLdfAcc][4]: |L|, |C|, changed
  int4* trace;
  LdfCtl* ctl;
};

// ---------------------------------------------------------------------------
// A short row (used < kSmallLen) of v by its own thread: mex of the
// predecessors' colours and the predecessor slots with colour cv.  At most 63
// predecessors, so the mex is below 64.
__device__ __forceinline__ void short_pred_scan(const GraphView& G, int32_t v, int32_t cv,
                                                int32_t& pm, int32_t& cc) {
  const uint2 r = G.row[v];
  const int32_t pv = G.prop[v];
  uint32_t lo = 0u, hi = 0u;   // colours 0..31, 32..63
  int32_t k = 0;
  for (uint32_t j0 = 0; j0 < r.y; j0 += 4) {
    int32_t w[4], c[4], p[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool in = j0 + q < r.y;
      w[q] = in ? G.col[(uint64_t)r.x + j0 + q] : -1;
      c[q] = in ? G.ccol[(uint64_t)r.x + j0 + q] : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) p[q] = w[q] >= 0 ? G.prop[w[q]] : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (w[q] < 0 || !precedes(nullptr, w[q], p[q], v, pv)) continue;
      const uint32_t cq = (uint32_t)c[q];
      lo |= cq < 32u ? 1u << cq : 0u;
      hi |= cq - 32u < 32u ? 1u << (cq - 32u) : 0u;
      k += c[q] == cv ? 1 : 0;
    }
  }
  pm = lo != ~0u ? __ffs(~lo) - 1 : 32 + __ffs(~hi) - 1;
  cc = k;
}

// The same for a long row by a whole wave: the mex over LDS windows of
// kWinBits colours (more than kWinBits predecessors only if the window is
// full), cc by a wave reduction.  Every lane returns both.
__device__ __forceinline__ void wave_pred_scan(const GraphView& G, int32_t v, int32_t cv,
                                               unsigned long long* ww, int32_t& pm, int32_t& cc) {
  const int lane = lane_id();
  const uint2 r = G.row[v];
  const int32_t pv = G.prop[v];
  int32_t k = 0;
  for (int32_t base = 0;; base += kWinBits) {
    for (int i = lane; i < kWinWords; i += 64) ww[i] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t j = lane; j < r.y; j += 64) {
      const int32_t w = G.col[(uint64_t)r.x + j];
      if (w < 0) continue;
      if (!precedes(nullptr, w, G.prop[w], v, pv)) continue;
      const int32_t c = G.ccol[(uint64_t)r.x + j];
      if (base == 0 && c == cv) ++k;
      const int32_t cw = c - base;
      if (cw >= 0 && cw < kWinBits) atomicOr(&ww[cw >> 6], 1ull << (cw & 63));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int32_t found = INT_MAX;
    for (int i = lane; i < kWinWords; i += 64)
      if (ww[i] != ~0ull && found == INT_MAX) found = i * 64 + __ffsll((long long)~ww[i]) - 1;
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
// batch set-up: every owned vertex's cc and (short rows) cached mex
__global__ __launch_bounds__(kBlock) void k_ldf_init_cc(LdfState S) {
  __shared__ unsigned long long win[kBlock / 64][kWinWords];
  const GraphView& G = S.G;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < G.n_owned;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (G.hidx[v] >= 0) continue;   // long rows below
    int32_t pm, cc;
    short_pred_scan(G, (int32_t)v, S.c0[v], pm, cc);
    S.pm[v] = pm;
    S.cc[v] = cc;
  }
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t i = gwave; i < S.n_hrows; i += nwaves) {
    const int32_t v = S.hrowv[i];
    int32_t pm, cc;
    wave_pred_scan(G, v, S.c0[v], win[threadIdx.x >> 6], pm, cc);
    if (lane == 0) {
      S.pm[v] = pm;
      S.cc[v] = cc;
    }
  }
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
// round r, launch 1: members, their mex (short rows), D and X flags
__global__ __launch_bounds__(kBlock) void k_ldf_members(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  const GraphView& G = S.G;
  const uint8_t* Lf = S.F[r & 1];
  uint8_t* Cf = S.F[(r + 1) & 1];
  uint32_t nl = 0, nc = 0, nx = 0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < G.n_owned;
       v += (int64_t)gridDim.x * blockDim.x) {
    const bool inl = Lf[v] != 0;
    const bool member = S.K[v] > 0 || (inl && S.cc[v] > 0);
    Cf[v] = member ? 1 : 0;
    uint8_t f = inl != member ? kFlD : 0;
    nl += inl ? 1u : 0u;
    nc += member ? 1u : 0u;
    if (member) {
      int32_t m = S.pm[v];
      if (m < 0) {
        if (G.hidx[v] < 0) {
          int32_t cc;
          short_pred_scan(G, (int32_t)v, S.c0[v], m, cc);
          S.pm[v] = m;
        } else {
          f |= kFlPm;   // long row: the next launch
        }
      }
      if (m >= 0 && m != S.c0[v]) {
        S.newc[v] = m;
        S.cc[v] = 0;   // its changed predecessors re-count it in the row pass
        f |= kFlX;
        ++nx;
      }
    }
    S.fl[v] = f;
  }
  ldf_sums(S, r & 1, nl, nc, nx);
}

// launch 2: stale mex of long rows, a wave per row of the long-row table
__global__ __launch_bounds__(kBlock) void k_ldf_long_mex(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  __shared__ unsigned long long win[kBlock / 64][kWinWords];
  const GraphView& G = S.G;
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  uint32_t nx = 0;
  for (uint32_t i = gwave; i < S.n_hrows; i += nwaves) {
    const int32_t v = S.hrowv[i];
    const uint8_t f = S.fl[v];
    if (!(f & kFlPm)) continue;
    int32_t m, cc;
    wave_pred_scan(G, v, S.c0[v], win[threadIdx.x >> 6], m, cc);
    if (lane == 0) {
      S.pm[v] = m;
      if (m != S.c0[v]) {
        S.newc[v] = m;
        S.cc[v] = 0;
        S.fl[v] = f | kFlX;
        ++nx;
      }
    }
  }
  ldf_sums(S, r & 1, 0u, 0u, nx);
}

// One lane's share of the row pass of x (slots lo, lo + stride, ...): to
// each successor y, D adds x's +-1 to K[y]; X moves cc[y] and drops y's
// cached mex.
__device__ __forceinline__ void ldf_row_part(const LdfState& S, int32_t x, uint8_t f, int32_t ksign,
                                             int32_t a, int32_t b, uint64_t rbase, uint32_t lo,
                                             uint32_t hi, uint32_t stride) {
  const GraphView& G = S.G;
  const int32_t px = G.prop[x];
  for (uint32_t j0 = lo; j0 < hi; j0 += 4 * stride) {
    int32_t y[4], p[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t j = j0 + q * stride;
      y[q] = j < hi ? G.col[rbase + j] : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) p[q] = y[q] >= 0 ? G.prop[y[q]] : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (y[q] < 0 || !precedes(nullptr, x, px, y[q], p[q])) continue;   // x precedes y
      if (f & kFlD) atomicAdd(&S.K[y[q]], ksign);
      if (f & kFlX) {
        const uint8_t fy = S.fl[y[q]];
        int32_t dlt;
        if (fy & kFlX) dlt = b == S.newc[y[q]] ? 1 : 0;
        else {
          const int32_t cy = S.c0[y[q]];
          dlt = (b == cy ? 1 : 0) - (a == cy ? 1 : 0);
        }
        if (dlt) atomicAdd(&S.cc[y[q]], dlt);
        S.pm[y[q]] = -1;
      }
    }
  }
}

// launch 3: the row passes of D and X (short rows by their own thread, long
// rows by a wave), and each X member's new colour into its neighbours' slot
// caches, histograms and all-neighbour mex caches
__global__ __launch_bounds__(kBlock) void k_ldf_rows(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  const GraphView& G = S.G;
  const uint8_t* Cf = S.F[(r + 1) & 1];
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < G.n_owned;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t f = S.fl[v];
    if (!(f & (kFlD | kFlX)) || G.hidx[v] >= 0) continue;
    const uint2 rw = G.row[v];
    const int32_t a = S.c0[v];
    const int32_t b = (f & kFlX) ? S.newc[v] : a;
    ldf_row_part(S, (int32_t)v, f, Cf[v] ? 1 : -1, a, b, rw.x, 0, rw.y, 1);
    if (f & kFlX) propagate_slots(G, b, a, rw.x, 0, rw.y, 1);
  }
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t i = gwave; i < S.n_hrows; i += nwaves) {
    const int32_t v = S.hrowv[i];
    const uint8_t f = S.fl[v];
    if (!(f & (kFlD | kFlX))) continue;
    const uint2 rw = G.row[v];
    const int32_t a = S.c0[v];
    const int32_t b = (f & kFlX) ? S.newc[v] : a;
    ldf_row_part(S, v, f, Cf[v] ? 1 : -1, a, b, rw.x, lane, rw.y, 64);
    if (f & kFlX) propagate_slots(G, b, a, rw.x, lane, rw.y, 64);
  }
}

// launch 4: commit the new colours; round r's counts into the trace; the
// loop ends after a round without a colour change
__global__ __launch_bounds__(kBlock) void k_ldf_commit(LdfState S, int32_t r) {
  if (S.ctl->done_round < r) return;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < S.G.n_owned;
       v += (int64_t)gridDim.x * blockDim.x)
    if (S.fl[v] & kFlX) {
      S.c0[v] = S.newc[v];
      S.c1[v] = S.newc[v];
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
};

static LdfHost* ldf_state(Graph* g) { return static_cast<LdfHost*>(g->ldf); }

void ldf_free(Graph* g) {
  LdfHost* s = ldf_state(g);
  if (!s) return;
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(s->S.K); f(s->S.cc); f(s->S.pm); f(s->S.newc); f(s->flags); f(s->S.acc); f(s->S.trace);
  f(s->S.ctl);
  if (s->hctl) (void)hipHostFree(s->hctl);
  delete s;
  g->ldf = nullptr;
}

static int ldf_alloc(Graph* g, LdfHost* s) {
  const size_t n = (size_t)g->n + 1;
  LdfState& S = s->S;
  DC_HIP(hipMalloc(&S.K, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.cc, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.pm, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.newc, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&s->flags, 3 * n));
  S.F[0] = s->flags;
  S.F[1] = s->flags + n;
  S.fl = s->flags + 2 * n;
  DC_HIP(hipMalloc(&S.acc, sizeof(unsigned long long) * 2 * kLdfAcc * 4));
  DC_HIP(hipMalloc(&S.ctl, sizeof(LdfCtl)));
  DC_HIP(hipHostMalloc(&s->hctl, sizeof(LdfCtl), hipHostMallocDefault));
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
    if (rc) return rc;
  }
  int rc = ldf_trace(g, s, 256);
  if (rc) return rc;
  LdfState& S = s->S;
  S.G = g->view();
  S.c0 = g->color[g->cur];
  S.c1 = g->color[g->cur ^ 1];
  S.hrowv = g->hrowv;
  S.n_hrows = g->n_hrows;
  const size_t n = (size_t)g->n + 1;
  DC_HIP(hipMemsetAsync(S.K, 0, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(s->flags, 0, 3 * n, g->stream));
  DC_HIP(hipMemsetAsync(S.acc, 0, sizeof(unsigned long long) * 2 * kLdfAcc * 4, g->stream));
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
  while (r < max_rounds && rounds < 0) {
    const int32_t nb = std::min(ldf_block_end(r), max_rounds) - r;
    if ((rc = ldf_trace(g, s, r + nb))) return rc;
    for (int32_t i = 0; i < nb; ++i) {
      hipLaunchKernelGGL(k_ldf_members, grid, blk, 0, g->stream, S, r + i);
      hipLaunchKernelGGL(k_ldf_long_mex, grid, blk, 0, g->stream, S, r + i);
      hipLaunchKernelGGL(k_ldf_rows, grid, blk, 0, g->stream, S, r + i);
      hipLaunchKernelGGL(k_ldf_commit, grid, blk, 0, g->stream, S, r + i);
    }
    DC_HIP(hipGetLastError());
    launches += 4 * nb;
    DC_HIP(hipMemcpyAsync(s->hctl, S.ctl, sizeof(LdfCtl), hipMemcpyDeviceToHost, g->stream));
    DC_HIP(hipStreamSynchronize(g->stream));
    r += nb;
    if (s->hctl->done_round != INT_MAX) rounds = s->hctl->done_round + 1;
  }
  if (rounds < 0) {
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
  return DC_OK;
}

}  // namespace dc
