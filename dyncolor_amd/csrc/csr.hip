// csr.hip -- device-resident undirected CSR with per-row spare capacity and
// tombstones: build, batch apply (insert / tombstone / LDF counter), regrow,
// compaction, and the scan used to lay rows out.
//
// Layout: row[v] = {start, used}; slots [start, start+used) hold neighbours,
// -1 marks a deleted slot; [used, cap[v]) is spare room for insertions.  Both
// directions of an edge are stored; ghost rows (partition share) are empty.
// This replaces the reference's host-flattened AoS {col, flag} rows that are
// rebuilt on every run (HI/supportingFunctions.cu:27-67).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include <hipcub/hipcub.hpp>

#include "dc_device.h"
#include "dc_internal.h"

namespace dc {

// ---------------------------------------------------------------- scan -----
// Exclusive scan of n u32 values into u64 (+ total at out[n]).  Three passes:
// per-block sums, one-block scan of the block sums, per-block rescan.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t x, uint64_t* total) {
  __shared__ uint64_t wsum[kBlock / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t inc = x;
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(inc, d);
    if (lane >= d) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (int i = 0; i < kBlock / 64; ++i) {
    if (i < w) off += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

__global__ void k_scan_partials(const uint32_t* in, int32_t n, uint64_t* part) {
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint64_t s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + (int64_t)i * kBlock + threadIdx.x;
    if (j < n) s += in[j];
  }
  uint64_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void k_scan_partials_scan(uint64_t* part, int32_t nb) {
  // single block: sequential over tiles of kBlock partials
  __shared__ uint64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int32_t t = 0; t < nb; t += kBlock) {
    int32_t j = t + threadIdx.x;
    uint64_t x = j < nb ? part[j] : 0;
    uint64_t tot;
    uint64_t ex = block_exclusive_scan(x, &tot);
    if (j < nb) part[j] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nb] = carry;
}

__global__ void k_scan_final(const uint32_t* in, int32_t n, const uint64_t* part, uint64_t* out) {
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + i;
    v[i] = j < n ? in[j] : 0u;
    s += v[i];
  }
  uint64_t tot;
  uint64_t ex = block_exclusive_scan(s, &tot) + part[blockIdx.x];
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + i;
    if (j < n) out[j] = ex;
    ex += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = part[gridDim.x];
}

int exclusive_scan_u32_to_u64(Graph* g, const uint32_t* in, uint64_t* out, int32_t n) {
  int32_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 0) nb = 1;
  if (g->scan_tmp_n < (size_t)nb + 1) {
    if (g->scan_tmp) DC_HIP(hipFree(g->scan_tmp));
    g->scan_tmp = nullptr;
    DC_HIP(hipMalloc(&g->scan_tmp, sizeof(uint64_t) * ((size_t)nb + 1)));
    g->scan_tmp_n = (size_t)nb + 1;
  }
  hipLaunchKernelGGL(k_scan_partials, dim3(nb), dim3(kBlock), 0, g->stream, in, n, g->scan_tmp);
  hipLaunchKernelGGL(k_scan_partials_scan, dim3(1), dim3(kBlock), 0, g->stream, g->scan_tmp, nb);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kBlock), 0, g->stream, in, n, g->scan_tmp, out);
  DC_HIP(hipGetLastError());
  return DC_OK;
}

// ------------------------------------------------------------- build -------
static inline int grid_for(int64_t items, int per_block = kBlock) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return (int)b;
}

__global__ void k_count_degree(int64_t m, const int32_t* u, const int32_t* v, int32_t n_owned,
                               uint32_t* deg, int32_t* prop) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t a = u[i], b = v[i];
    if (a < n_owned) atomicAdd(&deg[a], 1u);
    if (b < n_owned) atomicAdd(&deg[b], 1u);
    atomicAdd(&prop[a], 1);
    atomicAdd(&prop[b], 1);
  }
}

__device__ __forceinline__ uint32_t slack_for(uint32_t used, float frac, int32_t smin) {
  uint32_t s = (uint32_t)((float)used * frac + 0.999f);
  return s < (uint32_t)smin ? (uint32_t)smin : s;
}

// cap[v] <- max(cap[v], used + slack(used)) for owned rows, 0 for ghosts.
// With keep_old = 0 the input cap holds the degree and is replaced.
__global__ void k_set_cap(int32_t n, int32_t n_owned, uint32_t* cap, const uint2* row,
                          int keep_old, float frac, int32_t smin) {
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    if (v >= n_owned) {
      cap[v] = 0;
      continue;
    }
    if (keep_old) {
      uint32_t used = row[v].y;
      uint32_t c = cap[v];
      if (used > c) c = used + slack_for(used, frac, smin);
      cap[v] = c;
    } else {
      uint32_t d = cap[v];
      cap[v] = d + slack_for(d, frac, smin);
    }
  }
}

__global__ void k_set_rows(int32_t n, const uint64_t* start, uint2* row, int reset_used) {
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    uint2 r = row[v];
    r.x = (uint32_t)start[v];
    if (reset_used) r.y = 0;
    row[v] = r;
  }
}

// Both directions of an edge are appended by the same thread, so each slot
// learns its twin (the reverse slot's position) here.
__global__ void k_fill(int64_t m, const int32_t* u, const int32_t* v, GraphView G) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = u[i], b = v[i];
    const bool oa = a < G.n_owned, ob = b < G.n_owned;
    const uint32_t pa = oa ? atomicAdd(&G.row[a].y, 1u) : 0u;
    const uint32_t pb = ob ? atomicAdd(&G.row[b].y, 1u) : 0u;
    if (oa) {
      const uint64_t s = (uint64_t)G.row[a].x + pa;
      G.col[s] = b;
      G.twin[s] = ob ? (int32_t)pb : -1;
    }
    if (ob) {
      const uint64_t s = (uint64_t)G.row[b].x + pb;
      G.col[s] = a;
      G.twin[s] = oa ? (int32_t)pa : -1;
    }
  }
}

// ccol[j] <- colour of col[j] (tombstones -1) for every used slot of the rows
// below kMidLen slots of capacity.  A wave takes 64 vertices at a time: rows
// below kSmallLen by 8-lane groups (a lane per slot of a group's row), the
// longer ones by the whole wave, one after the other.
__global__ __launch_bounds__(kBlock) void k_ccol_rows(GraphView G, const int32_t* color) {
  __shared__ int32_t lists[kBlock / 64][64];
  int32_t* lst = lists[threadIdx.x >> 6];
  const int lane = lane_id();
  const int grp = lane >> 3, sub = lane & 7;
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  auto slots = [&](uint2 r, uint32_t lo, uint32_t stride) {
    for (uint32_t j = lo; j < r.y; j += stride) {
      const int32_t w = G.col[(uint64_t)r.x + j];
      G.ccol[(uint64_t)r.x + j] = w >= 0 ? color[w] : -1;
    }
  };
  for (int64_t base = gwave * 64; base < G.n_owned; base += nwaves * 64) {
    const int64_t v = base + lane;
    const uint32_t cv = v < G.n_owned ? G.cap[v] : kMidLen;
    const unsigned long long bs = __ballot(cv < kSmallLen);
    if (cv < kSmallLen) lst[__popcll(bs & ((1ull << lane) - 1ull))] = lane;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t ns = (uint32_t)__popcll(bs);
    for (uint32_t t = 0; t < ns; t += 8)
      if (t + (uint32_t)grp < ns) slots(G.row[base + lst[t + grp]], (uint32_t)sub, 8u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    unsigned long long bm = __ballot(cv >= kSmallLen && cv < kMidLen);
    while (bm) {
      const int src = __ffsll((long long)bm) - 1;
      bm &= bm - 1ull;
      slots(G.row[base + src], (uint32_t)lane, 64u);
    }
  }
}

// the same for the rows of kMidLen slots of capacity or more: a workgroup per
// entry of the hub chunk table (kChunk slots of one row)
__global__ __launch_bounds__(kBlock) void k_ccol_hubs(GraphView G, const int32_t* color,
                                                      const int2* hch, uint32_t n_hch) {
  for (uint32_t e = blockIdx.x; e < n_hch; e += gridDim.x) {
    const int2 hc = hch[e];
    const uint2 r = G.row[hc.x];
    const uint32_t hi = min(r.y, (uint32_t)(hc.y + 1) * kChunk);
    for (uint32_t j = (uint32_t)hc.y * kChunk + threadIdx.x; j < hi; j += kBlock) {
      const int32_t w = G.col[(uint64_t)r.x + j];
      G.ccol[(uint64_t)r.x + j] = w >= 0 ? color[w] : -1;
    }
  }
}

// Count slots breaking an invariant: ccol != colour of the neighbour, or a
// twin that does not point back; vertices whose cached mex (below the window
// size) is not the mex of their neighbours' colours, and rows with a
// histogram row iff their capacity is at least kSmallLen (tests only).
__global__ __launch_bounds__(kBlock) void k_check(GraphView G, const int32_t* color,
                                                  unsigned long long* bad) {
  __shared__ unsigned long long win[kBlock / 64][kWinWords];
  const int lane = lane_id();
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long* ww = win[threadIdx.x >> 6];
  unsigned long long cnt = 0;
  for (int64_t v = wave; v < G.n_owned; v += nwaves) {
    const uint2 r = G.row[v];
    const int32_t mc = G.mexc[v];
    if (lane == 0 && (G.cap[v] >= kSmallLen) != (G.hidx[v] >= 0)) ++cnt;
    if (mc >= 0 && mc < kWinBits) {
      ww[lane] = 0ull;
      ww[lane + 64] = 0ull;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (uint32_t j = lane; j < r.y; j += 64) {
        const int32_t w = G.col[(uint64_t)r.x + j];
        if (w < 0) continue;
        const int32_t c = color[w];
        if (c >= 0 && c < kWinBits) atomicOr(&ww[c >> 6], 1ull << (c & 63));
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      bool wrong = false;
      for (int wi = lane; wi < kWinWords; wi += 64) {
        const int32_t lo = wi * 64;
        const unsigned long long x = ww[wi];
        if (lo < mc) {
          const unsigned long long need = mc - lo >= 64 ? ~0ull : ((1ull << (mc - lo)) - 1ull);
          wrong |= (x & need) != need;
        }
        if (mc >= lo && mc < lo + 64) wrong |= ((x >> (mc - lo)) & 1ull) != 0ull;
      }
      const unsigned long long wb = __ballot(wrong);
      if (lane == 0 && wb) ++cnt;
      __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t j = lane; j < r.y; j += 64) {
      const uint64_t s = (uint64_t)r.x + j;
      const int32_t w = G.col[s];
      if (w < 0) {
        cnt += G.ccol[s] != -1;
        continue;
      }
      cnt += G.ccol[s] != color[w];
      const int32_t t = G.twin[s];
      if (w >= G.n_owned) {
        cnt += t != -1;
      } else {
        const uint2 rw = G.row[w];
        if (t < 0 || (uint32_t)t >= rw.y) {
          ++cnt;
        } else {
          const uint64_t s2 = (uint64_t)rw.x + (uint32_t)t;
          cnt += G.col[s2] != (int32_t)v;
          cnt += G.twin[s2] != (int32_t)j;
        }
      }
    }
  }
  wave_add_u64(cnt, bad);
}

// count vertices / chunks that can need the hub path (rows with cap >=
// kMidLen) and rows that keep a colour histogram (cap >= kSmallLen).  Launched
// with a small grid: one atomic per wave and counter.
// Capacity classes over all rows: {hub rows, their kChunk-slot chunks, rows
// with a histogram}.  Four loads in flight per lane, one atomic per
// workgroup and counter.
__global__ __launch_bounds__(kBlock) void k_hub_capacity(int32_t n, const uint32_t* cap,
                                                         uint32_t* out) {
  __shared__ uint32_t part[kBlock / 64][3];
  uint32_t hubs = 0, chunks = 0, hrows = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t v0 = blockIdx.x * blockDim.x + threadIdx.x; v0 < (uint32_t)n; v0 += 4 * stride) {
    uint32_t c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = v0 + k * stride < (uint32_t)n ? cap[v0 + k * stride] : 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (c[k] >= kMidLen) {
        hubs += 1;
        chunks += (c[k] + kChunk - 1) / kChunk;
      }
      hrows += c[k] >= kSmallLen ? 1u : 0u;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    hubs += __shfl_xor(hubs, off);
    chunks += __shfl_xor(chunks, off);
    hrows += __shfl_xor(hrows, off);
  }
  if (lane_id() == 0) {
    part[threadIdx.x >> 6][0] = hubs;
    part[threadIdx.x >> 6][1] = chunks;
    part[threadIdx.x >> 6][2] = hrows;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    uint32_t t = 0;
    for (int q = 0; q < kBlock / 64; ++q) t += part[q][threadIdx.x];
    if (t) atomicAdd(&out[threadIdx.x], t);
  }
}
constexpr int kCountGrid = 256;

// h = {hub rows, hub chunks, histogram rows} from the current capacities
static int hub_counts(Graph* g, uint32_t h[3]) {
  uint32_t* tmp = g->aux;
  DC_HIP(hipMemsetAsync(tmp, 0, sizeof(uint32_t) * 3, g->stream));
  hipLaunchKernelGGL(k_hub_capacity, dim3(kCountGrid), dim3(kBlock), 0, g->stream, g->n, g->cap,
                     tmp);
  DC_HIP(hipGetLastError());
  DC_HIP(hipMemcpyAsync(h, tmp, sizeof(uint32_t) * 3, hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  return DC_OK;
}

// frontier hub buffers for h (from hub_counts)
// (the staged batch's counters in g->ctr must survive this: it runs inside
// the apply when a row overflows)
static int ensure_hub_buffers(Graph* g, const uint32_t h[3]) {
  uint32_t need_hubs = h[0] + 1, need_chunks = h[1] + 1;
  if (need_hubs > g->hub_cap) g->hub_cap = need_hubs + need_hubs / 4 + 16;
  if (need_chunks > g->hchunk_cap) {
    if (g->hbuf) DC_HIP(hipFree(g->hbuf));
    if (g->partial) DC_HIP(hipFree(g->partial));
    g->hbuf = nullptr;
    g->partial = nullptr;
    uint32_t c = need_chunks + need_chunks / 4 + 16;
    DC_HIP(hipMalloc(&g->hbuf, sizeof(int4) * (size_t)c * 5));
    DC_HIP(hipMalloc(&g->partial, sizeof(unsigned long long) * 16 * (size_t)c));
    g->hchunk_cap = c;
    for (int i = 0; i < 3; ++i) g->F[i].h = g->hbuf + (size_t)c * i;
    g->T.h = g->hbuf + (size_t)c * 3;
    g->X.h = g->hbuf + (size_t)c * 4;
  }
  return DC_OK;
}

int alloc_state(Graph* g) {
  size_t n = (size_t)g->n + 1;
  DC_HIP(hipMalloc(&g->row, sizeof(uint2) * n));
  DC_HIP(hipMalloc(&g->cap, sizeof(uint32_t) * n));
  DC_HIP(hipMalloc(&g->prop, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&g->color[0], sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&g->color[1], sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&g->prev, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&g->tmark, sizeof(uint32_t) * n));
  DC_HIP(hipMalloc(&g->mmark, sizeof(uint32_t) * n));
  DC_HIP(hipMalloc(&g->pending, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&g->mexc, sizeof(int32_t) * n));
  DC_HIP(hipMemsetAsync(g->mexc, 0xff, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMalloc(&g->hidx, sizeof(int32_t) * n));
  DC_HIP(hipMemsetAsync(g->hidx, 0xff, sizeof(int32_t) * n, g->stream));
  g->cbits_words = n / 32 + 2;
  DC_HIP(hipMalloc(&g->cbits[0], sizeof(uint32_t) * 2 * g->cbits_words));
  g->cbits[1] = g->cbits[0] + g->cbits_words;
  DC_HIP(hipMemsetAsync(g->cbits[0], 0, sizeof(uint32_t) * 2 * g->cbits_words, g->stream));
  // 4 frontiers (F0, F1, F2=marked set, T) x 2 int lists (S, M)
  DC_HIP(hipMalloc(&g->fbuf, sizeof(int32_t) * n * 10));
  DC_HIP(hipMalloc(&g->fcnt, sizeof(uint32_t) * 20));
  for (int i = 0; i < 3; ++i) {
    g->F[i].s = g->fbuf + n * (2 * i);
    g->F[i].m = g->fbuf + n * (2 * i + 1);
    g->F[i].cnt = g->fcnt + 4 * i;
  }
  g->T.s = g->fbuf + n * 6;
  g->T.m = g->fbuf + n * 7;
  g->T.cnt = g->fcnt + 12;
  g->X.s = g->fbuf + n * 8;
  g->X.m = g->fbuf + n * 9;
  g->X.cnt = g->fcnt + 16;
  DC_HIP(hipMalloc(&g->ctr, sizeof(Counters)));
  DC_HIP(hipMalloc(&g->aux, sizeof(uint32_t) * 16));
  DC_HIP(hipMalloc(&g->tail, sizeof(unsigned long long)));
  DC_HIP(hipHostMalloc(&g->hctr, sizeof(Counters), hipHostMallocDefault));
  DC_HIP(hipMemsetAsync(g->row, 0, sizeof(uint2) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->cap, 0, sizeof(uint32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->prop, 0, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->color[0], 0, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->color[1], 0, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->prev, 0xff, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->tmark, 0, sizeof(uint32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->mmark, 0, sizeof(uint32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(g->fcnt, 0, sizeof(uint32_t) * 20, g->stream));
  DC_HIP(hipMemsetAsync(g->ctr, 0, sizeof(Counters), g->stream));
  for (int i = 0; i < 4; ++i) DC_HIP(hipEventCreate(&g->ev[i]));
  g->epoch = 0;
  return DC_OK;
}

void free_state(Graph* g) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  fused_free(g);
  loop_free(g);
  ldf_free(g);
  for (hipEvent_t e : g->evpool) (void)hipEventDestroy(e);
  g->evpool.clear();
  f(g->row); f(g->cap); f(g->col); f(g->prop); f(g->ccol); f(g->twin);
  f(g->color[0]); f(g->color[1]); f(g->prev);
  f(g->tmark); f(g->mmark); f(g->pending); f(g->cbits[0]); f(g->mexc);
  f(g->hidx); f(g->hist); f(g->hrowv); f(g->hch);
  f(g->fbuf); f(g->fcnt); f(g->hbuf); f(g->partial);
  f(g->del); f(g->ins); f(g->xdel); f(g->xins); f(g->ovf); f(g->rch);
  f(g->tkeys); f(g->tlong); f(g->sort_tmp);
  f(g->ctr); f(g->aux); f(g->tail); f(g->scan_tmp);
  f(g->scratch_batch.u); f(g->scratch_batch.v); f(g->scratch_batch.ins);
  if (g->hctr) (void)hipHostFree(g->hctr);
  if (g->pre_host) (void)hipHostFree(g->pre_host);
  f(g->vpart);
  f(g->pstate);
  g->vpart = nullptr;
  g->pstate = nullptr;
  g->nparts = 0;
  f(g->pre_dev);
  g->pre_dev = nullptr;
  g->pre_host = nullptr;
  for (int i = 0; i < 4; ++i)
    if (g->ev[i]) (void)hipEventDestroy(g->ev[i]);
}

// ------------------------------------------------------- hub histograms ----
// Histogram rows: cap >= kSmallLen.  hidx[v] = row of v's histogram or -1.
__global__ void k_hist_index(int32_t n, int32_t n_owned, const uint32_t* cap, int32_t* hidx,
                             int32_t* hrowv, uint32_t* counter) {
  const int lane = lane_id();
  // one atomic per wave for its histogram rows (the loop bound is the same
  // for every lane of a wave)
  for (int64_t b = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) & ~63ll; b < n;
       b += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = b + lane;
    const bool hub = v < n_owned && cap[v] >= kSmallLen;
    const unsigned long long bm = __ballot(hub);
    uint32_t base = 0u;
    if (lane == 0 && bm) base = atomicAdd(counter, (uint32_t)__popcll(bm));
    base = __shfl(base, 0);
    int32_t h = -1;
    if (hub) {
      h = (int32_t)(base + (uint32_t)__popcll(bm & ((1ull << lane) - 1ull)));
      hrowv[h] = (int32_t)v;
    }
    if (v < n) hidx[v] = h;
  }
}

// One workgroup per histogram row h0..n_hrows-1: count the live slots'
// colours below kHistW in LDS.
__global__ __launch_bounds__(kBlock) void k_hist_fill(GraphView G, const int32_t* hrowv,
                                                      uint32_t h0, uint32_t n_hrows) {
  __shared__ uint32_t cnt[kHistW];
  for (uint32_t h = h0 + blockIdx.x; h < n_hrows; h += gridDim.x) {
    for (int i = threadIdx.x; i < kHistW; i += blockDim.x) cnt[i] = 0u;
    __syncthreads();
    const uint2 r = G.row[hrowv[h]];
    for (uint32_t j = threadIdx.x; j < r.y; j += blockDim.x) {
      const int32_t c = G.ccol[(uint64_t)r.x + j];
      if (c >= 0 && c < kHistW) atomicAdd(&cnt[c], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistW; i += blockDim.x) G.hist[(uint64_t)h * kHistW + i] = cnt[i];
    __syncthreads();
  }
}

// the static hub chunk table: one entry per kChunk slots of a hub's capacity
// (a workgroup reserves its entries with one atomic: a prefix sum of its
// threads' chunk counts)
__global__ __launch_bounds__(kBlock) void k_hub_chunks(const int32_t* hrowv, uint32_t n_hrows,
                                                       const uint32_t* cap, int2* hch,
                                                       uint32_t* counter) {
  __shared__ uint32_t wsum[kBlock / 64];
  __shared__ uint32_t bbase;
  const int lane = lane_id();
  const int wib = threadIdx.x >> 6;
  for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n_hrows; b0 += gridDim.x * blockDim.x) {
    const uint32_t h = b0 + threadIdx.x;
    const int32_t v = h < n_hrows ? hrowv[h] : 0;
    const uint32_t cv = h < n_hrows ? cap[v] : 0u;
    const uint32_t nch = cv >= kMidLen ? (cv + kChunk - 1) / kChunk : 0u;
    uint32_t x = nch;   // inclusive scan over the wave
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wib] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int q = 0; q < kBlock / 64; ++q) t += wsum[q];
      bbase = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    uint32_t base = bbase + x - nch;
    for (int q = 0; q < wib; ++q) base += wsum[q];
    for (uint32_t c = 0; c < nch; ++c) hch[base + c] = make_int2(v, (int)c);
    __syncthreads();
  }
}

static int hub_chunk_table(Graph* g, uint32_t chunks);

static int hist_fill(Graph* g) {
  if (g->n_hrows == 0) return DC_OK;
  const uint32_t blocks = std::min<uint32_t>(g->n_hrows, 8192u);
  hipLaunchKernelGGL(k_hist_fill, dim3(blocks), dim3(kBlock), 0, g->stream, g->view(), g->hrowv, 0u,
                     g->n_hrows);
  DC_HIP(hipGetLastError());
  return DC_OK;
}

// histogram rows, the hub chunk table and (fill) the histograms for h (from
// hub_counts)
static int hist_reindex(Graph* g, const uint32_t h[3], bool fill = true) {
  uint32_t* tmp = g->aux;
  const uint32_t nh = h[2];
  if (nh > g->hist_cap) {
    if (g->hist) DC_HIP(hipFree(g->hist));
    if (g->hrowv) DC_HIP(hipFree(g->hrowv));
    g->hist = nullptr;
    g->hrowv = nullptr;
    const uint32_t c = nh + nh / 4 + 16;
    DC_HIP(hipMalloc(&g->hrowv, sizeof(int32_t) * (size_t)c));
    DC_HIP(hipMalloc(&g->hist, sizeof(uint32_t) * kHistW * (size_t)c));
    g->hist_cap = c;
  }
  DC_HIP(hipMemsetAsync(tmp + 4, 0, sizeof(uint32_t), g->stream));
  hipLaunchKernelGGL(k_hist_index, dim3(grid_for(g->n)), dim3(kBlock), 0, g->stream, g->n, g->n_owned,
                     g->cap, g->hidx, g->hrowv, tmp + 4);
  DC_HIP(hipGetLastError());
  DC_HIP(hipMemcpyAsync(&g->n_hrows, tmp + 4, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  int rc = hub_chunk_table(g, h[1]);
  if (rc) return rc;
  return fill ? hist_fill(g) : DC_OK;
}

int hist_reindex(Graph* g) {
  uint32_t h[3];
  int rc = hub_counts(g, h);
  if (rc) return rc;
  if ((rc = hist_reindex(g, h))) return rc;
  return ensure_hub_buffers(g, h);
}

// the hub chunk table from the current capacities (chunks = an upper bound
// of the entries, from k_hub_capacity)
static int hub_chunk_table(Graph* g, uint32_t chunks) {
  uint32_t* tmp = g->aux;
  if (chunks > g->hch_cap) {
    if (g->hch) DC_HIP(hipFree(g->hch));
    g->hch = nullptr;
    const uint32_t c = chunks + chunks / 4 + 16;
    DC_HIP(hipMalloc(&g->hch, sizeof(int2) * (size_t)c));
    g->hch_cap = c;
  }
  DC_HIP(hipMemsetAsync(tmp + 3, 0, sizeof(uint32_t), g->stream));
  if (g->n_hrows)
    hipLaunchKernelGGL(k_hub_chunks, dim3(grid_for(g->n_hrows)), dim3(kBlock), 0, g->stream, g->hrowv,
                       g->n_hrows, g->cap, g->hch, tmp + 3);
  DC_HIP(hipGetLastError());
  DC_HIP(hipMemcpyAsync(&g->n_hch, tmp + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  return DC_OK;
}

// tests: histogram entries that differ from a recount of the hub's row
__global__ __launch_bounds__(kBlock) void k_hist_check(GraphView G, const int32_t* hrowv,
                                                       uint32_t n_hrows, unsigned long long* bad) {
  __shared__ uint32_t cnt[kHistW];
  unsigned long long b = 0;
  for (uint32_t h = blockIdx.x; h < n_hrows; h += gridDim.x) {
    for (int i = threadIdx.x; i < kHistW; i += blockDim.x) cnt[i] = 0u;
    __syncthreads();
    const int32_t v = hrowv[h];
    const uint2 r = G.row[v];
    for (uint32_t j = threadIdx.x; j < r.y; j += blockDim.x) {
      const int32_t c = G.ccol[(uint64_t)r.x + j];
      if (c >= 0 && c < kHistW) atomicAdd(&cnt[c], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistW; i += blockDim.x)
      b += G.hist[(uint64_t)h * kHistW + i] != cnt[i];
    if (threadIdx.x == 0) b += G.hidx[v] != (int32_t)h;
    __syncthreads();
  }
  wave_add_u64(b, bad);
}

// Lay out rows from cap[] (already final) and allocate the slot arrays
// (col, ccol, twin: three int32 per slot).
static int layout_rows(Graph* g, int reset_used, int32_t** new_col, int32_t** new_ccol,
                       int32_t** new_twin) {
  uint64_t* start = nullptr;
  DC_HIP(hipMalloc(&start, sizeof(uint64_t) * ((size_t)g->n + 1)));
  int rc = exclusive_scan_u32_to_u64(g, g->cap, start, g->n);
  if (rc) {
    (void)hipFree(start);
    return rc;
  }
  uint64_t rows = 0;
  DC_HIP(hipMemcpyAsync(&rows, start + g->n, sizeof(uint64_t), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  // a spare tail of half the row slots (at least 1 Mi): rows that outgrow
  // their capacity move there one by one instead of relaying every row
  const uint64_t spare = g->opt.spare_slots > 0 ? (uint64_t)g->opt.spare_slots
                                                : std::max<uint64_t>(rows / 2, 1ull << 20);
  const uint64_t total = rows + spare;
  if (total >= 0xffffffffull) {
    (void)hipFree(start);
    set_error("graph exceeds 2^32 adjacency slots");
    return DC_ERR_NOMEM;
  }
  int32_t *col = nullptr, *ccol = nullptr, *twin = nullptr;
  hipError_t e = hipMalloc(&col, sizeof(int32_t) * (size_t)(total + 1));
  if (e == hipSuccess) e = hipMalloc(&ccol, sizeof(int32_t) * (size_t)(total + 1));
  if (e == hipSuccess) e = hipMalloc(&twin, sizeof(int32_t) * (size_t)(total + 1));
  if (e != hipSuccess) {
    (void)hipFree(start);
    if (col) (void)hipFree(col);
    if (ccol) (void)hipFree(ccol);
    return hip_fail(e, "hipMalloc(slot arrays)");
  }
  hipLaunchKernelGGL(k_set_rows, dim3(grid_for(g->n)), dim3(kBlock), 0, g->stream, g->n, start,
                     g->row, reset_used);
  DC_HIP(hipGetLastError());
  DC_HIP(hipFree(start));
  g->total_cap = total;
  g->rows_end = rows;
  const unsigned long long t0 = rows;
  DC_HIP(hipMemcpyAsync(g->tail, &t0, sizeof(t0), hipMemcpyHostToDevice, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  *new_col = col;
  *new_ccol = ccol;
  *new_twin = twin;
  return DC_OK;
}

int rebuild_ccol(Graph* g) {
  // every cached mex is stale once colours are replaced wholesale
  DC_HIP(hipMemsetAsync(g->mexc, 0xff, sizeof(int32_t) * ((size_t)g->n + 1), g->stream));
  hipLaunchKernelGGL(k_ccol_rows, dim3(grid_for(g->n_owned)), dim3(kBlock), 0, g->stream, g->view(),
                     g->color[g->cur]);
  if (g->n_hch)
    hipLaunchKernelGGL(k_ccol_hubs, dim3(std::min<uint32_t>(g->n_hch, 65535u)), dim3(kBlock), 0,
                       g->stream, g->view(), g->color[g->cur], g->hch, g->n_hch);
  DC_HIP(hipGetLastError());
  return hist_fill(g);
}

int check_consistency(Graph* g, int64_t* bad) {
  DC_HIP(hipMemsetAsync(g->aux, 0, sizeof(unsigned long long), g->stream));
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(g->aux);
  hipLaunchKernelGGL(k_check, dim3(grid_for((int64_t)g->n_owned * 64)), dim3(kBlock), 0, g->stream,
                     g->view(), g->color[g->cur], cnt);
  if (g->n_hrows)
    hipLaunchKernelGGL(k_hist_check, dim3(std::min<uint32_t>(g->n_hrows, 4096u)), dim3(kBlock), 0,
                       g->stream, g->view(), g->hrowv, g->n_hrows, cnt);
  DC_HIP(hipGetLastError());
  unsigned long long h = 0;
  DC_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  int64_t lbad = 0;
  int rc = loop_check(g, &lbad);
  if (rc) return rc;
  if (bad) *bad = (int64_t)h + lbad;
  return DC_OK;
}

int csr_build(Graph* g, int64_t m, const int32_t* du, const int32_t* dv) {
  // cap[] first holds the degree
  DC_HIP(hipMemsetAsync(g->cap, 0, sizeof(uint32_t) * ((size_t)g->n + 1), g->stream));
  if (m > 0)
    hipLaunchKernelGGL(k_count_degree, dim3(grid_for(m)), dim3(kBlock), 0, g->stream, m, du, dv,
                       g->n_owned, g->cap, g->prop);
  hipLaunchKernelGGL(k_set_cap, dim3(grid_for(g->n)), dim3(kBlock), 0, g->stream, g->n, g->n_owned,
                     g->cap, g->row, 0, g->opt.slack_frac, g->opt.slack_min);
  DC_HIP(hipGetLastError());
  int32_t *col = nullptr, *ccol = nullptr, *twin = nullptr;
  int rc = layout_rows(g, 1, &col, &ccol, &twin);
  if (rc) return rc;
  g->col = col;
  g->ccol = ccol;
  g->twin = twin;
  if (m > 0)
    hipLaunchKernelGGL(k_fill, dim3(grid_for(m)), dim3(kBlock), 0, g->stream, m, du, dv, g->view());
  DC_HIP(hipGetLastError());
  // hub rows and their chunk table first: the slot colours are written by them
  uint32_t h[3];
  if ((rc = hub_counts(g, h))) return rc;
  if ((rc = hist_reindex(g, h, false))) return rc;
  if ((rc = ensure_hub_buffers(g, h))) return rc;
  return rebuild_ccol(g);
}

// ------------------------------------------------------------- parts -------
// vertex -> part from the part offsets (binary search)
__global__ void k_part_map(int32_t n_owned, const int32_t* off, int32_t nparts, int32_t* vpart) {
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n_owned; v += gridDim.x * blockDim.x) {
    int32_t lo = 0, hi = nparts - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (off[mid] <= v) lo = mid;
      else hi = mid - 1;
    }
    vpart[v] = lo;
  }
}

// live slots whose two ends lie in different parts
__global__ void k_part_cross(GraphView G, const int32_t* vpart, unsigned long long* bad) {
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long b = 0;
  for (int64_t v0 = gwave; v0 < G.n_owned; v0 += nwaves) {
    const int32_t v = (int32_t)v0;
    const uint2 r = G.row[v];
    const int32_t pv = vpart[v];
    for (uint32_t j = lane; j < r.y; j += 64) {
      const int32_t w = G.col[(uint64_t)r.x + j];
      if (w >= 0 && vpart[w] != pv) ++b;
    }
  }
  wave_add_u64(b, bad);
}

// Declare the graph a union of disjoint parts (part p = owned vertices
// offsets[p] .. offsets[p+1]-1, no edge between parts).  Checks the offsets
// and that no live edge joins two parts.
int parts_set(Graph* g, int32_t nparts, const int32_t* offsets) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(g->vpart);
  f(g->pstate);
  g->vpart = nullptr;
  g->pstate = nullptr;
  g->nparts = 0;
  if (nparts == 0) return DC_OK;
  if (g->n_owned < g->n) {
    set_error("dc_set_parts: not for a partition share (ghost vertices)");
    return DC_ERR_INVALID;
  }
  if (offsets[0] != 0 || offsets[nparts] != g->n_owned) {
    set_error("dc_set_parts: offsets must run from 0 to the vertex count");
    return DC_ERR_INVALID;
  }
  for (int32_t p = 0; p < nparts; ++p)
    if (offsets[p + 1] <= offsets[p]) {
      set_error("dc_set_parts: every part needs at least one vertex, in increasing order");
      return DC_ERR_INVALID;
    }
  int32_t* doff = nullptr;
  DC_HIP(hipMalloc(&doff, sizeof(int32_t) * ((size_t)nparts + 1)));
  DC_HIP(hipMalloc(&g->vpart, sizeof(int32_t) * ((size_t)g->n_owned + 1)));
  DC_HIP(hipMalloc(&g->pstate, sizeof(uint32_t) * 4 * (size_t)nparts));
  DC_HIP(hipMemcpyAsync(doff, offsets, sizeof(int32_t) * ((size_t)nparts + 1), hipMemcpyHostToDevice,
                        g->stream));
  hipLaunchKernelGGL(k_part_map, dim3(grid_for(g->n_owned)), dim3(kBlock), 0, g->stream, g->n_owned,
                     doff, nparts, g->vpart);
  unsigned long long* bad = reinterpret_cast<unsigned long long*>(g->aux);
  DC_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned long long), g->stream));
  hipLaunchKernelGGL(k_part_cross, dim3(2048), dim3(kBlock), 0, g->stream, g->view(), g->vpart, bad);
  DC_HIP(hipGetLastError());
  unsigned long long h = 0;
  DC_HIP(hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  DC_HIP(hipFree(doff));
  if (h) {
    f(g->vpart);
    f(g->pstate);
    g->vpart = nullptr;
    g->pstate = nullptr;
    set_error("dc_set_parts: an edge joins two parts");
    return DC_ERR_INVALID;
  }
  g->nparts = nparts;
  return DC_OK;
}

// ------------------------------------------------------------- apply -------
// One thread per change line: append insertions to both rows (reserving a
// slot with an atomic on the row's used count), move the LDF counter, and
// route the line into the staged lists.  Single graph: an insertion is kept
// for the insertion round only when its endpoints had equal batch-start
// colors (HI/all_structure_undir.cuh:106-108, HB/all_structure_undir.cuh:
// 106-109).  Partition share: no pre-filter, cross edges (one ghost endpoint)
// go to their own lists as (owned, ghost) (MG/all_structure_undir.cuh:
// 100-188).
__global__ void k_apply(int64_t k, const int32_t* bu, const int32_t* bv, const uint8_t* bins,
                        GraphView G, const int32_t* color, int2* del, int2* ins, int2* xdel,
                        int2* xins, int4* ovf, Counters* ctr) {
  const bool partition = G.n_owned < G.n;
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < k;
       base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + threadIdx.x;
    bool ok = i < k;
    int32_t a = 0, b = 0;
    bool is_ins = false;
    if (ok) {
      a = bu[i];
      b = bv[i];
      is_ins = bins[i] != 0;
      if (a >= G.n_owned) {
        int32_t t = a;
        a = b;
        b = t;
      }
    }
    bool cross = ok && b >= G.n_owned;
    bool ovf_a = false, ovf_b = false;
    // both rows change (or may, for a deletion of an absent edge)
    if (ok) {
      G.mexc[a] = -1;
      if (!cross) G.mexc[b] = -1;
    }
    // an insertion between two parts of an arena ends their independence
    if (ok && is_ins && !cross && G.vpart && G.vpart[a] != G.vpart[b]) atomicOr(&ctr->xpart, 1u);
    // a new slot's colour enters its row's histogram (its slot may land
    // after a regrow, which recounts the histograms anyway)
    if (ok && is_ins) {
      const int32_t ha = G.hidx[a], cb = color[b];
      if (ha >= 0 && cb < kHistW) atomicAdd(&G.hist[(uint64_t)ha * kHistW + cb], 1u);
      if (!cross) {
        const int32_t hb = G.hidx[b], ca = color[a];
        if (hb >= 0 && ca < kHistW) atomicAdd(&G.hist[(uint64_t)hb * kHistW + ca], 1u);
      }
    }
    uint32_t pa = 0, pb = 0;
    if (ok && is_ins) {
      // reserve both slots first: each one records the other's position
      pa = atomicAdd(&G.row[a].y, 1u);
      if (!cross) pb = atomicAdd(&G.row[b].y, 1u);
      if (pa < G.cap[a]) {
        const uint64_t s = (uint64_t)G.row[a].x + pa;
        G.col[s] = b;
        G.ccol[s] = color[b];
        G.twin[s] = cross ? -1 : (int32_t)pb;
      } else {
        ovf_a = true;
      }
      if (!cross) {
        if (pb < G.cap[b]) {
          const uint64_t s = (uint64_t)G.row[b].x + pb;
          G.col[s] = a;
          G.ccol[s] = color[a];
          G.twin[s] = (int32_t)pa;
        } else {
          ovf_b = true;
        }
      }
      atomicAdd(&G.prop[a], 1);
      atomicAdd(&G.prop[b], 1);
    } else if (ok) {
      atomicSub(&G.prop[a], 1);
      atomicSub(&G.prop[b], 1);
    }
    // overflow entry: {row, slot position, neighbour, twin position}
    wave_append(ovf_a, make_int4(a, (int)pa, b, cross ? -1 : (int)pb), ovf, &ctr->overflow);
    wave_append(ovf_b, make_int4(b, (int)pb, a, (int)pa), ovf, &ctr->overflow);
    bool eq0 = ok && color[a] == color[b];
    int2 e = make_int2(a, b);
    wave_append(ok && !is_ins && !cross, e, del, &ctr->n_del);
    wave_append(ok && !is_ins && cross, e, xdel, &ctr->n_xdel);
    wave_append(ok && is_ins && !cross && (partition || eq0), e, ins, &ctr->n_ins);
    wave_append(ok && is_ins && cross, e, xins, &ctr->n_xins);
  }
}

// One wave per deletion line: tombstone every slot of the shorter row equal
// to the other endpoint together with its twin in the other row, i.e. every
// copy of the edge in both rows, including copies inserted by this batch
// (deleteEdgeFromAdj, HI/gpuFunctions_undir.cuh:118-142, scans both rows; the
// twin links make one scan of min(deg) slots enough).  A ghost endpoint has
// no row: the owned row is scanned.
// One kTombLanes-lane group per line, kTombUnroll slot loads in flight per
// lane: most scanned rows are short, and a wave per line would leave most
// lanes idle and keep one line's chain of dependent loads in flight instead
// of several.  Graphs with long rows take 16 lanes x 8 loads, the others 8 x
// 16 (measured, lanes x loads: R-MAT 22 with 50 000-line deletion batches,
// staging per batch 64x8 0.88 ms, 16x8 0.46, 8x16 0.47, 8x8 0.54, 4x16 0.71;
// a 512-graph arena of mostly short rows, 8.7 M deletion lines per step,
// step time 16x8 150 ms, 8x8 137, 8x16 137, 4x16 130; a quarter of R-MAT 24
// with 500 000-line deletion batches: 16x8 1 ms per step faster than 8x16).
template <int kTombLanes, int kTombUnroll>
__global__ void k_tombstone(const int2* del, uint32_t ndel, const int2* xdel, uint32_t nxdel,
                            GraphView G, Counters* ctr) {
  const int sub = lane_id() & (kTombLanes - 1);
  const int64_t grp = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kTombLanes;
  const int64_t ngrp = ((int64_t)gridDim.x * blockDim.x) / kTombLanes;
  const int64_t items = (int64_t)ndel + nxdel;
  unsigned long long tomb = 0, dslots = 0;
  for (int64_t it = grp; it < items; it += ngrp) {
    const bool cross = it >= (int64_t)ndel;
    const int2 e = cross ? xdel[it - ndel] : del[it];
    int32_t x = e.x, y = e.y;   // cross lines are (owned, ghost)
    uint2 r = G.row[x];
    uint2 ry = make_uint2(0u, 0u);
    if (!cross) {
      ry = G.row[y];
      if (ry.y < r.y) {
        const int32_t t = x;
        x = y;
        y = t;
        const uint2 u = r;
        r = ry;
        ry = u;
      }
    }
    if (sub == 0) dslots += (unsigned long long)r.y + ry.y;
    // Each slot is retired by the one thread that swaps its live value out
    // (the same edge can come twice, in either orientation), which also
    // takes its colour out of the row's histogram.
    const int32_t hx = G.hidx[x], hy = cross ? -1 : G.hidx[y];
    for (uint32_t j0 = sub; j0 < r.y; j0 += kTombUnroll * kTombLanes) {
      // kTombUnroll loads in flight per lane; matches are rare
      int32_t cv[kTombUnroll];
#pragma unroll
      for (int k = 0; k < kTombUnroll; ++k) {
        const uint32_t jj = j0 + k * kTombLanes;
        cv[k] = jj < r.y ? G.col[(uint64_t)r.x + jj] : -2;
      }
#pragma unroll
      for (int k = 0; k < kTombUnroll; ++k) {
        if (cv[k] != y) continue;
        const uint64_t s = (uint64_t)r.x + j0 + k * kTombLanes;
        if (atomicExch(&G.col[s], -1) == y) {
          const int32_t c = atomicExch(&G.ccol[s], -1);
          if (hx >= 0 && c >= 0 && c < kHistW) atomicSub(&G.hist[(uint64_t)hx * kHistW + c], 1u);
          ++tomb;
        }
        if (!cross) {
          const uint64_t s2 = (uint64_t)ry.x + (uint32_t)G.twin[s];
          if (atomicExch(&G.col[s2], -1) == x) {
            const int32_t c = atomicExch(&G.ccol[s2], -1);
            if (hy >= 0 && c >= 0 && c < kHistW) atomicSub(&G.hist[(uint64_t)hy * kHistW + c], 1u);
            ++tomb;
          }
        }
      }
    }
  }
  wave_add_u64(tomb, &ctr->tomb);
  wave_add_u64(dslots, &ctr->delslots);
}

// ---- deletions grouped by scanned row ----------------------------------------
// A row that many deletion lines scan (a hub of an R-MAT graph whose edges a
// batch deletes by the hundred) is read once for all of them instead of once
// per line: each line becomes the key (scanned row << 32 | other endpoint),
// the keys are sorted, and the first line of each row's run scans the row,
// matching every slot's neighbour against the run's sorted targets by binary
// search.  Runs over rows of kLongRow or more slots go to whole workgroups.
// The slots retired are those the per-line pass retires: every copy of each
// deleted pair in both rows.  The scanned row of a pair does not depend on the
// line's orientation (the shorter row, the lower ID on a tie), so every copy
// of a pair, and its twin, belongs to exactly one run: slots are retired with
// plain stores (only the histogram counts, which runs of different rows
// share, stay atomic).
constexpr uint32_t kLongRow = 4096;

__device__ __forceinline__ uint64_t tomb_key(int32_t x, int32_t y) {
  return ((uint64_t)(uint32_t)x << 32) | (uint32_t)y;
}

// scanned row and other endpoint of each line (the shorter row, the lower
// ID on a tie; the owned endpoint of a cross line)
__global__ void k_tomb_keys(const int2* del, uint32_t ndel, const int2* xdel, uint32_t nxdel,
                            GraphView G, uint64_t* keys, Counters* ctr) {
  const int64_t items = (int64_t)ndel + nxdel;
  unsigned long long dslots = 0;
  for (int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; it < items;
       it += (int64_t)gridDim.x * blockDim.x) {
    const bool cross = it >= (int64_t)ndel;
    const int2 e = cross ? xdel[it - ndel] : del[it];
    int32_t x = e.x, y = e.y;
    const uint32_t rx = G.row[x].y;
    uint32_t ry = 0;
    if (!cross) {
      ry = G.row[y].y;
      if (ry < rx || (ry == rx && y < x)) {
        x = e.y;
        y = e.x;
      }
    }
    dslots += (unsigned long long)rx + ry;
    keys[it] = tomb_key(x, y);
  }
  wave_add_u64(dslots, &ctr->delslots);
}

// a live slot of row x (owner hx) whose neighbour w is a target: retire it
// and its twin (the run of row x is their only writer)
__device__ __forceinline__ uint32_t tomb_retire(const GraphView& G, int32_t x, int32_t hx,
                                                uint64_t s, int32_t w) {
  const int32_t c = G.ccol[s];
  G.col[s] = -1;
  G.ccol[s] = -1;
  if (hx >= 0 && c >= 0 && c < kHistW) atomicSub(&G.hist[(uint64_t)hx * kHistW + c], 1u);
  if (w >= G.n_owned) return 1;
  const uint64_t s2 = (uint64_t)G.row[w].x + (uint32_t)G.twin[s];
  const int32_t c2 = G.ccol[s2];
  G.col[s2] = -1;
  G.ccol[s2] = -1;
  const int32_t hw = G.hidx[w];
  if (hw >= 0 && c2 >= 0 && c2 < kHistW) atomicSub(&G.hist[(uint64_t)hw * kHistW + c2], 1u);
  return 2;
}

// is w one of the run's targets (keys[lo, hi) of row x, sorted)?
__device__ __forceinline__ bool tomb_target(const uint64_t* keys, uint32_t lo, uint32_t hi,
                                            int32_t x, int32_t w) {
  const uint64_t k = tomb_key(x, w);
  const uint32_t end = hi;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo < end && keys[lo] == k;
}

// Runs of the sorted keys.  An 8-lane group owns 8 consecutive lines: its
// lanes flag the lines that start a run (a ballot gives the group's starts),
// then the group takes those runs in turn; a run that goes on past the 8
// lines is followed to its end (8 keys per step).  A short row (fewer than
// kMidRow slots) is read by the group in one round trip (8 lanes x 8
// consecutive slots); runs over longer rows go to k_tomb_mid (a wave per run)
// or k_tomb_long (a workgroup per run).
constexpr uint32_t kMidRow = 64;

__global__ __launch_bounds__(kBlock) void k_tomb_runs(const uint64_t* keys, uint32_t items,
                                                      GraphView G, int2* mids, uint32_t* nmid,
                                                      int2* longs, uint32_t* nlong, Counters* ctr) {
  constexpr int kL = 8;
  const int lane = lane_id();
  const int sub = lane & (kL - 1);
  const int gl0 = lane & ~(kL - 1);   // the group's first lane
  const uint64_t grp = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / kL;
  const uint64_t ngrp = ((uint64_t)gridDim.x * blockDim.x) / kL;
  unsigned long long tomb = 0;
  for (uint64_t base = grp * kL; base < items; base += ngrp * kL) {
    const uint64_t i = base + sub;
    const bool valid = i < items;
    const int64_t xi = valid ? (int64_t)(keys[i] >> 32) : -1;
    const int64_t xp = valid && i > 0 ? (int64_t)(keys[i - 1] >> 32) : -2;
    const bool start = valid && xp != xi;
    uint32_t m = (uint32_t)(__ballot(start) >> gl0) & 0xffu;
    while (m) {
      const int b = __ffs(m) - 1;
      m &= m - 1u;
      const int32_t x = (int32_t)__shfl((int)xi, gl0 + b);
      const uint32_t p = (uint32_t)(base + b);
      // the run's end: the next start in the group, else followed past it
      uint32_t end;
      if (m) {
        end = (uint32_t)base + (uint32_t)(__ffs(m) - 1);
      } else {
        end = (uint32_t)min<uint64_t>(base + kL, items);
        for (;;) {
          const uint32_t j = end + (uint32_t)sub;
          const bool same = j < items && (int32_t)(keys[j] >> 32) == x;
          uint32_t stop = same ? 0xffffffffu : j;
          for (int off = 1; off < kL; off <<= 1)
            stop = min(stop, (uint32_t)__shfl_xor((int)stop, off, kL));
          if (stop != 0xffffffffu) {
            end = stop;
            break;
          }
          end += kL;
        }
      }
      const uint2 r = G.row[x];
      if (r.y >= kMidRow) {
        if (sub == 0) {
          if (r.y >= kLongRow) longs[atomicAdd(nlong, 1u)] = make_int2((int32_t)p, (int32_t)end);
          else mids[atomicAdd(nmid, 1u)] = make_int2((int32_t)p, (int32_t)end);
        }
        continue;
      }
      const int32_t hx = G.hidx[x];
      const bool one = end == p + 1;
      const int32_t y0 = (int32_t)(uint32_t)keys[p];
      int32_t cv[kMidRow / kL];
#pragma unroll
      for (int q = 0; q < (int)(kMidRow / kL); ++q) {
        const uint32_t jj = (uint32_t)sub + q * kL;
        cv[q] = jj < r.y ? G.col[(uint64_t)r.x + jj] : -2;
      }
#pragma unroll
      for (int q = 0; q < (int)(kMidRow / kL); ++q) {
        if (cv[q] < 0) continue;
        if (one ? cv[q] != y0 : !tomb_target(keys, p, end, x, cv[q])) continue;
        tomb += tomb_retire(G, x, hx, (uint64_t)r.x + (uint32_t)sub + q * kL, cv[q]);
      }
    }
  }
  wave_add_u64(tomb, &ctr->tomb);
}

// runs over rows of kMidRow to kLongRow slots: one wave per run
__global__ __launch_bounds__(kBlock) void k_tomb_mid(const uint64_t* keys, const int2* mids,
                                                     const uint32_t* nmid, GraphView G,
                                                     Counters* ctr) {
  constexpr int kUnroll = 4;
  const int lane = lane_id();
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  unsigned long long tomb = 0;
  const uint32_t nm = *nmid;
  for (uint32_t q = gwave; q < nm; q += nwaves) {
    const int2 run = mids[q];
    const int32_t x = (int32_t)(keys[run.x] >> 32);
    const uint2 r = G.row[x];
    const int32_t hx = G.hidx[x];
    for (uint32_t j0 = lane; j0 < r.y; j0 += kUnroll * 64) {
      int32_t cv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint32_t jj = j0 + u * 64;
        cv[u] = jj < r.y ? G.col[(uint64_t)r.x + jj] : -2;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        if (cv[u] < 0 || !tomb_target(keys, (uint32_t)run.x, (uint32_t)run.y, x, cv[u])) continue;
        tomb += tomb_retire(G, x, hx, (uint64_t)r.x + j0 + u * 64, cv[u]);
      }
    }
  }
  wave_add_u64(tomb, &ctr->tomb);
}

// runs over long rows: one workgroup per run
__global__ __launch_bounds__(kBlock) void k_tomb_long(const uint64_t* keys, const int2* longs,
                                                      const uint32_t* nlong, GraphView G,
                                                      Counters* ctr) {
  constexpr int kUnroll = 8;
  unsigned long long tomb = 0;
  const uint32_t nl = *nlong;
  for (uint32_t q = blockIdx.x; q < nl; q += gridDim.x) {
    const int2 run = longs[q];
    const int32_t x = (int32_t)(keys[run.x] >> 32);
    const uint2 r = G.row[x];
    const int32_t hx = G.hidx[x];
    for (uint32_t j0 = threadIdx.x; j0 < r.y; j0 += kUnroll * kBlock) {
      int32_t cv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint32_t jj = j0 + u * kBlock;
        cv[u] = jj < r.y ? G.col[(uint64_t)r.x + jj] : -2;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        if (cv[u] < 0 || !tomb_target(keys, (uint32_t)run.x, (uint32_t)run.y, x, cv[u])) continue;
        tomb += tomb_retire(G, x, hx, (uint64_t)r.x + j0 + u * kBlock, cv[u]);
      }
    }
  }
  wave_add_u64(tomb, &ctr->tomb);
}

// regrow: copy every row's slots [0, min(used, old cap)) to the new layout
__global__ void k_copy_rows(int32_t n, const uint32_t* old_start, const uint32_t* old_cap,
                            const int32_t* old_col, const int32_t* old_ccol,
                            const int32_t* old_twin, const uint2* row, int32_t* col, int32_t* ccol,
                            int32_t* twin) {
  const int lane = lane_id();
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wave; v < n; v += nwaves) {
    uint32_t used = row[v].y;
    uint32_t c = old_cap[v];
    uint32_t lim = used < c ? used : c;
    uint64_t src = old_start[v], dst = row[v].x;
    for (uint32_t j = lane; j < lim; j += 64) {
      col[dst + j] = old_col[src + j];
      ccol[dst + j] = old_ccol[src + j];
      twin[dst + j] = old_twin[src + j];
    }
  }
}

__global__ void k_save_starts(int32_t n, const uint2* row, const uint32_t* cap, uint32_t* st,
                              uint32_t* oc) {
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    st[v] = row[v].x;
    oc[v] = cap[v];
  }
}

__global__ void k_fix_overflow(const int4* ovf, uint32_t novf, GraphView G, const int32_t* color) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < novf; i += gridDim.x * blockDim.x) {
    const int4 o = ovf[i];
    const uint64_t s = (uint64_t)G.row[o.x].x + (uint32_t)o.y;
    G.col[s] = o.z;
    G.ccol[s] = color[o.z];
    G.twin[s] = o.w;
  }
}

// Move each row that over-ran its capacity into the spare tail with room
// used + 2 x slack.  A thread per overflow entry: the first entry of a row
// claims it; a wave reserves its claimed rows' room at the tail with one
// atomic and lists their copies, one entry per kRelocChunk slots kept (one
// more atomic), for k_reloc_copy.  A row whose capacity crosses kSmallLen
// takes the next histogram row (filled by the caller once the row is whole).
// Twin links hold positions inside rows, so nothing else changes.
// flags[0]: the tail ran out (the caller relays every row instead);
// flags[1]: the histogram rows ran out (the caller re-indexes them all);
// flags[2]: histogram rows in use; flags[3]: copy entries listed.
constexpr uint32_t kRelocChunk = 2048;

__global__ __launch_bounds__(kBlock) void k_relocate(const int4* ovf, uint32_t novf, GraphView G,
                                                     uint32_t* mark, uint32_t epoch,
                                                     unsigned long long* tail,
                                                     unsigned long long tail_end, uint32_t* flags,
                                                     float frac, int32_t smin, int4* rch,
                                                     uint64_t rch_cap, int32_t* hrowv,
                                                     uint32_t hist_cap) {
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  // the loop bound is the same for every lane of a wave
  for (uint64_t b = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) & ~63ull; b < novf;
       b += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = b + lane;
    int32_t v = -1;
    if (i < novf) {
      v = ovf[i].x;
      if (!claim(mark, v, epoch)) v = -1;
    }
    uint2 r = make_uint2(0u, 0u);
    uint32_t oldcap = 0u, newcap = 0u;
    if (v >= 0) {
      r = G.row[v];
      oldcap = G.cap[v];
      newcap = r.y + 2 * slack_for(r.y, frac, smin);
    }
    // the wave's room at the tail: an inclusive scan of newcap
    unsigned long long x = newcap;
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned long long y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    unsigned long long base = 0ull;
    const unsigned long long tot = __shfl(x, 63);
    if (lane == 0 && tot) base = atomicAdd(tail, tot);
    const unsigned long long start = __shfl(base, 0) + x - newcap;
    const bool moved = v >= 0 && start + newcap <= tail_end;
    if (v >= 0 && !moved) flags[0] = 1u;
    const uint32_t keep = moved ? min(r.y, oldcap) : 0u;
    // copy entries
    const uint32_t nch = (keep + kRelocChunk - 1) / kRelocChunk;
    uint32_t cx = nch;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(cx, off);
      if (lane >= off) cx += y;
    }
    uint32_t cbase = 0u;
    const uint32_t ctot = __shfl(cx, 63);
    if (lane == 0 && ctot) cbase = atomicAdd(&flags[3], ctot);
    const uint64_t e0 = (uint64_t)__shfl(cbase, 0) + cx - nch;
    for (uint32_t c = 0; c < nch && e0 + c < rch_cap; ++c)
      rch[e0 + c] = make_int4((int)(r.x + c * kRelocChunk), (int)(uint32_t)(start + c * kRelocChunk),
                              (int)min(kRelocChunk, keep - c * kRelocChunk), 0);
    // histogram rows for the rows crossing kSmallLen
    const bool cross = moved && oldcap < kSmallLen && newcap >= kSmallLen;
    const unsigned long long bh = __ballot(cross);
    uint32_t hbase = 0u;
    if (lane == 0 && bh) hbase = atomicAdd(&flags[2], (uint32_t)__popcll(bh));
    hbase = __shfl(hbase, 0);
    if (cross) {
      const uint32_t h = hbase + (uint32_t)__popcll(bh & below);
      if (h < hist_cap) {
        hrowv[h] = v;
        G.hidx[v] = (int32_t)h;
      } else {
        flags[1] = 1u;
      }
    }
    if (moved) {
      G.row[v].x = (uint32_t)start;
      G.cap[v] = newcap;
    }
  }
}

// the copies listed by k_relocate: a wave per entry of at most kRelocChunk slots
__global__ __launch_bounds__(kBlock) void k_reloc_copy(GraphView G, const int4* rch,
                                                       const uint32_t* nrch, uint64_t rch_cap) {
  const int lane = lane_id();
  const uint64_t ne = min((uint64_t)*nrch, rch_cap);
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t e = wave; e < ne; e += nwaves) {
    const int4 c = rch[e];
    const uint64_t from = (uint32_t)c.x, to = (uint32_t)c.y;
    for (uint32_t j = lane; j < (uint32_t)c.z; j += 64) {
      G.col[to + j] = G.col[from + j];
      G.ccol[to + j] = G.ccol[from + j];
      G.twin[to + j] = G.twin[from + j];
    }
  }
}

static int regrow(Graph* g, uint32_t novf);

// Make room for a batch's overflowing insertions: relocate the rows into the
// spare tail, or relay every row when the tail is used up.
static int make_room(Graph* g, uint32_t novf, bool* relaid) {
  *relaid = false;
  // copy entries: at most one per claimed row plus one per kRelocChunk slots
  const uint64_t need = (uint64_t)novf + g->total_cap / kRelocChunk + 64;
  if (need > g->rch_cap) {
    if (g->rch) DC_HIP(hipFree(g->rch));
    g->rch = nullptr;
    g->rch_cap = 0;
    const uint64_t c = need + need / 4;
    DC_HIP(hipMalloc(&g->rch, sizeof(int4) * (size_t)c));
    g->rch_cap = c;
  }
  uint32_t* flags = g->aux + 8;
  DC_HIP(hipMemsetAsync(flags, 0, sizeof(uint32_t) * 4, g->stream));
  DC_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(flags + 2), (int)g->n_hrows, 1,
                           g->stream));
  int rc = ensure_epochs(g, 1);
  if (rc) return rc;
  const uint32_t e = next_epoch(g);
  hipLaunchKernelGGL(k_relocate, dim3(grid_for(novf)), dim3(kBlock), 0, g->stream, g->ovf, novf,
                     g->view(), g->tmark, e, g->tail, (unsigned long long)g->total_cap, flags,
                     g->opt.slack_frac, g->opt.slack_min, g->rch, g->rch_cap, g->hrowv, g->hist_cap);
  hipLaunchKernelGGL(k_reloc_copy, dim3(std::min<uint64_t>((need + 3) / 4, 8192)), dim3(kBlock), 0,
                     g->stream, g->view(), g->rch, flags + 3, g->rch_cap);
  DC_HIP(hipGetLastError());
  uint32_t hf[4];
  DC_HIP(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  if (hf[3] > g->rch_cap) {   // cannot happen with the bound above
    set_error("row relocation copy list overflow");
    return DC_ERR_NOMEM;
  }
  if (hf[0]) {
    *relaid = true;
    return regrow(g, novf);
  }
  hipLaunchKernelGGL(k_fix_overflow, dim3(grid_for(novf)), dim3(kBlock), 0, g->stream, g->ovf, novf,
                     g->view(), g->color[g->cur]);
  DC_HIP(hipGetLastError());
  // capacities grew: frontier hub buffers and the hub chunk table follow; a
  // row past kSmallLen got a histogram row above, filled here once its
  // overflow slots are in (the existing ones already count the batch)
  uint32_t h[3];
  if ((rc = hub_counts(g, h))) return rc;
  if (hf[1]) {
    if ((rc = hist_reindex(g, h))) return rc;
  } else {
    if (hf[2] > g->n_hrows) {
      const uint32_t blocks = std::min<uint32_t>(hf[2] - g->n_hrows, 8192u);
      hipLaunchKernelGGL(k_hist_fill, dim3(blocks), dim3(kBlock), 0, g->stream, g->view(), g->hrowv,
                         g->n_hrows, hf[2]);
      DC_HIP(hipGetLastError());
      g->n_hrows = hf[2];
    }
    if ((rc = hub_chunk_table(g, h[1]))) return rc;
  }
  return ensure_hub_buffers(g, h);
}

// Give every row that over-ran its capacity new room (used + slack), relay
// all rows, then write the slots that did not fit.
static int regrow(Graph* g, uint32_t novf) {
  uint32_t *old_start = nullptr, *old_cap = nullptr;
  DC_HIP(hipMalloc(&old_start, sizeof(uint32_t) * ((size_t)g->n + 1)));
  DC_HIP(hipMalloc(&old_cap, sizeof(uint32_t) * ((size_t)g->n + 1)));
  hipLaunchKernelGGL(k_save_starts, dim3(grid_for(g->n)), dim3(kBlock), 0, g->stream, g->n, g->row,
                     g->cap, old_start, old_cap);
  hipLaunchKernelGGL(k_set_cap, dim3(grid_for(g->n)), dim3(kBlock), 0, g->stream, g->n, g->n_owned,
                     g->cap, g->row, 1, g->opt.slack_frac, g->opt.slack_min);
  DC_HIP(hipGetLastError());
  int32_t *col = nullptr, *ccol = nullptr, *twin = nullptr;
  int rc = layout_rows(g, 0, &col, &ccol, &twin);
  if (rc) {
    (void)hipFree(old_start);
    (void)hipFree(old_cap);
    return rc;
  }
  hipLaunchKernelGGL(k_copy_rows, dim3(grid_for((int64_t)g->n * 64)), dim3(kBlock), 0, g->stream,
                     g->n, old_start, old_cap, g->col, g->ccol, g->twin, g->row, col, ccol, twin);
  DC_HIP(hipGetLastError());
  DC_HIP(hipStreamSynchronize(g->stream));
  DC_HIP(hipFree(g->col));
  DC_HIP(hipFree(g->ccol));
  DC_HIP(hipFree(g->twin));
  DC_HIP(hipFree(old_start));
  DC_HIP(hipFree(old_cap));
  g->col = col;
  g->ccol = ccol;
  g->twin = twin;
  hipLaunchKernelGGL(k_fix_overflow, dim3(grid_for(novf)), dim3(kBlock), 0, g->stream, g->ovf, novf,
                     g->view(), g->color[g->cur]);
  DC_HIP(hipGetLastError());
  // capacities changed: hub rows and their histograms are recomputed
  return hist_reindex(g);
}

static int ensure_lists(Graph* g, int64_t k) {
  if (k <= g->list_cap) return DC_OK;
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(g->del); f(g->ins); f(g->xdel); f(g->xins); f(g->ovf);
  g->del = g->ins = g->xdel = g->xins = nullptr;
  g->ovf = nullptr;
  int64_t c = k + k / 4 + 64;
  DC_HIP(hipMalloc(&g->del, sizeof(int2) * (size_t)c));
  DC_HIP(hipMalloc(&g->ins, sizeof(int2) * (size_t)c));
  DC_HIP(hipMalloc(&g->xdel, sizeof(int2) * (size_t)c));
  DC_HIP(hipMalloc(&g->xins, sizeof(int2) * (size_t)c));
  DC_HIP(hipMalloc(&g->ovf, sizeof(int4) * (size_t)(2 * c)));
  g->list_cap = c;
  // grouped deletions: keys, sorted keys, long-row runs, radix-sort storage
  f(g->tkeys); f(g->tlong); f(g->sort_tmp);
  g->tkeys = nullptr;
  g->tlong = nullptr;
  g->sort_tmp = nullptr;
  DC_HIP(hipMalloc(&g->tkeys, sizeof(uint64_t) * 2 * (size_t)c));
  DC_HIP(hipMalloc(&g->tlong, sizeof(int2) * 2 * (size_t)c));   // mid runs, then long runs
  size_t bytes = 0;
  DC_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, g->tkeys, g->tkeys + c, (int)c, 0, 63,
                                           g->stream));
  DC_HIP(hipMalloc(&g->sort_tmp, bytes));
  g->sort_tmp_bytes = bytes;
  return DC_OK;
}

int csr_apply(Graph* g, const Batch& b, int rule) {
  (void)rule;
  int rc = ensure_lists(g, b.k);
  if (rc) return rc;
  DC_HIP(hipMemsetAsync(g->ctr, 0, sizeof(Counters), g->stream));
  if (b.k > 0)
    hipLaunchKernelGGL(k_apply, dim3(grid_for(b.k)), dim3(kBlock), 0, g->stream, b.k, b.u, b.v,
                       b.ins, g->view(), g->color[g->cur], g->del, g->ins, g->xdel, g->xins, g->ovf,
                       g->ctr);
  DC_HIP(hipGetLastError());
  DC_HIP(hipMemcpyAsync(g->hctr, g->ctr, sizeof(Counters), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  Counters c = *g->hctr;
  // an edge between parts: they are no longer independent, the loop stops
  // treating them apart
  if (c.xpart && g->nparts > 0) g->nparts = 0;
  g->staged_n[0] = c.n_del;
  g->staged_n[1] = c.n_ins;
  g->staged_n[2] = c.n_xdel;
  g->staged_n[3] = c.n_xins;
  g->staged_regrows = 0;
  g->staged_moves = 0;
  if (c.overflow > 0) {
    bool relaid = false;
    rc = make_room(g, c.overflow, &relaid);
    if (rc) return rc;
    g->staged_regrows = relaid ? 1 : 0;
    g->staged_moves = 1;
  }
  uint32_t items = c.n_del + c.n_xdel;
  // row slots (with slack) per owned vertex: 32 or more -> wide groups
  const bool wide = g->rows_end >= 32ull * (uint64_t)std::max(g->n_owned, 1);
  const int32_t gd = g->tuning.group_deletions != 0 ? g->tuning.group_deletions
                                                   : g->opt.group_deletions;
  const uint32_t group_at = gd == 0 ? (1u << 16) : gd < 0 ? 0xffffffffu : (uint32_t)gd;
  if (items > 0 && items >= group_at) {
    // many lines: grouped by scanned row (each row read once)
    uint64_t* keys = g->tkeys;
    uint64_t* sorted = g->tkeys + g->list_cap;
    uint32_t* nlong = g->aux + 12;
    DC_HIP(hipMemsetAsync(nlong, 0, sizeof(uint32_t), g->stream));
    hipLaunchKernelGGL(k_tomb_keys, dim3(grid_for(items)), dim3(kBlock), 0, g->stream, g->del, c.n_del,
                       g->xdel, c.n_xdel, g->view(), keys, g->ctr);
    DC_HIP(hipGetLastError());
    size_t bytes = g->sort_tmp_bytes;
    DC_HIP(hipcub::DeviceRadixSort::SortKeys(g->sort_tmp, bytes, keys, sorted, (int)items, 0, 63,
                                             g->stream));
    uint32_t* nmid = g->aux + 13;
    DC_HIP(hipMemsetAsync(nmid, 0, sizeof(uint32_t), g->stream));
    hipLaunchKernelGGL(k_tomb_runs, dim3(grid_for(((int64_t)items + 7) / 8 * 8)), dim3(kBlock), 0,
                       g->stream, sorted, items,
                       g->view(), g->tlong, nmid, g->tlong + g->list_cap, nlong, g->ctr);
    hipLaunchKernelGGL(k_tomb_mid, dim3(2048), dim3(kBlock), 0, g->stream, sorted, g->tlong, nmid,
                       g->view(), g->ctr);
    hipLaunchKernelGGL(k_tomb_long, dim3(1024), dim3(kBlock), 0, g->stream, sorted,
                       g->tlong + g->list_cap, nlong, g->view(), g->ctr);
    DC_HIP(hipGetLastError());
  } else if (items > 0) {
    if (wide)
      hipLaunchKernelGGL((k_tombstone<16, 8>), dim3(grid_for((int64_t)items * 16)), dim3(kBlock), 0,
                         g->stream, g->del, c.n_del, g->xdel, c.n_xdel, g->view(), g->ctr);
    else
      hipLaunchKernelGGL((k_tombstone<8, 16>), dim3(grid_for((int64_t)items * 8)), dim3(kBlock), 0,
                         g->stream, g->del, c.n_del, g->xdel, c.n_xdel, g->view(), g->ctr);
    DC_HIP(hipGetLastError());
  }
  g->staged = true;
  return DC_OK;
}

// ------------------------------------------------------------ compact ------
// In-place compaction of each row's live slots (ballot/mbcnt within the wave;
// writes never pass the slot being read).
// (slots move together with their ccol and twin values; newpos records each
// old slot's new position so the twins can be re-pointed afterwards)
__global__ void k_compact_rows(GraphView G, int32_t* newpos) {
  const int lane = lane_id();
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wave; v < G.n_owned; v += nwaves) {
    uint2 r = G.row[v];
    uint32_t out = 0;
    for (uint32_t j0 = 0; j0 < r.y; j0 += 64) {
      const uint32_t j = j0 + lane;
      const uint64_t s = (uint64_t)r.x + j;
      const int32_t w = j < r.y ? G.col[s] : -1;
      const int32_t c = j < r.y ? G.ccol[s] : -1;
      const int32_t t = j < r.y ? G.twin[s] : -1;
      const bool live = w >= 0;
      const unsigned long long m = __ballot(live);
      const uint32_t np = out + lanes_below(m);
      if (j < r.y) newpos[s] = live ? (int32_t)np : -1;
      if (live) {
        G.col[(uint64_t)r.x + np] = w;
        G.ccol[(uint64_t)r.x + np] = c;
        G.twin[(uint64_t)r.x + np] = t;
      }
      out += (uint32_t)__popcll(m);
    }
    if (lane == 0) G.row[v].y = out;
  }
}

__global__ void k_remap_twins(GraphView G, const int32_t* newpos) {
  const int lane = lane_id();
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t v = wave; v < G.n_owned; v += nwaves) {
    const uint2 r = G.row[v];
    for (uint32_t j = lane; j < r.y; j += 64) {
      const uint64_t s = (uint64_t)r.x + j;
      const int32_t w = G.col[s];
      if (w < 0 || w >= G.n_owned) continue;
      G.twin[s] = newpos[(uint64_t)G.row[w].x + (uint32_t)G.twin[s]];
    }
  }
}

static int compact_with(Graph* g, int32_t* newpos) {
  const dim3 grid(grid_for((int64_t)g->n_owned * 64));
  hipLaunchKernelGGL(k_compact_rows, grid, dim3(kBlock), 0, g->stream, g->view(), newpos);
  hipLaunchKernelGGL(k_remap_twins, grid, dim3(kBlock), 0, g->stream, g->view(), newpos);
  hipError_t e = hipGetLastError();
  hipError_t s = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) return hip_fail(e, "compact");
  if (s != hipSuccess) return hip_fail(s, "compact");
  return DC_OK;
}

int csr_compact(Graph* g) {
  int32_t* newpos = nullptr;
  DC_HIP(hipMalloc(&newpos, sizeof(int32_t) * (size_t)(g->total_cap + 1)));
  const int rc = compact_with(g, newpos);
  (void)hipFree(newpos);
  return rc;
}

// ---- compaction without allocation -------------------------------------------
// The slot map of the compaction (old slot -> new position) in a file-scope
// device array instead of a per-call allocation, for callers that cannot
// allocate device memory.  Each GPU holds its own copy for the life of the
// process; every call on that GPU shares it, so calls that overlap on one GPU
// race.  Nothing in it needs to start at zero: k_compact_rows writes the entry
// of every used slot of every owned row before k_remap_twins reads any (a
// twin always points at a used slot of an owned row).
constexpr uint64_t kScratchSlots = 1ull << 23;   // 32 MiB of int32
static __device__ __align__(256) int32_t scratch[kScratchSlots];
constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static int32_t* scratch_dev[kMaxDevices] = {};   // the array's address on each device

uint64_t compact_noalloc_max_slots() { return kScratchSlots - 1; }

int csr_compact_noalloc(Graph* g) {
  if (g->total_cap + 1 > kScratchSlots) {
    set_error("dc_compact_noalloc: the graph has " + std::to_string(g->total_cap) +
              " row slots; the static slot map holds at most " +
              std::to_string(kScratchSlots - 1));
    return DC_ERR_INVALID;
  }
  if (g->device < 0 || g->device >= kMaxDevices) {
    set_error("dc_compact_noalloc: device ordinal out of range");
    return DC_ERR_INVALID;
  }
  int32_t* ptr = nullptr;
  {
    // once per device, with that device current (the caller set it): the
    // host-side name of the array is not a device address
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[g->device]) {
      void* p = nullptr;
      DC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      scratch_dev[g->device] = static_cast<int32_t*>(p);
    }
    ptr = scratch_dev[g->device];
  }
  return compact_with(g, ptr);
}

}  // namespace dc
