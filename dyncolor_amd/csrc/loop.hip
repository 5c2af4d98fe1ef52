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
// rose (R).  A mex comes from the per-vertex cache (mexc) unless a row edit
// or a neighbour's colour change made it stale.  Colours, traces and round
// counts are those of the full re-derivation.
//
// One round = two phases:
//   A(r): filter the candidates of round r-1 by K > 0 into C_r and evaluate
//         each member at once (cached mex, or a scan of its row by an 8-lane
//         group / a wave / a whole workgroup for hubs), commit; members of
//         L_r = C_{r-1} that dropped out join D_r.
//   C(r): counts over D_r's rows, the new colours of X_r (changed vertices)
//         into the slot cache, and L_r's bitmap cleared.  The recount reads
//         the colour of a C_r member from the colour array, so it does not
//         wait for the slot cache.
// The plain driver launches one kernel per phase (the host enqueues rounds in
// blocks and reads one control word per block; rounds past the last one
// return at once); the persistent driver runs both phases of every round in
// one cooperative launch with grid barriers between phases.
//
// Outside the loop K is zero everywhere, and so are the bitmaps (C/L members
// in B[0..1], D members in Db[0..1], R claims in Rb).
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "dc_phases.h"

namespace dc {

struct LoopCtl {
  int32_t done_round;     // the round whose C was empty (INT_MAX while running)
  uint32_t pad[3];
  unsigned long long rslots, rlive;   // rows scanned for a stale mex
  unsigned long long uslots, ulive;   // rows scanned by the count pass
};

// Everything a loop kernel needs, by value; per-round list roles are derived
// from the round number (lists rotate over three buffers).
struct LoopState {
  GraphView G;
  int32_t* col0;       // colour buffers: both are written, so they agree at the end
  int32_t* col1;
  int32_t* prev;
  int32_t* K;
  int2* dold;          // D members: {colour, (prev + 1) | in_L << 31} before the round
  uint32_t* B[2];      // C_r members in B[r & 1] (so L_r = C_{r-1} in the other)
  uint32_t* Db[2];     // D_r members in Db[r & 1]
  uint32_t* Rb;        // R claims
  Frontier C[3], D[3], X[3], R[3];
  Frontier L0;         // the marked set (plays C_{-1})
  uint32_t hub_cap;
  int4* trace;         // per round {|L|, |C|, changed, 0}
  LoopCtl* ctl;
};

struct IncState {
  LoopState S{};
  int32_t* lbuf = nullptr;
  int4* hb = nullptr;
  uint32_t hcap = 0;
  uint32_t* cnt = nullptr;
  uint32_t* bits = nullptr;   // Db[0], Db[1], Rb
  size_t words = 0;
  int32_t trace_cap = 0;
  LoopCtl* hctl = nullptr;
};

__device__ __forceinline__ bool has_bit(const uint32_t* b, int32_t v) {
  return (b[v >> 5] >> (v & 31)) & 1u;
}
__device__ __forceinline__ void set_bit(uint32_t* b, int32_t v) {
  atomicOr(&b[v >> 5], 1u << (v & 31));
}

__device__ __forceinline__ uint32_t fsize_dev(const Frontier& F) { return F.cnt[0] + F.cnt[1] + F.cnt[3]; }

// The marking predicate of findEligibleNeighbors for a slot (v -> w).
__device__ __forceinline__ bool marks(int32_t v, int32_t cv, int32_t pv, int32_t w, int32_t cw) {
  return (cv == cw && v < w) || pv < cw;
}

__device__ __forceinline__ Frontier c_list(const LoopState& S, int r) {
  return r < 0 ? S.L0 : S.C[r % 3];
}
__device__ __forceinline__ int rot(int r) { return (r + 3) % 3; }

// Rows of a frontier: fn(v, row, lo, hi, stride, width) for each lane's
// share; `width` lanes share (v, this call) for reductions (8 for S groups,
// 64 for M rows and for each wave of an H chunk).  Idle S lanes get v = -1.
template <class Fn>
__device__ __forceinline__ void visit_rows(const GraphView& G, const Frontier& F, Fn fn) {
  const int lane = lane_id();
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t nS = F.cnt[0], nM = F.cnt[1], nH = F.cnt[2];
  for (uint32_t ci = blockIdx.x; ci < nH; ci += gridDim.x) {
    const int4 it = F.h[ci];
    const uint2 r = G.row[it.x];
    const uint32_t lo = (uint32_t)it.y * kChunk;
    fn(it.x, r, lo + threadIdx.x, min(r.y, lo + kChunk), (uint32_t)kBlock, 64);
  }
  for (int64_t i = gwave; i < nM; i += nwaves) {
    const int32_t v = F.m[i];
    const uint2 r = G.row[v];
    fn(v, r, (uint32_t)lane, r.y, 64u, 64);
  }
  const int grp = lane >> 3, sub = lane & 7;
  for (int64_t b = gwave * 8; b < nS; b += nwaves * 8) {
    const int64_t idx = b + grp;
    const bool act = idx < nS;
    const int32_t v = act ? F.s[idx] : -1;
    const uint2 r = act ? G.row[v] : make_uint2(0u, 0u);
    fn(v, r, (uint32_t)sub, r.y, (uint32_t)kGroup, 8);
  }
}

// Each member of a frontier once (hubs through their first chunk entry), one
// thread per list entry; fn(v, valid) is called by whole waves.
template <class Fn>
__device__ __forceinline__ void visit_members(const Frontier& F, Fn fn) {
  const uint32_t nS = F.cnt[0], nM = F.cnt[1], nH = F.cnt[2];
  const uint32_t tot = nS + nM + nH;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < tot; base += stride) {
    const uint32_t i = base + threadIdx.x;
    int32_t v = -1;
    if (i < nS) {
      v = F.s[i];
    } else if (i < nS + nM) {
      v = F.m[i - nS];
    } else if (i < tot) {
      const int4 it = F.h[i - nS - nM];
      if (it.y == 0) v = it.x;
    }
    fn(v, v >= 0);
  }
}

__device__ __forceinline__ uint32_t group_sum(uint32_t x, int width) {
  for (int off = width >> 1; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

// ---------------------------------------------------------------------------
// set-up: L0's bits, then every marking slot of L0's rows counted into K; the
// marked vertices outside L0 become the candidates R_{-1}.
__device__ __forceinline__ void init_bits(const LoopState& S) {
  visit_members(S.L0, [&](int32_t v, bool valid) {
    if (valid) set_bit(S.B[1], v);
  });
}

__device__ __forceinline__ void init_push(const LoopState& S) {
  const GraphView& G = S.G;
  const Frontier Rn = S.R[2];
  visit_rows(G, S.L0, [&](int32_t v, uint2 r, uint32_t lo, uint32_t hi, uint32_t stride, int width) {
    (void)width;
    const int32_t cv = v >= 0 ? S.col0[v] : 0;
    const int32_t pv = v >= 0 ? S.prev[v] : 0;
    for (uint32_t j = lo; j < hi; j += stride) {
      const int32_t w = G.col[(uint64_t)r.x + j];
      const int32_t cw = G.ccol[(uint64_t)r.x + j];
      const bool t = w >= 0 && w < G.n_owned && marks(v, cv, pv, w, cw);
      if (t) atomicAdd(&S.K[w], 1);
      const bool push = t && !has_bit(S.B[1], w) && claim_bit(S.Rb, w);
      frontier_push(push, w, G, Rn, S.hub_cap);
    }
  });
}

// ---------------------------------------------------------------------------
// phase A
//
// Pushes go to per-workgroup queues in LDS and are appended to the global
// lists once per workgroup at the end of the phase (one global atomic per
// list and class per wave of the flush, instead of one per vertex).
constexpr int kQCap = 1024;
struct BlockQueues {
  int32_t v[3][kQCap];   // 0: C_r, 1: X_r, 2: D_r
  uint32_t n[3];
};

// Append a wave's pushes to a frontier; every lane of the wave must be
// active.  Hub chunk entries are reserved with one atomic per wave.
__device__ __forceinline__ void push_converged(bool p, int32_t v, const GraphView& G, Frontier F,
                                               uint32_t hub_cap) {
  const int lane = lane_id();
  const uint32_t len = p ? G.row[v].y : 0u;
  wave_append(p && len < kSmallLen, v, F.s, &F.cnt[0]);
  wave_append(p && len >= kSmallLen && len < kMidLen, v, F.m, &F.cnt[1]);
  const bool is_h = p && len >= kMidLen;
  const unsigned long long hm = __ballot(is_h);
  if (hm == 0ull) return;
  const uint32_t nch = is_h ? (len + kChunk - 1) / kChunk : 0u;
  uint32_t x = nch;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  const uint32_t total = __shfl(x, 63);
  const int leader = __ffsll((long long)hm) - 1;
  uint32_t sbase = 0, cbase = 0;
  if (lane == leader) {
    sbase = atomicAdd(&F.cnt[3], (uint32_t)__popcll(hm));
    cbase = atomicAdd(&F.cnt[2], total);
  }
  sbase = __shfl(sbase, leader);
  cbase = __shfl(cbase, leader);
  if (is_h) {
    const uint32_t slot = sbase + lanes_below(hm);
    const uint32_t base = cbase + x - nch;
    if (slot < hub_cap)
      for (uint32_t c = 0; c < nch; ++c) F.h[base + c] = make_int4(v, (int)c, (int)slot, (int)nch);
  }
}

// Queue a push (wave-level call, divergence allowed); a full queue falls back
// to a direct append.
__device__ __forceinline__ void q_push(BlockQueues& Q, int qi, bool p, int32_t v, const GraphView& G,
                                       const Frontier& F, uint32_t hub_cap) {
  bool direct = false;
  if (p) {
    const uint32_t i = atomicAdd(&Q.n[qi], 1u);
    if (i < (uint32_t)kQCap) Q.v[qi][i] = v;
    else direct = true;
  }
  frontier_push(direct, v, G, F, hub_cap);
}

struct RoundA {
  Frontier Cn, Dn, Xn;   // this round's C, D, X (written)
  const uint32_t* cbp;   // C_{r-1} = L_r members
  uint32_t* cb;          // C_r members
  uint32_t* db;          // D_r members
};

// A candidate's state, loaded at once.
struct Cand {
  int32_t k, cm, so, po;
  bool inl;
};
__device__ __forceinline__ Cand load_cand(const LoopState& S, const RoundA& P, int32_t w) {
  Cand c;
  c.k = S.K[w];
  c.cm = S.G.mexc[w];
  c.so = S.col0[w];
  c.po = S.prev[w];
  c.inl = has_bit(P.cbp, w);
  return c;
}
// 0: not a candidate here, 1: member of C_r, 2: member of L_r leaving it.
// src 0 = the C_{r-1} list, 1 = D_{r-1} (members of C_{r-1} are covered by
// their C entry), 2 = R_{r-1}.
__device__ __forceinline__ int role_of(int src, const Cand& c) {
  if (src == 1 && c.inl) return 0;
  if (c.k > 0) return 1;
  return src == 0 ? 2 : 0;
}

// Commit w's new colour m (one lane).  chv: the colour changed; dv: w's
// (in_L, colour, prev) state changed, i.e. w joins D.
__device__ __forceinline__ void commit_member(const LoopState& S, const RoundA& P, int32_t w,
                                              const Cand& c, int32_t m, bool& chv, bool& dv) {
  set_bit(P.cb, w);
  if (c.cm < 0) S.G.mexc[w] = m;
  S.col0[w] = m;
  S.col1[w] = m;
  S.prev[w] = c.so;
  chv = m != c.so;
  dv = !c.inl || c.po != c.so || chv;
  if (dv) {
    S.dold[w] = make_int2(c.so, (c.po + 1) | (c.inl ? (int32_t)0x80000000 : 0));
    S.K[w] = 0;
    set_bit(P.db, w);
  }
}

__device__ __forceinline__ void leave_member(const LoopState& S, const RoundA& P, int32_t x,
                                             const Cand& c) {
  S.dold[x] = make_int2(c.so, (c.po + 1) | (int32_t)0x80000000);
  set_bit(P.db, x);
}

// M and H candidates of one source list: one wave per vertex.  A stale mex
// comes from the hub's colour histogram, or a scan of the (< 512-slot) row.
template <bool HUB>
__device__ __forceinline__ void phase_a_wave(const LoopState& S, const RoundA& P, int src,
                                             const Frontier& F, BlockQueues& Q, uint32_t& changed,
                                             unsigned long long& slots, unsigned long long& live) {
  __shared__ unsigned long long wwin[kBlock / 64][HUB ? kWinWords : 8];
  const GraphView& G = S.G;
  RecolorArgs RA{};
  RA.G = G;
  const int lane = lane_id();
  unsigned long long* hw = wwin[threadIdx.x >> 6];
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t n = HUB ? F.cnt[2] : F.cnt[1];
  for (int64_t i = gwave; i < n; i += nwaves) {
    int32_t w;
    if constexpr (HUB) {
      const int4 it = F.h[i];
      if (it.y != 0) continue;
      w = it.x;
    } else {
      w = F.m[i];
    }
    const Cand c = load_cand(S, P, w);
    const int role = role_of(src, c);
    if (role == 0) continue;
    if (role == 2) {
      if (lane == 0) leave_member(S, P, w, c);
      q_push(Q, 2, lane == 0, w, G, P.Dn, S.hub_cap);
      continue;
    }
    int32_t m = c.cm;
    if (m < 0) {
      if constexpr (HUB) {
        m = hist_mex(G.hist + (uint64_t)G.hidx[w] * kHistW);
        if (m < 0) m = wave_row_mex<M_LOOP_HI>(RA, w, hw);
      } else {
        const uint2 rw = G.row[w];
        if (lane < 8) hw[lane] = 0ull;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        RegMask mk;
        mk.clear();
        mex_slots<M_LOOP_HI>(RA, w, 0, rw.x, lane, rw.y, 64, mk, hw, 512, slots, live);
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
      }
    }
    bool chv = false, dv = false;
    if (lane == 0) {
      commit_member(S, P, w, c, m, chv, dv);
      changed += chv ? 1u : 0u;
    }
    q_push(Q, 0, lane == 0, w, G, P.Cn, S.hub_cap);
    q_push(Q, 1, chv, w, G, P.Xn, S.hub_cap);
    q_push(Q, 2, dv, w, G, P.Dn, S.hub_cap);
  }
}

// Small candidates of one source list: 8-lane groups.
__device__ __forceinline__ void phase_a_small(const LoopState& S, const RoundA& P, int src,
                                              const Frontier& F, BlockQueues& Q, uint32_t& changed,
                                              unsigned long long& slots, unsigned long long& live) {
  const GraphView& G = S.G;
  RecolorArgs RA{};
  RA.G = G;
  const int lane = lane_id();
  const int grp = lane >> 3, sub = lane & 7;
  const int64_t gwave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t nS = F.cnt[0];
  for (int64_t b = gwave * 8; b < nS; b += nwaves * 8) {
    const int64_t idx = b + grp;
    const bool act = idx < nS;
    const int32_t w = act ? F.s[idx] : 0;
    Cand c{};
    if (act) c = load_cand(S, P, w);
    const int role = act ? role_of(src, c) : 0;
    const uint2 rw = (role == 1 && c.cm < 0) ? G.row[w] : make_uint2(0u, 0u);
    RegMask mk;
    mk.clear();
    mex_slots<M_LOOP_HI>(RA, w, 0, rw.x, sub, rw.y, kGroup, mk, nullptr, 0, slots, live);
    unsigned long long mask = mk.m0;
    mask |= __shfl_xor(mask, 1);
    mask |= __shfl_xor(mask, 2);
    mask |= __shfl_xor(mask, 4);
    const bool head = act && sub == 0;
    bool chv = false, dv = false;
    if (head && role == 1) {
      const int32_t m = c.cm >= 0 ? c.cm : __ffsll((long long)~mask) - 1;
      commit_member(S, P, w, c, m, chv, dv);
      changed += chv ? 1u : 0u;
    } else if (head && role == 2) {
      leave_member(S, P, w, c);
      dv = true;
    }
    q_push(Q, 0, head && role == 1, w, G, P.Cn, S.hub_cap);
    q_push(Q, 1, chv, w, G, P.Xn, S.hub_cap);
    q_push(Q, 2, dv, w, G, P.Dn, S.hub_cap);
  }
}

__device__ __forceinline__ void round_a(const LoopState& S, int r) {
  __shared__ BlockQueues Q;
  const Frontier Cp = c_list(S, r - 1), Dp = S.D[rot(r - 1)], Rp = S.R[rot(r - 1)];
  RoundA P;
  P.Cn = S.C[r % 3];
  P.Dn = S.D[r % 3];
  P.Xn = S.X[r % 3];
  P.cbp = S.B[(r + 1) & 1];
  P.cb = S.B[r & 1];
  P.db = S.Db[r & 1];
  uint32_t* dbp = S.Db[(r + 1) & 1];
  if (threadIdx.x < 3) Q.n[threadIdx.x] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.trace[r].x = (int32_t)fsize_dev(Cp);
    const Frontier z[4] = {S.C[(r + 1) % 3], S.D[(r + 1) % 3], S.X[(r + 1) % 3], S.R[r % 3]};
    for (int k = 0; k < 4; ++k)
      for (int i = 0; i < 4; ++i) z[k].cnt[i] = 0u;
  }
  __syncthreads();
  // D_{r-1}'s bits and R_{r-1}'s claims are no longer read
  visit_members(Dp, [&](int32_t x, bool valid) {
    if (valid) dbp[x >> 5] = 0u;
  });
  visit_members(Rp, [&](int32_t y, bool valid) {
    if (valid) S.Rb[y >> 5] = 0u;
  });
  uint32_t changed = 0;
  unsigned long long slots = 0, live = 0;
  const Frontier* src[3] = {&Cp, &Dp, &Rp};
  for (int k = 0; k < 3; ++k) phase_a_wave<true>(S, P, k, *src[k], Q, changed, slots, live);
  for (int k = 0; k < 3; ++k) phase_a_wave<false>(S, P, k, *src[k], Q, changed, slots, live);
  for (int k = 0; k < 3; ++k) phase_a_small(S, P, k, *src[k], Q, changed, slots, live);
  wave_add_u32(changed, reinterpret_cast<uint32_t*>(&S.trace[r].z));
  wave_add_u64(slots, &S.ctl->rslots);
  wave_add_u64(live, &S.ctl->rlive);
  // flush the queues
  __syncthreads();
  const Frontier* dst[3] = {&P.Cn, &P.Xn, &P.Dn};
  for (int k = 0; k < 3; ++k) {
    const uint32_t cnt = min(Q.n[k], (uint32_t)kQCap);
    for (uint32_t base = 0; base < cnt; base += blockDim.x) {
      const uint32_t i = base + threadIdx.x;
      const bool valid = i < cnt;
      push_converged(valid, valid ? Q.v[k][i] : 0, S.G, *dst[k], S.hub_cap);
    }
  }
}

// ---------------------------------------------------------------------------
// phase C.  For x in D_r and each slot (x, y):
//   y outside D_r: K[y] += (x's new term) - (x's old term); a rise at a
//                  non-member of C_r makes y a candidate (R_r);
//   y in C_r (= L_{r+1}): y's term towards x is part of x's recount.
__device__ __forceinline__ void round_c(const LoopState& S, int r) {
  const GraphView& G = S.G;
  const Frontier Cr = S.C[r % 3], Dr = S.D[r % 3], Xr = S.X[r % 3], Rr = S.R[r % 3];
  const Frontier Cp = c_list(S, r - 1);
  const uint32_t* cb = S.B[r & 1];
  uint32_t* cbp = S.B[(r + 1) & 1];
  uint32_t* db = S.Db[r & 1];
  const uint32_t nC = fsize_dev(Cr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.trace[r].y = (int32_t)nC;
    if (nC == 0) S.ctl->done_round = r;
  }
  visit_members(Cp, [&](int32_t x, bool valid) {
    if (valid) cbp[x >> 5] = 0u;
  });
  if (nC == 0) {
    visit_members(Dr, [&](int32_t x, bool valid) {
      if (valid) db[x >> 5] = 0u;
    });
    return;
  }
  unsigned long long slots = 0, live = 0;
  visit_rows(G, Dr, [&](int32_t x, uint2 rw, uint32_t lo, uint32_t hi, uint32_t stride, int width) {
    int32_t so = 0, po = 0, sn = 0, pn = 0;
    bool inl = false, inc = false;
    if (x >= 0) {
      const int2 o = S.dold[x];
      so = o.x;
      po = (o.y & 0x7fffffff) - 1;
      inl = o.y < 0;
      sn = S.col0[x];
      pn = S.prev[x];
      inc = has_bit(cb, x);
    }
    uint32_t cnt = 0;
    for (uint32_t j = lo; j < hi; j += kUnroll * stride) {
      int32_t y[kUnroll], c[kUnroll];
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) {
        const uint32_t jj = j + k * stride;
        const bool in = jj < hi;
        y[k] = in ? G.col[(uint64_t)rw.x + jj] : -2;
        c[k] = in ? G.ccol[(uint64_t)rw.x + jj] : -1;
      }
#pragma unroll
      for (int k = 0; k < kUnroll; ++k) {
        slots += y[k] != -2 ? 1ull : 0ull;
        bool pushr = false;
        if (y[k] >= 0) {
          ++live;
          const int32_t yy = y[k];
          if (yy < G.n_owned && !has_bit(db, yy)) {
            // y's colour is unchanged this round, so its slot value is current
            const bool to = inl && marks(x, so, po, yy, c[k]);
            const bool tn = inc && marks(x, sn, pn, yy, c[k]);
            if (to != tn) {
              atomicAdd(&S.K[yy], tn ? 1 : -1);
              pushr = tn && !has_bit(cb, yy) && claim_bit(S.Rb, yy);
            }
          }
          if (has_bit(cb, yy) && marks(yy, S.col0[yy], S.prev[yy], x, sn)) ++cnt;
        }
        frontier_push(pushr, y[k], G, Rr, S.hub_cap);
      }
    }
    cnt = group_sum(cnt, width);
    if ((lane_id() & (width - 1)) == 0 && x >= 0 && cnt) atomicAdd(&S.K[x], (int32_t)cnt);
  });
  // the new colours into the slot cache (the count pass above never reads a
  // slot of a changed neighbour: changed vertices are in D_r)
  visit_rows(G, Xr, [&](int32_t v, uint2 rw, uint32_t lo, uint32_t hi, uint32_t stride, int width) {
    (void)width;
    if (v >= 0) propagate_slots(G, S.col0[v], S.dold[v].x, rw.x, lo, hi, stride);
  });
  wave_add_u64(slots, &S.ctl->uslots);
  wave_add_u64(live, &S.ctl->ulive);
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_loop_init_bits(LoopState S) { init_bits(S); }
__global__ __launch_bounds__(kBlock) void k_loop_init_push(LoopState S) { init_push(S); }

__global__ __launch_bounds__(kBlock) void k_round_a(LoopState S, int r) {
  if (S.ctl->done_round < r) return;
  round_a(S, r);
}

__global__ __launch_bounds__(kBlock) void k_round_c(LoopState S, int r) {
  if (S.ctl->done_round < r) return;
  round_c(S, r);
}

// Persistent form: every round in one cooperative launch.  The exit test
// reads the control word after a grid barrier, so every workgroup takes the
// same branch.
__global__ __launch_bounds__(kBlock) void k_loop_persistent(LoopState S, int r0, int r1) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  if (r0 == 0) {
    init_bits(S);
    grid.sync();
    init_push(S);
    grid.sync();
  }
  for (int r = r0; r < r1; ++r) {
    round_a(S, r);
    grid.sync();
    round_c(S, r);
    grid.sync();
    if (__atomic_load_n(&S.ctl->done_round, __ATOMIC_RELAXED) != INT_MAX) break;
  }
}

// K, the D/R bitmaps and the C/L bitmaps are zero between batches.
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
  s->words = n / 32 + 2;
  DC_HIP(hipMalloc(&S.K, sizeof(int32_t) * n));
  DC_HIP(hipMalloc(&S.dold, sizeof(int2) * n));
  DC_HIP(hipMalloc(&s->bits, sizeof(uint32_t) * 3 * s->words));
  S.Db[0] = s->bits;
  S.Db[1] = s->bits + s->words;
  S.Rb = s->bits + 2 * s->words;
  DC_HIP(hipMalloc(&s->lbuf, sizeof(int32_t) * n * 24));
  DC_HIP(hipMalloc(&s->cnt, sizeof(uint32_t) * 48));
  DC_HIP(hipMalloc(&S.ctl, sizeof(LoopCtl)));
  DC_HIP(hipHostMalloc(&s->hctl, sizeof(LoopCtl), hipHostMallocDefault));
  DC_HIP(hipMemsetAsync(S.K, 0, sizeof(int32_t) * n, g->stream));
  DC_HIP(hipMemsetAsync(s->bits, 0, sizeof(uint32_t) * 3 * s->words, g->stream));
  Frontier* fr[12] = {&S.C[0], &S.C[1], &S.C[2], &S.D[0], &S.D[1], &S.D[2],
                      &S.X[0], &S.X[1], &S.X[2], &S.R[0], &S.R[1], &S.R[2]};
  for (int i = 0; i < 12; ++i) {
    fr[i]->s = s->lbuf + n * (2 * i);
    fr[i]->m = s->lbuf + n * (2 * i + 1);
    fr[i]->cnt = s->cnt + 4 * i;
  }
  return DC_OK;
}

static int inc_hubs(Graph* g, IncState* s) {
  if (s->hcap >= g->hchunk_cap && s->hb) return DC_OK;
  if (s->hb) DC_HIP(hipFree(s->hb));
  s->hb = nullptr;
  const uint32_t c = g->hchunk_cap;
  DC_HIP(hipMalloc(&s->hb, sizeof(int4) * (size_t)c * 12));
  s->hcap = c;
  LoopState& S = s->S;
  Frontier* fr[12] = {&S.C[0], &S.C[1], &S.C[2], &S.D[0], &S.D[1], &S.D[2],
                      &S.X[0], &S.X[1], &S.X[2], &S.R[0], &S.R[1], &S.R[2]};
  for (int i = 0; i < 12; ++i) fr[i]->h = s->hb + (size_t)c * i;
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
  f(s->S.K); f(s->S.dold); f(s->bits); f(s->lbuf); f(s->hb); f(s->cnt); f(s->S.trace);
  f(s->S.ctl);
  if (s->hctl) (void)hipHostFree(s->hctl);
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
                       (size_t)g->n, s->bits, 3 * s->words, cnt);
  DC_HIP(hipGetLastError());
  unsigned long long h = 0;
  DC_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, g->stream));
  DC_HIP(hipStreamSynchronize(g->stream));
  *bad = (int64_t)h;
  return DC_OK;
}

static int loop_reset(Graph* g, IncState* s) {
  DC_HIP(hipMemsetAsync(s->S.K, 0, sizeof(int32_t) * ((size_t)g->n + 1), g->stream));
  DC_HIP(hipMemsetAsync(s->bits, 0, sizeof(uint32_t) * 3 * s->words, g->stream));
  DC_HIP(hipMemsetAsync(g->cbits[0], 0, sizeof(uint32_t) * 2 * g->cbits_words, g->stream));
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
  *grid = std::max(1, std::min(bpm, g->fused_blocks_per_cu)) * cus;
  return DC_OK;
}

int loop_hi(Graph* g, Frontier L0, bool persistent, LoopResult* res) {
  IncState* s = inc_state(g);
  if (!s) {
    s = new IncState();
    g->inc = s;
    int rc = inc_alloc(g, s);
    if (rc) return rc;
  }
  int rc = inc_hubs(g, s);
  if (rc) return rc;
  const int32_t max_rounds = g->opt.max_rounds;
  if ((rc = inc_trace(g, s, 256))) return rc;

  LoopState& S = s->S;
  S.G = g->view();
  S.col0 = g->color[g->cur];
  S.col1 = g->color[g->cur ^ 1];
  S.prev = g->prev;
  S.B[0] = g->cbits[0];
  S.B[1] = g->cbits[1];
  S.L0 = L0;
  S.hub_cap = g->hub_cap;

  DC_HIP(hipMemsetAsync(s->cnt, 0, sizeof(uint32_t) * 48, g->stream));
  *s->hctl = LoopCtl{};
  s->hctl->done_round = INT_MAX;
  DC_HIP(hipMemcpyAsync(S.ctl, s->hctl, sizeof(LoopCtl), hipMemcpyHostToDevice, g->stream));
  int64_t launches = 0;
  double kms = 0.0;
  hipEvent_t e0 = g->ev[2], e1 = g->ev[3];
  int32_t done = INT_MAX;

  if (persistent) {
    // rounds in launches of up to 4096 (the trace grows between launches)
    int grid = 0;
    if ((rc = persistent_grid(g, &grid))) return rc;
    int32_t r0 = 0;
    while (r0 < max_rounds && done == INT_MAX) {
      int32_t r1 = std::min(max_rounds, r0 + 4096);
      if ((rc = inc_trace(g, s, r1))) return rc;
      DC_HIP(hipMemsetAsync(S.trace + r0, 0, sizeof(int4) * (size_t)(r1 - r0), g->stream));
      void* args[] = {&S, &r0, &r1};
      if (g->profiling) DC_HIP(hipEventRecord(e0, g->stream));
      DC_HIP(hipLaunchCooperativeKernel((const void*)k_loop_persistent, dim3(grid), dim3(kBlock), args,
                                        0, g->stream));
      if (g->profiling) DC_HIP(hipEventRecord(e1, g->stream));
      DC_HIP(hipMemcpyAsync(s->hctl, S.ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost, g->stream));
      DC_HIP(hipStreamSynchronize(g->stream));
      if (g->profiling) {
        float x = 0.f;
        DC_HIP(hipEventElapsedTime(&x, e0, e1));
        kms += x;
      }
      ++launches;
      done = s->hctl->done_round;
      r0 = r1;
    }
  } else {
    const dim3 grid(2048), blk(kBlock);
    hipLaunchKernelGGL(k_loop_init_bits, grid, blk, 0, g->stream, S);
    hipLaunchKernelGGL(k_loop_init_push, grid, blk, 0, g->stream, S);
    DC_HIP(hipGetLastError());
    launches = 2;
    int32_t r = 0, block = 8;
    while (r < max_rounds) {
      const int32_t nb = std::min(block, max_rounds - r);
      if ((rc = inc_trace(g, s, r + nb))) return rc;
      DC_HIP(hipMemsetAsync(S.trace + r, 0, sizeof(int4) * (size_t)nb, g->stream));
      if (g->profiling) DC_HIP(hipEventRecord(e0, g->stream));
      for (int32_t i = 0; i < nb; ++i, ++r) {
        hipLaunchKernelGGL(k_round_a, grid, blk, 0, g->stream, S, r);
        hipLaunchKernelGGL(k_round_c, grid, blk, 0, g->stream, S, r);
        launches += 2;
      }
      DC_HIP(hipGetLastError());
      if (g->profiling) DC_HIP(hipEventRecord(e1, g->stream));
      DC_HIP(hipMemcpyAsync(s->hctl, S.ctl, sizeof(LoopCtl), hipMemcpyDeviceToHost, g->stream));
      DC_HIP(hipStreamSynchronize(g->stream));
      if (g->profiling) {
        float x = 0.f;
        DC_HIP(hipEventElapsedTime(&x, e0, e1));
        kms += x;
      }
      done = s->hctl->done_round;
      if (done != INT_MAX) break;
      block = std::min(block * 2, 32);
    }
  }
  if (done == INT_MAX) {
    int rc2 = loop_reset(g, s);
    if (rc2) return rc2;
    set_error("recolor loop hit max_rounds");
    return DC_ERR_ROUNDS;
  }
  const int32_t rounds = done + 1;
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
  res->rslots = s->hctl->rslots;
  res->rlive = s->hctl->rlive;
  res->slots = s->hctl->uslots;
  res->live = s->hctl->ulive;
  res->kernel_ms = kms;
  res->launches = launches;
  return DC_OK;
}

}  // namespace dc
