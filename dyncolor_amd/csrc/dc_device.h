// dc_device.h -- wave64 building blocks used by every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dc_internal.h"

namespace dc {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane (v_mbcnt_lo / v_mbcnt_hi).
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Stream compaction of one wave's predicates into `list`: ballot the
// predicate, one atomic per wave reserves the block of slots, each active
// lane writes at base + (set lanes below it).  Safe under divergence: only
// the lanes executing the call take part.
template <typename T>
__device__ __forceinline__ void wave_append(bool p, T val, T* list, uint32_t* counter) {
  unsigned long long m = __ballot(p);
  if (m == 0ull) return;
  int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (p) list[base + lanes_below(m)] = val;
}

// wave_append into a list of `cap` entries: the counter counts every append,
// entries past cap are not written (a count above cap marks the list as
// incomplete).
template <typename T>
__device__ __forceinline__ void wave_append_capped(bool p, T val, T* list, uint32_t* counter,
                                                   uint32_t cap) {
  unsigned long long m = __ballot(p);
  if (m == 0ull) return;
  int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  const uint32_t pos = base + lanes_below(m);
  if (p && pos < cap) list[pos] = val;
}

// Wave-aggregated add of a per-lane value into a 64-bit counter.
__device__ __forceinline__ void wave_add_u64(unsigned long long v, unsigned long long* ctr) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  unsigned long long act = __ballot(true);
  int leader = __ffsll((long long)act) - 1;
  if (lane_id() == leader && v) atomicAdd(ctr, v);
}

__device__ __forceinline__ void wave_add_u32(uint32_t v, uint32_t* ctr) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  unsigned long long act = __ballot(true);
  int leader = __ffsll((long long)act) - 1;
  if (lane_id() == leader && v) atomicAdd(ctr, v);
}

// LDF priority: u precedes v  (prop desc, then lower ID).
__device__ __forceinline__ bool precedes(const int32_t* prop, int32_t u, int32_t pu,
                                         int32_t v, int32_t pv) {
  (void)prop;
  return pu > pv || (pu == pv && u < v);
}

// Claim v for the epoch; true for exactly one caller per epoch.
__device__ __forceinline__ bool claim(uint32_t* mark, int32_t v, uint32_t epoch) {
  if (__atomic_load_n(&mark[v], __ATOMIC_RELAXED) == epoch) return false;
  return atomicExch(&mark[v], epoch) != epoch;
}

// Claim v in a 1-bit-per-vertex bitmap (small enough to stay in L2); true for
// exactly one caller until the bit is cleared.  The plain load only skips the
// atomic when the bit is already visible.
__device__ __forceinline__ bool claim_bit(uint32_t* bits, int32_t v) {
  uint32_t* p = bits + (v >> 5);
  const uint32_t b = 1u << (v & 31);
  if (__atomic_load_n(p, __ATOMIC_RELAXED) & b) return false;
  return (atomicOr(p, b) & b) == 0u;
}

// Append a claimed vertex to the class list its row length selects.  Hub rows
// get one entry per kChunk-slot chunk and a hub slot for the mex reduction.
// Must be called by whole waves (divergence allowed) so the class appends can
// be wave-aggregated.
__device__ __forceinline__ void frontier_push(bool p, int32_t v, const GraphView& G,
                                             Frontier F, uint32_t hub_cap) {
  uint32_t len = 0;
  if (p) len = G.row[v].y;
  bool is_s = p && len < kSmallLen;
  bool is_m = p && len >= kSmallLen && len < kMidLen;
  bool is_h = p && len >= kMidLen;
  wave_append(is_s, v, F.s, &F.cnt[0]);
  wave_append(is_m, v, F.m, &F.cnt[1]);
  if (is_h) {
    uint32_t nch = (len + kChunk - 1) / kChunk;
    uint32_t slot = atomicAdd(&F.cnt[3], 1u);
    if (slot < hub_cap) {
      uint32_t base = atomicAdd(&F.cnt[2], nch);
      for (uint32_t c = 0; c < nch; ++c)
        F.h[base + c] = make_int4(v, (int)c, (int)slot, (int)nch);
    }
  }
}

}  // namespace dc
