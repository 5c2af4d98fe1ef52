// gen.cpp -- deterministic synthetic inputs: Erdos-Renyi G(n,m), R-MAT, and
// update-batch streams.  Every random draw comes from splitmix64 keyed by
// (seed, item index), so the output does not depend on the thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dyncolor.h"
#include "dc_guard.h"

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, n) (multiply-shift; bias < 2^-32 for n < 2^32)
  uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

int nthreads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  return (int)std::min(h, 16u);
}

template <typename F>
void parallel_for(int64_t n, F f) {
  int t = nthreads();
  if (n < 100000) t = 1;
  std::vector<std::thread> th;
  for (int i = 0; i < t; ++i) {
    int64_t lo = n * i / t, hi = n * (i + 1) / t;
    th.emplace_back([=, &f] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// parallel sort: sort T chunks, then merge pairwise
void psort(std::vector<uint64_t>& a) {
  int t = nthreads();
  int64_t n = (int64_t)a.size();
  if (n < 1000000 || t == 1) {
    std::sort(a.begin(), a.end());
    return;
  }
  std::vector<int64_t> cut(t + 1);
  for (int i = 0; i <= t; ++i) cut[i] = n * i / t;
  {
    std::vector<std::thread> th;
    for (int i = 0; i < t; ++i)
      th.emplace_back([&, i] { std::sort(a.begin() + cut[i], a.begin() + cut[i + 1]); });
    for (auto& x : th) x.join();
  }
  std::vector<uint64_t> b(n);
  std::vector<int64_t> c = cut;  // run boundaries
  while (c.size() > 2) {
    const size_t nr = c.size() - 1;
    std::vector<int64_t> nc{0};
    std::vector<std::thread> th;
    for (size_t i = 0; i < nr; i += 2) {
      const int64_t lo = c[i];
      if (i + 1 < nr) {
        const int64_t mid = c[i + 1], hi = c[i + 2];
        th.emplace_back([&, lo, mid, hi] {
          std::merge(a.begin() + lo, a.begin() + mid, a.begin() + mid, a.begin() + hi,
                     b.begin() + lo);
        });
        nc.push_back(hi);
      } else {
        const int64_t hi = c[i + 1];
        th.emplace_back([&, lo, hi] { std::copy(a.begin() + lo, a.begin() + hi, b.begin() + lo); });
        nc.push_back(hi);
      }
    }
    for (auto& x : th) x.join();
    a.swap(b);
    c.swap(nc);
  }
}

inline uint64_t key_of(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  return ((uint64_t)lo << 32) | hi;
}

// Open-addressing set of edge keys with the edge's position in the edge array
// (linear probing, backward-shift deletion).
struct EdgeIndex {
  std::vector<uint64_t> key;
  std::vector<uint32_t> pos;
  uint64_t mask = 0;
  static constexpr uint64_t kEmpty = ~0ull;
  explicit EdgeIndex(uint64_t expected) {
    uint64_t c = 16;
    while (c < expected * 2 + 16) c <<= 1;
    key.assign(c, kEmpty);
    pos.assign(c, 0);
    mask = c - 1;
  }
  static uint64_t h(uint64_t k) { return splitmix64(k); }
  int64_t find(uint64_t k) const {
    for (uint64_t i = h(k) & mask;; i = (i + 1) & mask) {
      if (key[i] == k) return (int64_t)i;
      if (key[i] == kEmpty) return -1;
    }
  }
  void insert(uint64_t k, uint32_t p) {
    uint64_t i = h(k) & mask;
    while (key[i] != kEmpty) {
      if (key[i] == k) {
        pos[i] = p;
        return;
      }
      i = (i + 1) & mask;
    }
    key[i] = k;
    pos[i] = p;
  }
  // concurrent bulk insert of distinct keys
  void insert_concurrent(uint64_t k, uint32_t p) {
    auto* ak = reinterpret_cast<std::atomic<uint64_t>*>(key.data());
    uint64_t i = h(k) & mask;
    for (;;) {
      uint64_t e = kEmpty;
      if (ak[i].compare_exchange_strong(e, k, std::memory_order_relaxed)) {
        pos[i] = p;
        return;
      }
      i = (i + 1) & mask;
    }
  }
  void erase_at(uint64_t i) {
    key[i] = kEmpty;
    uint64_t j = i;
    for (;;) {
      j = (j + 1) & mask;
      if (key[j] == kEmpty) return;
      uint64_t home = h(key[j]) & mask;
      // move key[j] into the hole if its home is not in (i, j]
      bool between = (i <= j) ? (home > i && home <= j) : (home > i || home <= j);
      if (!between) {
        key[i] = key[j];
        pos[i] = pos[j];
        key[j] = kEmpty;
        i = j;
      }
    }
  }
};

}  // namespace

extern "C" {

int dc_gen_er(int32_t n, int64_t m, uint64_t seed, int32_t* u, int32_t* v) {
  return dc::guarded("dc_gen_er", [&]() -> int {
    if (n < 2 || m < 0 || (m > 0 && (!u || !v)) || (double)m > (double)n * (n - 1) / 2.0) {
      dc::set_error("dc_gen_er: invalid argument");
      return DC_ERR_INVALID;
    }
    // oversample candidate pairs in rounds; keep first occurrences in draw order
    std::vector<uint64_t> chosen;  // (draw index << 0) keys in draw order
    chosen.reserve((size_t)m);
    std::vector<uint64_t> seen;    // sorted keys already chosen
    uint64_t drawn = 0;
    while ((int64_t)chosen.size() < m) {
      int64_t need = m - (int64_t)chosen.size();
      int64_t cnt = need + need / 8 + 1024;
      std::vector<uint64_t> cand(cnt);
      uint64_t base = drawn;
      parallel_for(cnt, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
          Rng r(splitmix64(seed ^ 0x5851F42D4C957F2Dull) + (base + i) * 0x9E3779B97F4A7C15ull * 2);
          uint32_t a = (uint32_t)r.below((uint64_t)n), b = (uint32_t)r.below((uint64_t)n);
          cand[i] = a == b ? ~0ull : key_of(a, b);
        }
      });
      drawn += cnt;
      // (key, draw order) sorted by key to find first occurrences
      std::vector<std::pair<uint64_t, int64_t>> kp(cnt);
      for (int64_t i = 0; i < cnt; ++i) kp[i] = {cand[i], i};
      std::sort(kp.begin(), kp.end());
      std::vector<int64_t> firsts;
      for (int64_t i = 0; i < cnt; ++i) {
        if (kp[i].first == ~0ull) continue;
        if (i > 0 && kp[i].first == kp[i - 1].first) continue;
        if (std::binary_search(seen.begin(), seen.end(), kp[i].first)) continue;
        firsts.push_back(kp[i].second);
      }
      std::sort(firsts.begin(), firsts.end());
      for (int64_t idx : firsts) {
        if ((int64_t)chosen.size() >= m) break;
        chosen.push_back(cand[idx]);
      }
      seen = chosen;
      std::sort(seen.begin(), seen.end());
    }
    for (int64_t i = 0; i < m; ++i) {
      u[i] = (int32_t)(chosen[i] >> 32);
      v[i] = (int32_t)(chosen[i] & 0xffffffffu);
    }
    return DC_OK;
  });
}

int dc_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                int32_t* u, int32_t* v, int64_t* m_out) {
  return dc::guarded("dc_gen_rmat", [&]() -> int {
    if (scale < 1 || scale > 30 || edge_factor < 1 || !u || !v || !m_out || a < 0 || b < 0 ||
        c < 0 || a + b + c > 1.0) {
      dc::set_error("dc_gen_rmat: invalid argument");
      return DC_ERR_INVALID;
    }
    const int64_t samples = (int64_t)edge_factor << scale;
    std::vector<uint64_t> keys(samples);
    const double ab = a + b, abc = a + b + c;
    parallel_for(samples, [&](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) {
        Rng r(splitmix64(seed) + (uint64_t)i * 0xD1B54A32D192ED03ull);
        uint32_t x = 0, y = 0;
        for (int l = 0; l < scale; ++l) {
          double p = r.unit();
          uint32_t bx = p >= ab ? 1u : 0u;
          uint32_t by = ((p >= a && p < ab) || p >= abc) ? 1u : 0u;
          x = (x << 1) | bx;
          y = (y << 1) | by;
        }
        keys[i] = x == y ? ~0ull : key_of(x, y);
      }
    });
    psort(keys);
    int64_t m = 0;
    for (int64_t i = 0; i < samples; ++i) {
      if (keys[i] == ~0ull) break;
      if (i > 0 && keys[i] == keys[i - 1]) continue;
      u[m] = (int32_t)(keys[i] >> 32);
      v[m] = (int32_t)(keys[i] & 0xffffffffu);
      ++m;
    }
    *m_out = m;
    return DC_OK;
  });
}

int dc_gen_batches(int32_t n, int64_t m, const int32_t* u, const int32_t* v, int32_t nb, int64_t bs,
                   double ins_frac, uint64_t seed, int32_t* bu, int32_t* bv, uint8_t* bins) {
  return dc::guarded("dc_gen_batches", [&]() -> int {
    if (n < 2 || m < 0 || nb < 0 || bs < 0 || ins_frac < 0 || ins_frac > 1 ||
        (m > 0 && (!u || !v)) || (nb > 0 && bs > 0 && (!bu || !bv || !bins))) {
      dc::set_error("dc_gen_batches: invalid argument");
      return DC_ERR_INVALID;
    }
    const int64_t n_ins = (int64_t)std::llround((double)bs * ins_frac);
    const int64_t n_del = bs - n_ins;
    std::vector<uint64_t> E((size_t)m);
    for (int64_t i = 0; i < m; ++i) E[i] = key_of((uint32_t)u[i], (uint32_t)v[i]);
    EdgeIndex idx((uint64_t)m + (uint64_t)nb * (uint64_t)n_ins + 16);
    parallel_for(m, [&](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) idx.insert_concurrent(E[i], (uint32_t)i);
    });
    if ((int64_t)E.size() != m) return DC_ERR_INVALID;
    Rng r(splitmix64(seed ^ 0xA0761D6478BD642Full));
    std::vector<uint64_t> ins_keys, del_keys;
    std::vector<uint64_t> batch_new;
    for (int32_t bi = 0; bi < nb; ++bi) {
      if (n_del > (int64_t)E.size()) {
        dc::set_error("dc_gen_batches: not enough edges to delete");
        return DC_ERR_INVALID;
      }
      // insertions: uniform pairs that are neither edges nor already chosen
      ins_keys.clear();
      EdgeIndex fresh((uint64_t)n_ins + 16);
      while ((int64_t)ins_keys.size() < n_ins) {
        uint32_t x = (uint32_t)r.below((uint64_t)n), y = (uint32_t)r.below((uint64_t)n);
        if (x == y) continue;
        uint64_t k = key_of(x, y);
        if (idx.find(k) >= 0 || fresh.find(k) >= 0) continue;
        fresh.insert(k, 0);
        ins_keys.push_back(k);
      }
      // deletions: distinct edges of the pre-batch graph, uniformly
      del_keys.clear();
      for (int64_t d = 0; d < n_del; ++d) {
        uint64_t p = r.below((uint64_t)E.size());
        uint64_t k = E[p];
        int64_t slot = idx.find(k);
        // swap-remove from E, keep the index consistent
        uint64_t last = E.back();
        E[p] = last;
        E.pop_back();
        if (last != k) {
          int64_t ls = idx.find(last);
          idx.pos[ls] = (uint32_t)p;
        }
        idx.erase_at((uint64_t)slot);
        del_keys.push_back(k);
      }
      for (uint64_t k : ins_keys) {
        idx.insert(k, (uint32_t)E.size());
        E.push_back(k);
      }
      int64_t o = (int64_t)bi * bs;
      for (int64_t i = 0; i < n_del; ++i, ++o) {
        bu[o] = (int32_t)(del_keys[i] >> 32);
        bv[o] = (int32_t)(del_keys[i] & 0xffffffffu);
        bins[o] = 0;
      }
      for (int64_t i = 0; i < n_ins; ++i, ++o) {
        bu[o] = (int32_t)(ins_keys[i] >> 32);
        bv[o] = (int32_t)(ins_keys[i] & 0xffffffffu);
        bins[o] = 1;
      }
    }
    return DC_OK;
  });
}

}  // extern "C"
