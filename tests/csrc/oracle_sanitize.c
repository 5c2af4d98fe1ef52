/*
 * oracle_sanitize.c -- drives the CPU oracle (oracle/dc_oracle.c) over random
 * multigraphs, partitions and update batches for a build with the address and
 * undefined-behaviour sanitizers (tests/test_oracle_sanitize.py).  Test
 * infrastructure only.  Exits 0 when every result is a valid colouring.
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/dc_oracle.h"

static uint64_t rng_state = 88172645463325252ull;
static uint32_t rnd(uint32_t n) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (uint32_t)(rng_state % n);
}

int main(void) {
  for (int trial = 0; trial < 3000; ++trial) {
    const int32_t n = 3 + (int32_t)rnd(60);
    const int32_t n_owned = trial % 3 == 0 ? 1 + (int32_t)rnd((uint32_t)n) : n;
    const int rule = n_owned < n ? ORC_RULE_INDEPENDENT : (int)(trial & 1);
    const int64_t m = rnd((uint32_t)(4 * n));
    int32_t* u = malloc(sizeof(int32_t) * (size_t)(m + 1));
    int32_t* v = malloc(sizeof(int32_t) * (size_t)(m + 1));
    int64_t k = 0;
    for (int64_t i = 0; i < m; ++i) {
      int32_t a = (int32_t)rnd((uint32_t)n), b = (int32_t)rnd((uint32_t)n);
      if (a == b || (a >= n_owned && b >= n_owned)) continue;
      u[k] = a;
      v[k] = b;
      ++k;
    }
    orc_graph* g = orc_graph_create(n, n_owned, k, u, v);
    if (!g) {
      fprintf(stderr, "trial %d: create failed\n", trial);
      return 1;
    }
    int32_t* color = malloc(sizeof(int32_t) * (size_t)n);
    int32_t* prev = malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t i = 0; i < n; ++i) color[i] = (int32_t)rnd(4);
    orc_static_greedy(g, rule, color);
    int64_t trace[4 * 512];
    for (int batch = 0; batch < 4; ++batch) {
      const int64_t nl = rnd(30);
      int32_t bu[32], bv[32];
      uint8_t bi[32];
      int64_t kk = 0;
      for (int64_t i = 0; i < nl; ++i) {
        int32_t a = (int32_t)rnd((uint32_t)n), b = (int32_t)rnd((uint32_t)n);
        if (a == b || (a >= n_owned && b >= n_owned)) continue;
        bu[kk] = a;
        bv[kk] = b;
        bi[kk] = (uint8_t)rnd(2);
        ++kk;
      }
      if (orc_apply_batch(g, color, kk, bu, bv, bi) != ORC_OK) {
        fprintf(stderr, "trial %d: apply failed\n", trial);
        return 1;
      }
      int32_t nph = 0;
      if (orc_recolor(g, rule, color, prev, 100000, trace, 512, &nph) != ORC_OK) {
        fprintf(stderr, "trial %d: recolor failed\n", trial);
        return 1;
      }
      int32_t mc = 0;
      const int64_t mono = orc_validate(g, color, &mc);
      /* the independent rule leaves conflicts no batch line touched; the
       * static colouring above made the start proper, so none remain */
      if (mono != 0) {
        fprintf(stderr, "trial %d batch %d: %lld monochromatic\n", trial, batch, (long long)mono);
        return 1;
      }
    }
    int32_t* nb = malloc(sizeof(int32_t) * (size_t)(8 * n + 8));
    for (int32_t x = 0; x < n; ++x) {
      const int64_t d = orc_live_degree(g, x);
      if (orc_neighbors(g, x, nb, 8 * n + 8) != d) return 1;
    }
    free(nb);
    orc_graph_destroy(g);
    free(u);
    free(v);
    free(color);
    free(prev);
  }
  printf("ok\n");
  return 0;
}
