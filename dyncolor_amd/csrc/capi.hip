// capi.hip -- the extern "C" entry points declared in include/dyncolor.h.
// Validates host inputs, owns the per-graph HIP stream and device state, and
// maps every failure to a DC_ERR_* code (nothing throws across the ABI).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "dc_guard.h"
#include "dc_internal.h"

namespace dc {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* where) {
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? DC_ERR_NOMEM : DC_ERR_HIP;
}

static int check_edges(int32_t n, int32_t n_owned, int64_t m, const int32_t* u, const int32_t* v) {
  for (int64_t i = 0; i < m; ++i) {
    int32_t a = u[i], b = v[i];
    if (a < 0 || b < 0 || a >= n || b >= n) {
      set_error("vertex ID out of range at line " + std::to_string(i));
      return DC_ERR_VERTEX;
    }
    if (a == b) {
      set_error("self loop at line " + std::to_string(i));
      return DC_ERR_VERTEX;
    }
    if (a >= n_owned && b >= n_owned) {
      set_error("edge between two ghost vertices at line " + std::to_string(i));
      return DC_ERR_VERTEX;
    }
  }
  return DC_OK;
}

template <typename T>
static int upload(Graph* g, T** dst, const T* src, int64_t count) {
  if (count <= 0) return DC_OK;
  DC_HIP(hipMalloc(dst, sizeof(T) * (size_t)count));
  DC_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * (size_t)count, hipMemcpyHostToDevice, g->stream));
  return DC_OK;
}

static int create_common(int32_t n, int32_t n_owned, int64_t m, const int32_t* u, const int32_t* v,
                         const dc_options* opt, dc_graph** out) {
  if (!out || n < 0 || n_owned < 0 || n_owned > n || m < 0 || (m > 0 && (!u || !v))) {
    set_error("dc_graph_create: invalid argument");
    return DC_ERR_INVALID;
  }
  *out = nullptr;
  int rc = check_edges(n, n_owned, m, u, v);
  if (rc) return rc;
  Graph* g = new (std::nothrow) Graph();
  if (!g) return DC_ERR_NOMEM;
  dc_options_default(&g->opt);
  if (opt) g->opt = *opt;
  dc_tuning_default(&g->tuning);
  if (g->opt.slack_frac < 0.f) g->opt.slack_frac = 0.f;
  if (g->opt.slack_min < 1) g->opt.slack_min = 1;
  if (g->opt.spare_slots < 0) g->opt.spare_slots = 0;
  if (g->opt.max_rounds < 1) g->opt.max_rounds = 1;
  g->n = n;
  g->n_owned = n_owned;
  g->device = g->opt.device;
  auto fail = [&](int code) {
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    free_state(g);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return code;
  };
  hipError_t e = hipSetDevice(g->device);
  if (e != hipSuccess) return fail(hip_fail(e, "hipSetDevice"));
  e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(hip_fail(e, "hipStreamCreate"));
  if ((rc = alloc_state(g))) return fail(rc);
  int32_t *du = nullptr, *dv = nullptr;
  if ((rc = upload(g, &du, u, m)) || (rc = upload(g, &dv, v, m))) {
    if (du) (void)hipFree(du);
    if (dv) (void)hipFree(dv);
    return fail(rc);
  }
  rc = csr_build(g, m, du, dv);
  (void)hipStreamSynchronize(g->stream);
  if (du) (void)hipFree(du);
  if (dv) (void)hipFree(dv);
  if (rc) return fail(rc);
  *out = reinterpret_cast<dc_graph*>(g);
  return DC_OK;
}

}  // namespace dc

using dc::Graph;

static inline Graph* G_(dc_graph* h) { return reinterpret_cast<Graph*>(h); }

struct dc_batch {
  dc::Batch b;
  int32_t n = 0;
  int device = 0;
  int64_t n_ins = 0;
};

extern "C" {

void dc_options_default(dc_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->slack_frac = 0.25f;
  opt->slack_min = 4;
  opt->max_rounds = 1000000;
  opt->driver = DC_DRIVER_PLAIN;
  opt->device = 0;
}

const char* dc_last_error(void) { return dc::g_last_error.c_str(); }

void dc_tuning_default(dc_tuning* t) {
  if (!t) return;
  std::memset(t, 0, sizeof(*t));
  t->loop_blocks = 2048;
  t->light_blocks = 0;
  t->light_c = 4096;
  t->list_c = 1 << 17;
  t->list_slots = 1ll << 22;
  t->heavy_div = 8;
}

int dc_set_tuning(dc_graph* h, const dc_tuning* t) {
  return dc::guarded("dc_set_tuning", [&]() -> int {
    if (!h || !t || t->loop_blocks < 1 || t->loop_blocks > 4096 || t->light_blocks < 0 ||
        t->light_blocks > 4096 || t->light_c < 0 || t->list_c < -1 || t->list_slots < 0 ||
        t->heavy_div < 1 || t->ldf_loop < 0 || t->ldf_loop > 1) {
      dc::set_error("dc_set_tuning: parameter out of range");
      return DC_ERR_INVALID;
    }
    G_(h)->tuning = *t;
    return DC_OK;
  });
}

int dc_get_tuning(const dc_graph* h, dc_tuning* t) {
  return dc::guarded("dc_get_tuning", [&]() -> int {
    if (!h || !t) return DC_ERR_INVALID;
    *t = reinterpret_cast<const dc::Graph*>(h)->tuning;
    return DC_OK;
  });
}

int dc_device_count(int32_t* count) {
  return dc::guarded("dc_device_count", [&]() -> int {
    if (!count) return DC_ERR_INVALID;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
      *count = 0;
      return dc::hip_fail(e, "hipGetDeviceCount");
    }
    *count = c;
    return DC_OK;
  });
}

int dc_device_info(int32_t device, char* name, int32_t cap, char* arch, int32_t arch_cap,
                   int32_t* compute_units) {
  return dc::guarded("dc_device_info", [&]() -> int {
    hipDeviceProp_t p;
    DC_HIP(hipGetDeviceProperties(&p, device));
    if (name && cap > 0) {
      std::strncpy(name, p.name, (size_t)cap - 1);
      name[cap - 1] = 0;
    }
    if (arch && arch_cap > 0) {
      std::strncpy(arch, p.gcnArchName, (size_t)arch_cap - 1);
      arch[arch_cap - 1] = 0;
    }
    if (compute_units) *compute_units = p.multiProcessorCount;
    return DC_OK;
  });
}

int dc_graph_create(int32_t n, int64_t m, const int32_t* u, const int32_t* v, const dc_options* opt,
                    dc_graph** out) {
  return dc::guarded("dc_graph_create", [&]() -> int {
    return dc::create_common(n, n, m, u, v, opt, out);
  });
}

int dc_partition_create(int32_t n_owned, int32_t n_ghost, int64_t m, const int32_t* u,
                        const int32_t* v, const int32_t* ghost_colors, const dc_options* opt,
                        dc_graph** out) {
  return dc::guarded("dc_partition_create", [&]() -> int {
    if (n_owned < 0 || n_ghost < 0 || (int64_t)n_owned + n_ghost > 0x7ffffffe ||
        (n_ghost > 0 && !ghost_colors)) {
      dc::set_error("dc_partition_create: invalid argument");
      return DC_ERR_INVALID;
    }
    int32_t n = n_owned + n_ghost;
    int rc = dc::create_common(n, n_owned, m, u, v, opt, out);
    if (rc) return rc;
    Graph* g = G_(*out);
    if (n_ghost > 0) {
      for (int b = 0; b < 2; ++b) {
        hipError_t e = hipMemcpyAsync(g->color[b] + n_owned, ghost_colors, sizeof(int32_t) * n_ghost,
                                      hipMemcpyHostToDevice, g->stream);
        if (e != hipSuccess) {
          dc_graph_destroy(*out);
          *out = nullptr;
          return dc::hip_fail(e, "upload ghost colors");
        }
      }
      rc = dc::rebuild_ccol(g);
      (void)hipStreamSynchronize(g->stream);
      if (rc) {
        dc_graph_destroy(*out);
        *out = nullptr;
        return rc;
      }
    }
    return DC_OK;
  });
}

int dc_graph_destroy(dc_graph* h) {
  return dc::guarded("dc_graph_destroy", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    dc::free_state(g);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return DC_OK;
  });
}

int dc_graph_info(const dc_graph* h, int32_t* n, int32_t* n_owned, int64_t* live_entries,
                  int64_t* capacity) {
  return dc::guarded("dc_graph_info", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    const Graph* g = reinterpret_cast<const Graph*>(h);
    if (n) *n = g->n;
    if (n_owned) *n_owned = g->n_owned;
    if (capacity) *capacity = (int64_t)g->total_cap;
    if (live_entries) *live_entries = -1;  // not tracked; use dc_neighbors per vertex
    return DC_OK;
  });
}

int dc_colors_set(dc_graph* h, const int32_t* colors) {
  return dc::guarded("dc_colors_set", [&]() -> int {
    if (!h || !colors) return DC_ERR_INVALID;
    Graph* g = G_(h);
    for (int32_t i = 0; i < g->n; ++i)
      if (colors[i] < 0) {
        dc::set_error("negative color at vertex " + std::to_string(i));
        return DC_ERR_INVALID;
      }
    DC_HIP(hipSetDevice(g->device));
    for (int b = 0; b < 2; ++b)
      DC_HIP(hipMemcpyAsync(g->color[b], colors, sizeof(int32_t) * (size_t)g->n,
                            hipMemcpyHostToDevice, g->stream));
    int rc = dc::rebuild_ccol(g);
    if (rc) return rc;
    DC_HIP(hipStreamSynchronize(g->stream));
    return DC_OK;
  });
}

int dc_check_consistency(dc_graph* h, int64_t* bad_slots) {
  return dc::guarded("dc_check_consistency", [&]() -> int {
    if (!h || !bad_slots) return DC_ERR_INVALID;
    Graph* g = G_(h);
    if (g->staged) {
      dc::set_error("dc_check_consistency: a batch is staged");
      return DC_ERR_STATE;
    }
    DC_HIP(hipSetDevice(g->device));
    return dc::check_consistency(g, bad_slots);
  });
}

int dc_colors_get(dc_graph* h, int32_t* colors) {
  return dc::guarded("dc_colors_get", [&]() -> int {
    if (!h || !colors) return DC_ERR_INVALID;
    Graph* g = G_(h);
    DC_HIP(hipSetDevice(g->device));
    DC_HIP(hipMemcpyAsync(colors, g->color[g->cur], sizeof(int32_t) * (size_t)g->n,
                          hipMemcpyDeviceToHost, g->stream));
    DC_HIP(hipStreamSynchronize(g->stream));
    return DC_OK;
  });
}

int dc_static_color(dc_graph* h, int32_t rule) {
  return dc::guarded("dc_static_color", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    DC_HIP(hipSetDevice(g->device));
    return dc::static_color(g, rule);
  });
}

static int validate_batch(Graph* g, int64_t k, const int32_t* u, const int32_t* v,
                          const uint8_t* ins) {
  if (k < 0 || (k > 0 && (!u || !v || !ins))) {
    dc::set_error("batch: invalid argument");
    return DC_ERR_INVALID;
  }
  if (k > 0x7fffffffLL / 2) {
    dc::set_error("batch too large");
    return DC_ERR_INVALID;
  }
  for (int64_t i = 0; i < k; ++i)
    if (ins[i] > 1) {
      dc::set_error("is_insert must be 0 or 1 at line " + std::to_string(i));
      return DC_ERR_INVALID;
    }
  return dc::check_edges(g->n, g->n_owned, k, u, v);
}

int dc_apply_batch(dc_graph* h, int64_t k, const int32_t* u, const int32_t* v,
                   const uint8_t* is_insert) {
  return dc::guarded("dc_apply_batch", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    int rc = validate_batch(g, k, u, v, is_insert);
    if (rc) return rc;
    DC_HIP(hipSetDevice(g->device));
    dc::Batch& sb = g->scratch_batch;
    if (k > sb.k) {
      if (sb.u) DC_HIP(hipFree(sb.u));
      if (sb.v) DC_HIP(hipFree(sb.v));
      if (sb.ins) DC_HIP(hipFree(sb.ins));
      sb.u = sb.v = nullptr;
      sb.ins = nullptr;
      DC_HIP(hipMalloc(&sb.u, sizeof(int32_t) * (size_t)k));
      DC_HIP(hipMalloc(&sb.v, sizeof(int32_t) * (size_t)k));
      DC_HIP(hipMalloc(&sb.ins, (size_t)k));
      sb.k = k;
    }
    dc::Batch b;
    b.k = k;
    b.u = sb.u;
    b.v = sb.v;
    b.ins = sb.ins;
    if (k > 0) {
      DC_HIP(hipMemcpyAsync(b.u, u, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice, g->stream));
      DC_HIP(hipMemcpyAsync(b.v, v, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice, g->stream));
      DC_HIP(hipMemcpyAsync(b.ins, is_insert, (size_t)k, hipMemcpyHostToDevice, g->stream));
    }
    int64_t ni = 0;
    for (int64_t i = 0; i < k; ++i) ni += is_insert[i];
    g->staged_k = k;
    g->staged_ins = ni;
    g->staged_del = k - ni;
    return dc::csr_apply(g, b, 0);
  });
}

int dc_batch_create(dc_graph* h, int64_t k, const int32_t* u, const int32_t* v,
                    const uint8_t* is_insert, dc_batch** out) {
  return dc::guarded("dc_batch_create", [&]() -> int {
    if (!h || !out) return DC_ERR_INVALID;
    Graph* g = G_(h);
    int rc = validate_batch(g, k, u, v, is_insert);
    if (rc) return rc;
    DC_HIP(hipSetDevice(g->device));
    dc_batch* b = new (std::nothrow) dc_batch();
    if (!b) return DC_ERR_NOMEM;
    b->n = g->n;
    b->device = g->device;
    b->b.k = k;
    if (k > 0) {
      hipError_t e;
      if ((e = hipMalloc(&b->b.u, sizeof(int32_t) * (size_t)k)) != hipSuccess ||
          (e = hipMalloc(&b->b.v, sizeof(int32_t) * (size_t)k)) != hipSuccess ||
          (e = hipMalloc(&b->b.ins, (size_t)k)) != hipSuccess) {
        dc_batch_destroy(b);
        return dc::hip_fail(e, "dc_batch_create");
      }
      (void)hipMemcpy(b->b.u, u, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice);
      (void)hipMemcpy(b->b.v, v, sizeof(int32_t) * (size_t)k, hipMemcpyHostToDevice);
      (void)hipMemcpy(b->b.ins, is_insert, (size_t)k, hipMemcpyHostToDevice);
    }
    for (int64_t i = 0; i < k; ++i) b->n_ins += is_insert[i];
    *out = b;
    return DC_OK;
  });
}

int dc_batch_destroy(dc_batch* b) {
  return dc::guarded("dc_batch_destroy", [&]() -> int {
    if (!b) return DC_ERR_INVALID;
    (void)hipSetDevice(b->device);
    if (b->b.u) (void)hipFree(b->b.u);
    if (b->b.v) (void)hipFree(b->b.v);
    if (b->b.ins) (void)hipFree(b->b.ins);
    delete b;
    return DC_OK;
  });
}

int dc_apply_device_batch(dc_graph* h, const dc_batch* b) {
  return dc::guarded("dc_apply_device_batch", [&]() -> int {
    if (!h || !b) return DC_ERR_INVALID;
    Graph* g = G_(h);
    if (b->n != g->n) {
      dc::set_error("batch was validated for a different graph size");
      return DC_ERR_INVALID;
    }
    DC_HIP(hipSetDevice(g->device));
    g->staged_k = b->b.k;
    g->staged_ins = b->n_ins;
    g->staged_del = b->b.k - b->n_ins;
    return dc::csr_apply(g, b->b, 0);
  });
}

int dc_recolor(dc_graph* h, int32_t rule, dc_stats* st) {
  return dc::guarded("dc_recolor", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    DC_HIP(hipSetDevice(g->device));
    if (st) {
      std::memset(st, 0, sizeof(*st));
      st->updates = g->staged_k;
    }
    // independent rule: the incremental loop (one launch per phase, or one
    // cooperative launch for the fused driver); ldf rule: the fused driver runs
    // the whole recolor in one cooperative launch
    const bool fused = g->opt.driver == DC_DRIVER_FUSED;
    int rc = rule == DC_RULE_INDEPENDENT ? dc::recolor_plain(g, rule, st, fused)
             : fused                     ? dc::recolor_fused(g, rule, st)
                                         : dc::recolor_plain(g, rule, st);
    if (st) {
      st->insertions = g->staged_ins;
      st->deletions = g->staged_del;
      st->regrows = g->staged_moves;   // batches whose rows outgrew their capacity
    }
    return rc;
  });
}

int dc_process_batch(dc_graph* h, int32_t rule, int64_t k, const int32_t* u, const int32_t* v,
                     const uint8_t* is_insert, dc_stats* st) {
  return dc::guarded("dc_process_batch", [&]() -> int {
    int rc = dc_apply_batch(h, k, u, v, is_insert);
    if (rc) return rc;
    return dc_recolor(h, rule, st);
  });
}

int dc_validate(dc_graph* h, int64_t* mono, int32_t* max_color) {
  return dc::guarded("dc_validate", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    DC_HIP(hipSetDevice(g->device));
    return dc::validate(g, mono, max_color);
  });
}

int dc_trace_get(dc_graph* h, int64_t* trace, int32_t cap, int32_t* n_phases) {
  return dc::guarded("dc_trace_get", [&]() -> int {
    if (!h || cap < 0 || (cap > 0 && !trace)) return DC_ERR_INVALID;
    Graph* g = G_(h);
    int32_t np = (int32_t)(g->trace.size() / 4);
    int32_t c = np < cap ? np : cap;
    if (c > 0) std::memcpy(trace, g->trace.data(), sizeof(int64_t) * 4 * (size_t)c);
    if (n_phases) *n_phases = np;
    return DC_OK;
  });
}

int dc_neighbors(dc_graph* h, int32_t vtx, int32_t* out, int64_t cap, int64_t* deg) {
  return dc::guarded("dc_neighbors", [&]() -> int {
    if (!h || cap < 0 || (cap > 0 && !out)) return DC_ERR_INVALID;
    Graph* g = G_(h);
    if (vtx < 0 || vtx >= g->n) return DC_ERR_VERTEX;
    DC_HIP(hipSetDevice(g->device));
    uint2 r;
    DC_HIP(hipMemcpy(&r, g->row + vtx, sizeof(uint2), hipMemcpyDeviceToHost));
    int32_t* tmp = new (std::nothrow) int32_t[r.y + 1];
    if (!tmp) return DC_ERR_NOMEM;
    if (r.y > 0) {
      hipError_t e = hipMemcpy(tmp, g->col + r.x, sizeof(int32_t) * r.y, hipMemcpyDeviceToHost);
      if (e != hipSuccess) {
        delete[] tmp;
        return dc::hip_fail(e, "dc_neighbors");
      }
    }
    int64_t d = 0;
    for (uint32_t j = 0; j < r.y; ++j)
      if (tmp[j] >= 0) {
        if (d < cap) out[d] = tmp[j];
        ++d;
      }
    delete[] tmp;
    if (deg) *deg = d;
    return DC_OK;
  });
}

int dc_prop_get(dc_graph* h, int32_t* prop) {
  return dc::guarded("dc_prop_get", [&]() -> int {
    if (!h || !prop) return DC_ERR_INVALID;
    Graph* g = G_(h);
    DC_HIP(hipSetDevice(g->device));
    DC_HIP(hipMemcpy(prop, g->prop, sizeof(int32_t) * (size_t)g->n, hipMemcpyDeviceToHost));
    return DC_OK;
  });
}

int dc_compact(dc_graph* h) {
  return dc::guarded("dc_compact", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    if (g->staged) {
      dc::set_error("dc_compact: a batch is staged");
      return DC_ERR_STATE;
    }
    DC_HIP(hipSetDevice(g->device));
    return dc::csr_compact(g);
  });
}

int dc_compact_noalloc(dc_graph* h) {
  return dc::guarded("dc_compact_noalloc", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    if (g->staged) {
      dc::set_error("dc_compact_noalloc: a batch is staged");
      return DC_ERR_STATE;
    }
    DC_HIP(hipSetDevice(g->device));
    return dc::csr_compact_noalloc(g);
  });
}

int dc_set_driver(dc_graph* h, int32_t driver, int32_t blocks_per_cu) {
  return dc::guarded("dc_set_driver", [&]() -> int {
    if (!h || (driver != DC_DRIVER_PLAIN && driver != DC_DRIVER_FUSED && driver != DC_DRIVER_GRAPH) ||
        blocks_per_cu < 0)
      return DC_ERR_INVALID;
    Graph* g = G_(h);
    g->opt.driver = driver;
    if (blocks_per_cu > 0 && blocks_per_cu != g->fused_blocks_per_cu) {
      g->fused_blocks_per_cu = blocks_per_cu;
      dc::fused_free(g);  // grid size is recomputed on the next fused launch
    }
    return DC_OK;
  });
}

int dc_set_parts(dc_graph* h, int32_t nparts, const int32_t* offsets) {
  return dc::guarded("dc_set_parts", [&]() -> int {
    if (!h || nparts < 0 || (nparts > 0 && !offsets)) return DC_ERR_INVALID;
    Graph* g = G_(h);
    if (g->staged) {
      dc::set_error("dc_set_parts: a batch is staged");
      return DC_ERR_STATE;
    }
    DC_HIP(hipSetDevice(g->device));
    return dc::parts_set(g, nparts, offsets);
  });
}

int dc_set_profiling(dc_graph* h, int32_t on) {
  return dc::guarded("dc_set_profiling", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    G_(h)->profiling = on != 0;
    return DC_OK;
  });
}

int dc_synchronize(dc_graph* h) {
  return dc::guarded("dc_synchronize", [&]() -> int {
    if (!h) return DC_ERR_INVALID;
    Graph* g = G_(h);
    DC_HIP(hipSetDevice(g->device));
    DC_HIP(hipStreamSynchronize(g->stream));
    return DC_OK;
  });
}

}  // extern "C"
