// serve.cpp -- continuous-batching request loop over resident colored graphs.
//
// The resident graphs live in one device arena: a single dc_graph holding
// their disjoint union (graph g owns vertex IDs [off[g], off[g+1])).  The
// recolor rules only look at a vertex's own neighbourhood, so one recolor of
// the arena gives every graph exactly the colors it would get alone for the
// same batch (tests/test_gpu_serve.py checks this).
//
// Loop: requests (graph, lines, arrival time) are replayed against the wall
// clock.  Arrived requests are admitted in arrival order while the lines in
// flight fit the memory budget; a request larger than the whole budget is
// rejected.  Each step takes every graph with admitted requests, coalesces
// that graph's requests (arrival order) into one batch, applies the union of
// those batches to the arena and recolors it in one launch, so different
// graphs' work overlaps on the device.  Requests arriving during a step join
// the next one.  Latency = step completion - arrival.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dyncolor.h"
#include "dc_guard.h"

struct dc_server {
  dc_graph* arena = nullptr;
  std::vector<int32_t> off;
  int32_t rule = DC_RULE_INDEPENDENT;
  int64_t budget_lines = 0;
};

extern "C" {

int dc_server_create(dc_graph* arena, int32_t ngraphs, const int32_t* offsets,
                     int64_t budget_lines, int32_t rule, dc_server** out) {
  return dc::guarded("dc_server_create", [&]() -> int {
    if (!arena || ngraphs < 1 || !offsets || budget_lines < 1 || !out ||
        (rule != DC_RULE_INDEPENDENT && rule != DC_RULE_LDF)) {
      dc::set_error("dc_server_create: invalid argument");
      return DC_ERR_INVALID;
    }
    int32_t n = 0, n_owned = 0;
    int rc = dc_graph_info(arena, &n, &n_owned, nullptr, nullptr);
    if (rc) return rc;
    if (offsets[0] != 0 || offsets[ngraphs] != n) {
      dc::set_error("dc_server_create: offsets must cover the arena's vertices");
      return DC_ERR_INVALID;
    }
    for (int32_t i = 0; i < ngraphs; ++i)
      if (offsets[i + 1] < offsets[i]) {
        dc::set_error("dc_server_create: offsets must be non-decreasing");
        return DC_ERR_INVALID;
      }
    // the graphs are the arena's parts (empty graphs dropped): a graph whose
    // colours settle leaves the loop while the others go on.  An arena with
    // edges between its graphs keeps one loop (parts are only an optimisation).
    {
      std::vector<int32_t> parts{0};
      for (int32_t i = 1; i <= ngraphs; ++i)
        if (offsets[i] > parts.back()) parts.push_back(offsets[i]);
      if (n_owned == n && dc_set_parts(arena, (int32_t)parts.size() - 1, parts.data()) != DC_OK)
        (void)dc_set_parts(arena, 0, nullptr);
    }
    dc_server* s = new (std::nothrow) dc_server();
    if (!s) return DC_ERR_NOMEM;
    s->arena = arena;
    s->off.assign(offsets, offsets + ngraphs + 1);
    s->rule = rule;
    s->budget_lines = budget_lines;
    *out = s;
    return DC_OK;
  });
}

int dc_server_destroy(dc_server* s) {
  return dc::guarded("dc_server_destroy", [&]() -> int {
    if (!s) return DC_ERR_INVALID;
    delete s;
    return DC_OK;
  });
}

int dc_server_run(dc_server* s, int64_t nreq, const int32_t* gid, const int64_t* roff,
                  const int32_t* u, const int32_t* v, const uint8_t* ins, const double* arrival,
                  double* latency, int32_t* step_of_req, dc_serve_stats* st) {
  return dc::guarded("dc_server_run", [&]() -> int {
    if (!s || nreq < 0 || (nreq > 0 && (!gid || !roff || !arrival || !latency)) ||
        (nreq > 0 && roff[nreq] > roff[0] && (!u || !v || !ins))) {
      dc::set_error("dc_server_run: invalid argument");
      return DC_ERR_INVALID;
    }
    const int32_t ng = (int32_t)s->off.size() - 1;
    for (int64_t i = 0; i < nreq; ++i) {
      if (gid[i] < 0 || gid[i] >= ng || roff[i + 1] < roff[i] || (i > 0 && arrival[i] < arrival[i - 1])) {
        dc::set_error("dc_server_run: bad request " + std::to_string(i));
        return DC_ERR_INVALID;
      }
      const int32_t gn = s->off[gid[i] + 1] - s->off[gid[i]];
      for (int64_t j = roff[i]; j < roff[i + 1]; ++j)
        if (u[j] < 0 || v[j] < 0 || u[j] >= gn || v[j] >= gn || u[j] == v[j] || ins[j] > 1) {
          dc::set_error("dc_server_run: bad line in request " + std::to_string(i));
          return DC_ERR_VERTEX;
        }
    }
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };

    std::vector<std::vector<int64_t>> pending(ng);
    int64_t next = 0, finished = 0, inflight = 0;
    int32_t steps = 0;
    int64_t rejected = 0, lines_done = 0, max_step_lines = 0;
    double busy = 0.0;
    std::vector<int32_t> bu, bv;
    std::vector<uint8_t> bi;
    std::vector<int64_t> members;
    for (int64_t i = 0; i < nreq; ++i) {
      latency[i] = -1.0;
      if (step_of_req) step_of_req[i] = -1;
    }

    while (finished + rejected < nreq) {
      const double t = now();
      // admission in arrival order under the line budget
      while (next < nreq && arrival[next] <= t) {
        const int64_t lines = roff[next + 1] - roff[next];
        if (lines > s->budget_lines) {
          ++rejected;
          ++next;
          continue;
        }
        if (inflight + lines > s->budget_lines) break;
        pending[gid[next]].push_back(next);
        inflight += lines;
        ++next;
      }
      bool any = false;
      for (int32_t g = 0; g < ng && !any; ++g) any = !pending[g].empty();
      if (!any) {
        if (next < nreq) {
          const double wait = arrival[next] - now();
          if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
        }
        continue;
      }
      // one step: every graph with admitted requests, each coalesced
      bu.clear();
      bv.clear();
      bi.clear();
      members.clear();
      for (int32_t g = 0; g < ng; ++g) {
        for (int64_t r : pending[g]) {
          for (int64_t j = roff[r]; j < roff[r + 1]; ++j) {
            bu.push_back(u[j] + s->off[g]);
            bv.push_back(v[j] + s->off[g]);
            bi.push_back(ins[j]);
          }
          members.push_back(r);
        }
        pending[g].clear();
      }
      const double ts = now();
      int rc = dc_apply_batch(s->arena, (int64_t)bu.size(), bu.data(), bv.data(), bi.data());
      if (rc == DC_OK) rc = dc_recolor(s->arena, s->rule, nullptr);
      if (rc != DC_OK) return rc;
      const double te = now();
      busy += te - ts;
      for (int64_t r : members) {
        latency[r] = te - arrival[r];
        if (step_of_req) step_of_req[r] = steps;
      }
      finished += (int64_t)members.size();
      lines_done += (int64_t)bu.size();
      max_step_lines = std::max<int64_t>(max_step_lines, (int64_t)bu.size());
      inflight -= (int64_t)bu.size();
      ++steps;
    }
    if (st) {
      std::memset(st, 0, sizeof(*st));
      st->steps = steps;
      st->served = finished;
      st->rejected = rejected;
      st->lines = lines_done;
      st->max_step_lines = max_step_lines;
      st->elapsed_s = now();
      st->busy_s = busy;
    }
    return DC_OK;
  });
}

}  // extern "C"
