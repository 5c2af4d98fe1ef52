// serve.cpp -- continuous batching over resident coloured graphs.
//
// The resident graphs live in lanes: each lane is one device arena (a
// dc_graph holding the disjoint union of its graphs; graph g owns arena IDs
// [first[g], first[g] + size[g]) in lane lane_of[g]) with its own HIP stream
// and its own host worker thread.  The recolor rules only look at a vertex's
// own neighbourhood, so one recolor of an arena gives every graph exactly the
// colours it would get alone for the same batch (tests/test_gpu_serve.py).
//
// Requests (graph, lines, arrival time) are replayed against the wall clock.
// The admission thread takes them in arrival order as they arrive and admits
// each while the device bytes in flight (staged lines, their change lists and
// the row slots their insertions may take) fit the budget; otherwise it waits
// until a step completes and frees its bytes.  A request larger than the
// whole budget is rejected.  An admitted request joins its lane's pending
// set at once, also while that lane's step runs.  Each lane worker loops:
// take every pending request, coalesce them per graph (arrival order) into
// one batch, apply it to the arena and recolor it (one launch sequence for
// all its graphs), then release the bytes.  Lanes run concurrently, so their
// steps overlap on the device, and a request waits only for its own lane's
// current step.  Latency = completion of the request's step - arrival.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dyncolor.h"
#include "dc_guard.h"

namespace {

// Device bytes a request may occupy while in flight: per line its staged copy
// (u, v, flag: 9 B), its entries in the four change lists and the overflow
// list (64 B of list capacity), its two deletion sort keys (16 B), and per
// insertion the two row slots it adds
// (neighbour, slot colour, reverse link: 2 x 12 B) in case they land in the
// spare tail.
constexpr int64_t kBytesPerLine = 9 + 64 + 16;
constexpr int64_t kBytesPerInsertion = 24;

struct Lane {
  dc_graph* arena = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  std::vector<int64_t> pending;   // admitted requests, arrival order
  bool stop = false;
};

}  // namespace

struct dc_server {
  std::vector<dc_graph*> arenas;
  std::vector<int32_t> lane_of, first, size;
  int32_t rule = DC_RULE_INDEPENDENT;
  int64_t budget = 0;
};

extern "C" {

int dc_server_create(int32_t nlanes, dc_graph* const* arenas, int32_t ngraphs,
                     const int32_t* lane_of, const int32_t* first, const int32_t* size,
                     int64_t budget_bytes, int32_t rule, dc_server** out) {
  return dc::guarded("dc_server_create", [&]() -> int {
    if (nlanes < 1 || !arenas || ngraphs < 1 || !lane_of || !first || !size || budget_bytes < 0 ||
        !out || (rule != DC_RULE_INDEPENDENT && rule != DC_RULE_LDF)) {
      dc::set_error("dc_server_create: invalid argument");
      return DC_ERR_INVALID;
    }
    *out = nullptr;
    std::vector<int32_t> ln(nlanes, 0), lno(nlanes, 0);
    for (int32_t l = 0; l < nlanes; ++l) {
      if (!arenas[l]) {
        dc::set_error("dc_server_create: null arena");
        return DC_ERR_INVALID;
      }
      int rc = dc_graph_info(arenas[l], &ln[l], &lno[l], nullptr, nullptr);
      if (rc) return rc;
    }
    // per lane, its graphs sorted by first vertex must not overlap
    std::vector<std::vector<std::pair<int32_t, int32_t>>> spans(nlanes);
    for (int32_t g = 0; g < ngraphs; ++g) {
      const int32_t l = lane_of[g];
      if (l < 0 || l >= nlanes || first[g] < 0 || size[g] < 0 ||
          (int64_t)first[g] + size[g] > ln[l]) {
        dc::set_error("dc_server_create: graph " + std::to_string(g) + " is outside its arena");
        return DC_ERR_INVALID;
      }
      spans[l].push_back({first[g], size[g]});
    }
    for (int32_t l = 0; l < nlanes; ++l) {
      auto& sp = spans[l];
      std::sort(sp.begin(), sp.end());
      for (size_t i = 1; i < sp.size(); ++i)
        if (sp[i - 1].first + sp[i - 1].second > sp[i].first) {
          dc::set_error("dc_server_create: graphs overlap in lane " + std::to_string(l));
          return DC_ERR_INVALID;
        }
      // the lane's graphs are its arena's parts when they tile it: a graph
      // whose colours settle leaves the loop while the others go on (an arena
      // with edges between its graphs keeps one loop: parts only save work)
      std::vector<int32_t> parts{0};
      bool tiles = lno[l] == ln[l];
      for (auto& s : sp) {
        if (s.second == 0) continue;
        tiles = tiles && s.first == parts.back();
        parts.push_back(s.first + s.second);
      }
      tiles = tiles && parts.back() == ln[l];
      if (tiles && parts.size() > 1) {
        if (dc_set_parts(arenas[l], (int32_t)parts.size() - 1, parts.data()) != DC_OK)
          (void)dc_set_parts(arenas[l], 0, nullptr);
      }
    }
    int64_t budget = budget_bytes;
    if (budget == 0) {
      // a quarter of the device memory free now
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr == 0) {
        dc::set_error("dc_server_create: cannot read free device memory (pass a budget)");
        return DC_ERR_HIP;
      }
      budget = (int64_t)(fr / 4);
    }
    auto s = std::make_unique<dc_server>();
    s->arenas.assign(arenas, arenas + nlanes);
    s->lane_of.assign(lane_of, lane_of + ngraphs);
    s->first.assign(first, first + ngraphs);
    s->size.assign(size, size + ngraphs);
    s->rule = rule;
    s->budget = budget;
    *out = s.release();
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

int64_t dc_server_request_bytes(int64_t lines, int64_t insertions) {
  if (lines < 0 || insertions < 0 || insertions > lines) return -1;
  return kBytesPerLine * lines + kBytesPerInsertion * insertions;
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
    const int32_t ng = (int32_t)s->lane_of.size();
    const int32_t nl = (int32_t)s->arenas.size();
    std::vector<int64_t> bytes((size_t)nreq, 0);
    for (int64_t i = 0; i < nreq; ++i) {
      if (gid[i] < 0 || gid[i] >= ng || roff[i + 1] < roff[i] ||
          (i > 0 && arrival[i] < arrival[i - 1])) {
        dc::set_error("dc_server_run: bad request " + std::to_string(i));
        return DC_ERR_INVALID;
      }
      const int32_t gn = s->size[gid[i]];
      int64_t ni = 0;
      for (int64_t j = roff[i]; j < roff[i + 1]; ++j) {
        if (u[j] < 0 || v[j] < 0 || u[j] >= gn || v[j] >= gn || u[j] == v[j] || ins[j] > 1) {
          dc::set_error("dc_server_run: bad line in request " + std::to_string(i));
          return DC_ERR_VERTEX;
        }
        ni += ins[j];
      }
      bytes[(size_t)i] = dc_server_request_bytes(roff[i + 1] - roff[i], ni);
    }
    for (int64_t i = 0; i < nreq; ++i) {
      latency[i] = -1.0;
      if (step_of_req) step_of_req[i] = -1;
    }
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };

    std::vector<Lane> lanes((size_t)nl);
    for (int32_t l = 0; l < nl; ++l) lanes[(size_t)l].arena = s->arenas[(size_t)l];
    // admission state: bytes in flight, guarded by amu; a completed step
    // signals acv
    std::mutex amu;
    std::condition_variable acv;
    int64_t inflight = 0, max_inflight = 0;
    std::atomic<int32_t> step_ids{0};
    std::atomic<int64_t> lines_done{0}, max_step_lines{0}, served{0};
    std::atomic<int> failure{DC_OK};
    std::string fail_msg;
    std::vector<double> busy((size_t)nl, 0.0);

    auto worker = [&](int32_t l) {
      Lane& L = lanes[(size_t)l];
      std::vector<int64_t> take;
      std::vector<std::vector<int64_t>> per;
      std::vector<int32_t> bu, bv;
      std::vector<uint8_t> bi;
      for (;;) {
        {
          std::unique_lock<std::mutex> lk(L.mu);
          L.cv.wait(lk, [&] { return L.stop || !L.pending.empty(); });
          if (L.pending.empty()) return;   // stop with nothing left
          take.swap(L.pending);
          L.pending.clear();
        }
        // coalesce per graph, arrival order within a graph
        per.assign((size_t)ng, {});
        for (int64_t r : take) per[(size_t)gid[r]].push_back(r);
        bu.clear();
        bv.clear();
        bi.clear();
        int64_t freed = 0;
        for (int32_t g = 0; g < ng; ++g) {
          const int32_t off = s->first[(size_t)g];
          for (int64_t r : per[(size_t)g]) {
            for (int64_t j = roff[r]; j < roff[r + 1]; ++j) {
              bu.push_back(u[j] + off);
              bv.push_back(v[j] + off);
              bi.push_back(ins[j]);
            }
            freed += bytes[(size_t)r];
          }
        }
        const double ts = now();
        int rc = dc_apply_batch(L.arena, (int64_t)bu.size(), bu.data(), bv.data(), bi.data());
        if (rc == DC_OK) rc = dc_recolor(L.arena, s->rule, nullptr);
        const double te = now();
        if (rc != DC_OK) {
          int expect = DC_OK;
          if (failure.compare_exchange_strong(expect, rc)) fail_msg = dc_last_error();
        }
        busy[(size_t)l] += te - ts;
        const int32_t sid = step_ids.fetch_add(1);
        for (int64_t r : take) {
          latency[r] = te - arrival[r];
          if (step_of_req) step_of_req[r] = sid;
        }
        served += (int64_t)take.size();
        lines_done += (int64_t)bu.size();
        int64_t m = max_step_lines.load();
        while ((int64_t)bu.size() > m && !max_step_lines.compare_exchange_weak(m, (int64_t)bu.size())) {
        }
        take.clear();
        {
          std::lock_guard<std::mutex> lk(amu);
          inflight -= freed;
        }
        acv.notify_all();
        if (rc != DC_OK) return;
      }
    };
    std::vector<std::thread> threads;
    threads.reserve((size_t)nl);
    for (int32_t l = 0; l < nl; ++l) threads.emplace_back(worker, l);

    int64_t rejected = 0, waits = 0;
    for (int64_t i = 0; i < nreq && failure.load() == DC_OK; ++i) {
      const double wait = arrival[i] - now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      if (bytes[(size_t)i] > s->budget) {
        ++rejected;
        continue;
      }
      {
        std::unique_lock<std::mutex> lk(amu);
        if (inflight + bytes[(size_t)i] > s->budget) {
          ++waits;
          acv.wait(lk, [&] {
            return inflight + bytes[(size_t)i] <= s->budget || failure.load() != DC_OK;
          });
        }
        inflight += bytes[(size_t)i];
        max_inflight = std::max(max_inflight, inflight);
      }
      Lane& L = lanes[(size_t)s->lane_of[(size_t)gid[i]]];
      {
        std::lock_guard<std::mutex> lk(L.mu);
        L.pending.push_back(i);
      }
      L.cv.notify_one();
    }
    for (auto& L : lanes) {
      {
        std::lock_guard<std::mutex> lk(L.mu);
        L.stop = true;
      }
      L.cv.notify_one();
    }
    for (auto& t : threads) t.join();
    if (failure.load() != DC_OK) {
      dc::set_error("dc_server_run: a step failed: " + fail_msg);
      return failure.load();
    }
    if (st) {
      std::memset(st, 0, sizeof(*st));
      st->steps = step_ids.load();
      st->lanes = nl;
      st->served = served.load();
      st->rejected = rejected;
      st->lines = lines_done.load();
      st->max_step_lines = max_step_lines.load();
      st->elapsed_s = now();
      for (double b : busy) st->busy_s += b;
      st->budget_bytes = s->budget;
      st->max_inflight_bytes = max_inflight;
      st->admit_waits = waits;
    }
    return DC_OK;
  });
}

int dc_lanes_step(int32_t nlanes, dc_graph* const* arenas, const dc_batch* const* batches,
                  int32_t rule, dc_stats* stats) {
  return dc::guarded("dc_lanes_step", [&]() -> int {
    if (nlanes < 1 || !arenas || !batches) {
      dc::set_error("dc_lanes_step: invalid argument");
      return DC_ERR_INVALID;
    }
    for (int32_t l = 0; l < nlanes; ++l) {
      if (!arenas[l] || !batches[l]) {
        dc::set_error("dc_lanes_step: null arena or batch");
        return DC_ERR_INVALID;
      }
      for (int32_t k = 0; k < l; ++k)
        if (arenas[k] == arenas[l]) {
          dc::set_error("dc_lanes_step: an arena appears twice");
          return DC_ERR_INVALID;
        }
    }
    // lane l on its own host thread (lane 0 on the caller's): each arena has
    // its own stream, so the lanes' launches overlap on the device
    std::vector<int> rc((size_t)nlanes, DC_OK);
    std::vector<std::string> msg((size_t)nlanes);
    auto lane = [&](int32_t l) {
      int r = dc_apply_device_batch(arenas[l], batches[l]);
      if (r == DC_OK) r = dc_recolor(arenas[l], rule, stats ? &stats[l] : nullptr);
      if (r == DC_OK) r = dc_synchronize(arenas[l]);
      rc[(size_t)l] = r;
      if (r != DC_OK) msg[(size_t)l] = dc_last_error();
    };
    std::vector<std::thread> th;
    th.reserve((size_t)nlanes - 1);
    try {
      for (int32_t l = 1; l < nlanes; ++l) th.emplace_back(lane, l);
    } catch (...) {
      for (auto& t : th) t.join();
      throw;
    }
    lane(0);
    for (auto& t : th) t.join();
    for (int32_t l = 0; l < nlanes; ++l)
      if (rc[(size_t)l] != DC_OK) {
        dc::set_error("dc_lanes_step: lane " + std::to_string(l) + ": " + msg[(size_t)l]);
        return rc[(size_t)l];
      }
    return DC_OK;
  });
}

}  // extern "C"
