// partition.cpp -- one partition's share from global vertex IDs (host only).
//
// The multi-GPU variant of the reference gives every rank the list of its
// vertices (partition<r>.txt, one global ID per line: line i becomes local ID
// i, MG/Supporting.h:29-46), the part of every vertex (partitionAll.txt,
// MG/Supporting.h:49-63), and the whole graph, colour and change files.  Each
// rank then keeps the edges with an owned endpoint, splitting them into an
// internal CSR and a border CSR towards "other part vertices" that get local
// IDs in order of first appearance (read_graphEdges, MG/all_structure_undir.
// cuh:229-312), splits the change lines into internal/cross insertion and
// deletion lists with cross lines as (owned, other) (readin_changes,
// MG/all_structure_undir.cuh:78-192), and looks the colours of both kinds up
// (read_Input_Color, MG/all_structure_undir.cuh:199-222).
//
// Here the same maps become a plan: local IDs [0, n_owned) are the owned
// vertices in partition-file order, [n_owned, n_owned + n_ghost) the other
// parts' vertices in order of first appearance (graph lines, then change
// lines; within a line the non-owned endpoint of the first owned one), and the
// edge and change lists in those local IDs, ready for dc_partition_create and
// dc_apply_batch.  Lookups go through one flat global -> local array instead
// of std::map.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dyncolor.h"
#include "dc_guard.h"

struct dc_partition_plan {
  int32_t n_global = 0, rank = 0, n_owned = 0, n_ghost = 0, n_parts = 0;
  std::vector<int32_t> l2g;            // local -> global
  std::vector<int32_t> u, v;           // local edges: internal, then cross (owned, ghost), file order
  std::vector<int32_t> bu, bv;         // local change lines, file order, cross lines (owned, ghost)
  std::vector<uint8_t> bins;
  int64_t m_internal = 0, m_cross = 0, m_dropped = 0;
  int64_t k_del = 0, k_ins = 0, k_xdel = 0, k_xins = 0, k_dropped = 0;
  int32_t n_border = 0;
  std::vector<int64_t> border_per_part;  // owned vertices with a neighbour in part p
};

namespace {

int bad(int code, const std::string& msg) {
  dc::set_error(msg);
  return code;
}

}  // namespace

extern "C" {

int dc_partition_plan_create(int32_t n_global, int32_t rank, int32_t n_owned,
                             const int32_t* owned_ids, const int32_t* part_of, int64_t m,
                             const int32_t* u, const int32_t* v, int64_t k, const int32_t* cu,
                             const int32_t* cv, const uint8_t* cins, dc_partition_plan** out) {
  return dc::guarded("dc_partition_plan_create", [&]() -> int {
    if (!out || n_global < 1 || n_owned < 0 || n_owned > n_global || m < 0 || k < 0 ||
        (n_owned > 0 && !owned_ids) || (m > 0 && (!u || !v)) || (k > 0 && (!cu || !cv || !cins)))
      return bad(DC_ERR_INVALID, "dc_partition_plan_create: invalid argument");
    *out = nullptr;
    auto p = std::make_unique<dc_partition_plan>();
    p->n_global = n_global;
    p->rank = rank;
    p->n_owned = n_owned;
    // global -> local; -1 = not seen
    std::vector<int32_t> g2l((size_t)n_global, -1);
    p->l2g.reserve((size_t)n_owned);
    for (int32_t i = 0; i < n_owned; ++i) {
      const int32_t x = owned_ids[i];
      if (x < 0 || x >= n_global)
        return bad(DC_ERR_VERTEX, "partition vertex " + std::to_string(x) + " out of range at line " +
                                      std::to_string(i));
      if (g2l[(size_t)x] >= 0)
        return bad(DC_ERR_INVALID, "vertex " + std::to_string(x) + " listed twice in the partition");
      g2l[(size_t)x] = i;
      p->l2g.push_back(x);
    }
    if (part_of) {
      int32_t hi = -1;
      for (int32_t x = 0; x < n_global; ++x) {
        if (part_of[x] < 0)
          return bad(DC_ERR_INVALID, "negative part ID for vertex " + std::to_string(x));
        hi = std::max(hi, part_of[x]);
      }
      p->n_parts = hi + 1;
      for (int32_t i = 0; i < n_owned; ++i)
        if (part_of[owned_ids[i]] != rank)
          return bad(DC_ERR_INVALID, "vertex " + std::to_string(owned_ids[i]) +
                                         " is in the partition file but in part " +
                                         std::to_string(part_of[owned_ids[i]]) + ", not " +
                                         std::to_string(rank));
    }
    auto owned = [&](int32_t x) { return g2l[(size_t)x] >= 0 && g2l[(size_t)x] < n_owned; };
    auto ghost_id = [&](int32_t x) {
      int32_t& l = g2l[(size_t)x];
      if (l < 0) {
        l = n_owned + p->n_ghost++;
        p->l2g.push_back(x);
      }
      return l;
    };
    auto check_line = [&](int32_t a, int32_t b, const char* what, int64_t i) -> int {
      if (a < 0 || b < 0 || a >= n_global || b >= n_global)
        return bad(DC_ERR_VERTEX, std::string(what) + " line " + std::to_string(i) +
                                      ": vertex ID out of range");
      if (a == b)
        return bad(DC_ERR_VERTEX, std::string(what) + " line " + std::to_string(i) + ": self loop");
      return DC_OK;
    };
    // graph lines: internal edges, then cross edges in file order
    std::vector<int32_t> xu, xv;
    for (int64_t i = 0; i < m; ++i) {
      const int32_t a = u[i], b = v[i];
      if (int rc = check_line(a, b, "graph", i)) return rc;
      const bool oa = owned(a), ob = owned(b);
      if (oa && ob) {
        p->u.push_back(g2l[(size_t)a]);
        p->v.push_back(g2l[(size_t)b]);
        ++p->m_internal;
      } else if (oa || ob) {
        const int32_t o = oa ? a : b, x = oa ? b : a;
        xu.push_back(g2l[(size_t)o]);
        xv.push_back(ghost_id(x));
        ++p->m_cross;
      } else {
        ++p->m_dropped;
      }
    }
    p->u.insert(p->u.end(), xu.begin(), xu.end());
    p->v.insert(p->v.end(), xv.begin(), xv.end());
    // change lines in file order; lines without an owned endpoint belong to
    // other partitions
    for (int64_t i = 0; i < k; ++i) {
      const int32_t a = cu[i], b = cv[i];
      if (int rc = check_line(a, b, "change", i)) return rc;
      if (cins[i] > 1)
        return bad(DC_ERR_INVALID, "change line " + std::to_string(i) + ": insert flag must be 0 or 1");
      const bool oa = owned(a), ob = owned(b);
      if (!oa && !ob) {
        ++p->k_dropped;
        continue;
      }
      const bool cross = !(oa && ob);
      const int32_t la = oa ? g2l[(size_t)a] : g2l[(size_t)b];
      const int32_t lb = cross ? ghost_id(oa ? b : a) : g2l[(size_t)b];
      p->bu.push_back(la);
      p->bv.push_back(lb);
      p->bins.push_back(cins[i]);
      if (cins[i]) {
        ++(cross ? p->k_xins : p->k_ins);
      } else {
        ++(cross ? p->k_xdel : p->k_del);
      }
    }
    // border vertices: owned vertices with a ghost neighbour (a graph edge or
    // an inserted cross line), in total and per part of that neighbour
    // (distinct (part, owned vertex) pairs)
    std::vector<uint8_t> is_border((size_t)n_owned, 0);
    std::vector<uint64_t> pairs;
    auto cross_pair = [&](int32_t lo, int32_t lg) {
      is_border[(size_t)lo] = 1;
      if (p->n_parts > 0)
        pairs.push_back(((uint64_t)(uint32_t)part_of[p->l2g[(size_t)lg]] << 32) | (uint32_t)lo);
    };
    for (size_t i = (size_t)p->m_internal; i < p->u.size(); ++i) cross_pair(p->u[i], p->v[i]);
    for (size_t i = 0; i < p->bu.size(); ++i)
      if (p->bins[i] && p->bv[i] >= n_owned) cross_pair(p->bu[i], p->bv[i]);
    for (uint8_t b : is_border) p->n_border += b;
    if (p->n_parts > 0) {
      p->border_per_part.assign((size_t)p->n_parts, 0);
      std::sort(pairs.begin(), pairs.end());
      pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
      for (uint64_t q : pairs) ++p->border_per_part[(size_t)(q >> 32)];
    }
    if ((int64_t)n_owned + p->n_ghost > 0x7ffffffe)
      return bad(DC_ERR_INVALID, "partition share exceeds int32 local IDs");
    *out = p.release();
    return DC_OK;
  });
}

int dc_partition_plan_info(const dc_partition_plan* p, dc_partition_info* info) {
  return dc::guarded("dc_partition_plan_info", [&]() -> int {
    if (!p || !info) return bad(DC_ERR_INVALID, "dc_partition_plan_info: invalid argument");
    std::memset(info, 0, sizeof(*info));
    info->n_global = p->n_global;
    info->rank = p->rank;
    info->n_owned = p->n_owned;
    info->n_ghost = p->n_ghost;
    info->n_border = p->n_border;
    info->n_parts = p->n_parts;
    info->m_local = (int64_t)p->u.size();
    info->m_internal = p->m_internal;
    info->m_cross = p->m_cross;
    info->m_dropped = p->m_dropped;
    info->k_local = (int64_t)p->bu.size();
    info->k_del = p->k_del;
    info->k_ins = p->k_ins;
    info->k_xdel = p->k_xdel;
    info->k_xins = p->k_xins;
    info->k_dropped = p->k_dropped;
    return DC_OK;
  });
}

int dc_partition_plan_arrays(const dc_partition_plan* p, int32_t* local_to_global, int32_t* u,
                             int32_t* v, int32_t* bu, int32_t* bv, uint8_t* bins) {
  return dc::guarded("dc_partition_plan_arrays", [&]() -> int {
    if (!p) return bad(DC_ERR_INVALID, "dc_partition_plan_arrays: invalid argument");
    auto put = [](auto* dst, const auto& src) {
      if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(src[0]) * src.size());
    };
    put(local_to_global, p->l2g);
    put(u, p->u);
    put(v, p->v);
    put(bu, p->bu);
    put(bv, p->bv);
    put(bins, p->bins);
    return DC_OK;
  });
}

int dc_partition_plan_border(const dc_partition_plan* p, int64_t* per_part, int32_t nparts) {
  return dc::guarded("dc_partition_plan_border", [&]() -> int {
    if (!p || nparts < 0 || (nparts > 0 && !per_part))
      return bad(DC_ERR_INVALID, "dc_partition_plan_border: invalid argument");
    if (p->n_parts == 0) return bad(DC_ERR_STATE, "dc_partition_plan_border: no part vector given");
    for (int32_t i = 0; i < nparts; ++i)
      per_part[i] = i < p->n_parts ? p->border_per_part[(size_t)i] : 0;
    return DC_OK;
  });
}

int dc_partition_create_global(const dc_partition_plan* p, const int32_t* colors_global,
                               const dc_options* opt, dc_graph** out) {
  return dc::guarded("dc_partition_create_global", [&]() -> int {
    if (!p || !colors_global || !out)
      return bad(DC_ERR_INVALID, "dc_partition_create_global: invalid argument");
    *out = nullptr;
    const size_t n = p->l2g.size();
    std::vector<int32_t> local(n);
    for (size_t i = 0; i < n; ++i) {
      const int32_t c = colors_global[p->l2g[i]];
      if (c < 0)
        return bad(DC_ERR_INVALID, std::string("no colour for ") + (i < (size_t)p->n_owned ? "owned" : "ghost") +
                                       " vertex " + std::to_string(p->l2g[i]));
      local[i] = c;
    }
    dc_graph* g = nullptr;
    int rc = dc_partition_create(p->n_owned, p->n_ghost, (int64_t)p->u.size(), p->u.data(), p->v.data(),
                                 local.data() + p->n_owned, opt, &g);
    if (rc) return rc;
    if ((rc = dc_colors_set(g, local.data()))) {
      dc_graph_destroy(g);
      return rc;
    }
    *out = g;
    return DC_OK;
  });
}

int dc_partition_plan_destroy(dc_partition_plan* p) {
  return dc::guarded("dc_partition_plan_destroy", [&]() -> int {
    if (!p) return DC_ERR_INVALID;
    delete p;
    return DC_OK;
  });
}

}  // extern "C"
