// dyncolor_cli -- command-line drop-in for the reference program.
//
//   dyncolor_cli <graph file> <nodes> <edges> <color file> <change file>
//                [--rule independent|ldf] [--driver plain|fused] [--out colors.txt]
//   dyncolor_cli <graph file> <color file> <change file>
//                [--rank r] [--dir partition-file directory] [--driver plain|fused]
//                [--out colors.txt]
//
// Five positional arguments: the single-GPU programs (HI/Color_main.cu:21-38,
// HB/Color_main.cu:21-38).  Three: one rank of the multi-GPU program
// (MG/kernel.cu:43-49), which also reads partition<rank>.txt (the rank's
// vertices, one global ID per line) and partitionAll.txt (the part of every
// vertex) from the working directory (MG/kernel.cu:140-160); the rank comes
// from --rank or the MPI launcher's environment (OMPI_COMM_WORLD_RANK,
// PMI_RANK, PMIX_RANK, SLURM_PROCID), else 0, and the partition runs the
// heuristic-independent rules with its ghost colours frozen.  Text formats
// as in the reference: graph lines "a b w" (read_graphEdges,
// HI/all_structure_undir.cuh:156-181), color lines "node color"
// (read_Input_Color, :130-149), change lines "a b w inst" with inst 1 =
// insert, 0 = delete (readin_changes, :78-123).  Edge weights are read and
// ignored, as in the reference.  Prints the reference's phase timing lines;
// --out writes the final coloring as "node color" lines (global IDs of the
// owned vertices in the multi-GPU mode; the reference computes the colours
// but never writes them, HI/Color_main.cu:211-216).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/dyncolor.h"

using clk = std::chrono::high_resolution_clock;

static double us_since(clk::time_point t) {
  return std::chrono::duration<double, std::micro>(clk::now() - t).count();
}

static bool read_ints(const char* path, int want, std::vector<std::vector<long long>>& rows) {
  FILE* f = std::fopen(path, "r");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path);
    return false;
  }
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    long long x[4] = {0, 0, 0, 0};
    int got = std::sscanf(line, "%lld %lld %lld %lld", &x[0], &x[1], &x[2], &x[3]);
    if (got <= 0) continue;  // blank line
    if (got < want) {
      std::fprintf(stderr, "%s: malformed line: %s", path, line);
      std::fclose(f);
      return false;
    }
    rows.emplace_back(x, x + want);
  }
  std::fclose(f);
  return true;
}

static int die(int rc, const char* what) {
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, dc_last_error());
  return 1;
}

static void print_trace(dc_graph* h, const dc_stats& st) {
  std::vector<int64_t> trace((size_t)st.phases * 4 + 4);
  int32_t np = 0;
  dc_trace_get(h, trace.data(), st.phases, &np);
  for (int32_t i = 0; i < np && i < st.phases; ++i) {
    const int64_t* p = &trace[(size_t)i * 4];
    if (p[0] == 1) std::printf("deletion round: %lld targets, %lld recolored\n", (long long)p[1], (long long)p[2]);
    if (p[0] == 2) std::printf("insertion round: %lld targets, %lld changed\n", (long long)p[1], (long long)p[2]);
    if (p[0] == 3) std::printf("change-edge round: %lld endpoints, %lld recolored\n", (long long)p[1], (long long)p[2]);
    if (p[0] == 4) std::printf("cross deletion round: %lld targets, %lld recolored\n", (long long)p[1], (long long)p[2]);
    if (p[0] == 5) std::printf("cross insertion round: %lld targets, %lld changed\n", (long long)p[1], (long long)p[2]);
  }
  std::printf("recolor loop: %d rounds, %lld evaluations, %lld changes\n", st.loop_rounds,
              (long long)st.evaluated, (long long)st.changed);
}

static int single_gpu(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr,
                 "usage: %s <graph> <nodes> <edges> <colors> <changes> [--rule independent|ldf] "
                 "[--driver plain|fused] [--out file]\n",
                 argv[0]);
    return 2;
  }
  const long long nodes = std::atoll(argv[2]);
  int32_t rule = DC_RULE_INDEPENDENT, driver = DC_DRIVER_PLAIN;
  const char* out_path = nullptr;
  for (int i = 6; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--rule") && i + 1 < argc) {
      rule = !std::strcmp(argv[++i], "ldf") ? DC_RULE_LDF : DC_RULE_INDEPENDENT;
    } else if (!std::strcmp(argv[i], "--driver") && i + 1 < argc) {
      driver = !std::strcmp(argv[++i], "fused") ? DC_DRIVER_FUSED : DC_DRIVER_PLAIN;
    } else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) {
      out_path = argv[++i];
    } else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    }
  }
  if (nodes <= 0 || nodes > 0x7ffffffeLL) {
    std::fprintf(stderr, "bad node count %s\n", argv[2]);
    return 2;
  }
  const int32_t n = (int32_t)nodes;

  std::printf("Reading input graph...\n");
  auto t = clk::now();
  std::vector<std::vector<long long>> g, c, ch;
  if (!read_ints(argv[1], 2, g)) return 1;
  std::printf("Reading input graph completed\n");
  std::printf("Time taken to read input graph: %.0f microseconds\n", us_since(t));
  if (!read_ints(argv[4], 2, c)) return 1;
  std::printf("Reading input changed edges data...\n");
  t = clk::now();
  if (!read_ints(argv[5], 4, ch)) return 1;
  long long total_ins = 0;
  for (auto& r : ch) total_ins += r[3] == 1;
  std::printf("Reading input changed edges data completed. totalInsertion:%lld\n", total_ins);
  std::printf("Time taken to read input changed edges: %.0f microseconds\n", us_since(t));

  std::vector<int32_t> u, v, colors((size_t)n, 0), bu, bv;
  std::vector<uint8_t> bi;
  for (auto& r : g) {
    u.push_back((int32_t)r[0]);
    v.push_back((int32_t)r[1]);
  }
  int32_t max_color = -1;
  for (auto& r : c) {
    if (r[0] < 0 || r[0] >= n || r[1] < 0) {
      std::fprintf(stderr, "bad color line %lld %lld\n", r[0], r[1]);
      return 1;
    }
    colors[(size_t)r[0]] = (int32_t)r[1];
    if (r[1] > max_color) max_color = (int32_t)r[1];
  }
  std::printf("Max color id in input graph: %d\n", max_color);
  for (auto& r : ch) {
    if (r[3] != 0 && r[3] != 1) continue;  // other status values are ignored, as in the reference
    bu.push_back((int32_t)r[0]);
    bv.push_back((int32_t)r[1]);
    bi.push_back((uint8_t)r[3]);
  }

  std::printf("Transferring graph data from CPU to GPU\n");
  t = clk::now();
  dc_options opt;
  dc_options_default(&opt);
  opt.driver = driver;
  dc_graph* h = nullptr;
  int rc = dc_graph_create(n, (int64_t)u.size(), u.data(), v.data(), &opt, &h);
  if (rc) return die(rc, "dc_graph_create");
  if ((rc = dc_colors_set(h, colors.data()))) return die(rc, "dc_colors_set");
  if ((rc = dc_apply_batch(h, (int64_t)bu.size(), bu.data(), bv.data(), bi.data())))
    return die(rc, "dc_apply_batch");
  const double transfer_ms = us_since(t) / 1e3;
  std::printf("**Time taken to transfer graph data from CPU to GPU: %.3f milliseconds**\n", transfer_ms);

  t = clk::now();
  dc_stats st;
  if ((rc = dc_recolor(h, rule, &st))) return die(rc, "dc_recolor");
  const double recolor_ms = us_since(t) / 1e3;
  print_trace(h, st);
  std::printf("**Time taken for processing affected neighbors: %.3f milliseconds**\n", recolor_ms);
  std::printf("****Total Time for Vertex Color Update: %.3f milliseconds****\n", recolor_ms);

  if ((rc = dc_colors_get(h, colors.data()))) return die(rc, "dc_colors_get");
  int64_t mono = 0;
  int32_t maxc = 0;
  if ((rc = dc_validate(h, &mono, &maxc))) return die(rc, "dc_validate");
  std::printf("highest color id used: %d\n", maxc);
  if (mono) std::printf("##Error found: %lld monochromatic edges\n", (long long)mono);
  if (out_path) {
    FILE* f = std::fopen(out_path, "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", out_path);
      return 1;
    }
    for (int32_t i = 0; i < n; ++i) std::fprintf(f, "%d %d\n", i, colors[(size_t)i]);
    std::fclose(f);
  }
  dc_graph_destroy(h);
  return mono ? 3 : 0;
}

// ---- one rank of the multi-GPU program ---------------------------------------
static const char* env_rank() {
  for (const char* k : {"OMPI_COMM_WORLD_RANK", "PMI_RANK", "PMIX_RANK", "SLURM_PROCID"}) {
    const char* v = std::getenv(k);
    if (v && *v) return v;
  }
  return nullptr;
}

static int multi_gpu_rank(const char* graph_path, const char* color_path, const char* change_path,
                          int32_t rank, const std::string& dir, int32_t driver, const char* out_path) {
  std::printf("rank %d: reading partition files\n", rank);
  std::vector<std::vector<long long>> part, all, g, c, ch;
  const std::string pfile = dir + "/partition" + std::to_string(rank) + ".txt";
  const std::string afile = dir + "/partitionAll.txt";
  if (!read_ints(pfile.c_str(), 1, part)) return 1;
  bool have_all = false;
  if (FILE* f = std::fopen(afile.c_str(), "r")) {
    std::fclose(f);
    if (!read_ints(afile.c_str(), 1, all)) return 1;
    have_all = true;
  }
  std::printf("Reading input graph...\n");
  auto t = clk::now();
  if (!read_ints(graph_path, 2, g)) return 1;
  std::printf("Reading input graph completed\n");
  std::printf("Time taken to read input graph: %.0f microseconds\n", us_since(t));
  std::printf("Reading input changed edges data...\n");
  t = clk::now();
  if (!read_ints(change_path, 4, ch)) return 1;
  std::printf("Reading input changed edges data completed.\n");
  std::printf("Time taken to read input changed edges: %.0f microseconds\n", us_since(t));
  if (!read_ints(color_path, 2, c)) return 1;

  // global vertex count: partitionAll.txt's length (one line per vertex),
  // else one past the highest ID any file names
  long long hi = -1;
  auto see = [&](long long x) { hi = x > hi ? x : hi; };
  for (auto& r : part) see(r[0]);
  for (auto& r : g) { see(r[0]); see(r[1]); }
  for (auto& r : ch) { see(r[0]); see(r[1]); }
  for (auto& r : c) see(r[0]);
  const long long n_global = have_all ? (long long)all.size() : hi + 1;
  if (n_global <= 0 || n_global > 0x7ffffffeLL || hi >= n_global) {
    std::fprintf(stderr, "vertex ID %lld out of range (%lld vertices)\n", hi, n_global);
    return 1;
  }
  std::vector<int32_t> owned, part_of, u, v, cu, cv, colors((size_t)n_global, -1);
  std::vector<uint8_t> ci;
  for (auto& r : part) owned.push_back((int32_t)r[0]);
  for (auto& r : all) part_of.push_back((int32_t)r[0]);
  for (auto& r : g) {
    u.push_back((int32_t)r[0]);
    v.push_back((int32_t)r[1]);
  }
  for (auto& r : c) {
    if (r[0] < 0 || r[1] < 0) {
      std::fprintf(stderr, "bad color line %lld %lld\n", r[0], r[1]);
      return 1;
    }
    colors[(size_t)r[0]] = (int32_t)r[1];
  }
  for (auto& r : ch) {
    if (r[3] != 0 && r[3] != 1) continue;  // other status values are ignored, as in the reference
    cu.push_back((int32_t)r[0]);
    cv.push_back((int32_t)r[1]);
    ci.push_back((uint8_t)r[3]);
  }
  dc_partition_plan* plan = nullptr;
  int rc = dc_partition_plan_create((int32_t)n_global, rank, (int32_t)owned.size(), owned.data(),
                                    have_all ? part_of.data() : nullptr, (int64_t)u.size(), u.data(),
                                    v.data(), (int64_t)cu.size(), cu.data(), cv.data(), ci.data(), &plan);
  if (rc) return die(rc, "dc_partition_plan_create");
  dc_partition_info inf;
  dc_partition_plan_info(plan, &inf);
  std::printf("rank %d: local vertices %d, other part vertices %d, border vertices %d\n", rank,
              inf.n_owned, inf.n_ghost, inf.n_border);
  std::printf("rank %d: local edges %lld, cross edges %lld\n", rank, (long long)inf.m_internal,
              (long long)inf.m_cross);
  std::printf("rank %d: Ins edges(local) %lld, Del edges(local) %lld, Ins edges(cross) %lld, "
              "Del edges(cross) %lld\n", rank, (long long)inf.k_ins, (long long)inf.k_del,
              (long long)inf.k_xins, (long long)inf.k_xdel);
  std::vector<int32_t> l2g((size_t)inf.n_owned + inf.n_ghost + 1), bu((size_t)inf.k_local + 1),
      bv((size_t)inf.k_local + 1);
  std::vector<uint8_t> bi((size_t)inf.k_local + 1);
  dc_partition_plan_arrays(plan, l2g.data(), nullptr, nullptr, bu.data(), bv.data(), bi.data());

  std::printf("Transferring graph data from CPU to GPU\n");
  t = clk::now();
  dc_options opt;
  dc_options_default(&opt);
  opt.driver = driver;
  dc_graph* h = nullptr;
  rc = dc_partition_create_global(plan, colors.data(), &opt, &h);
  dc_partition_plan_destroy(plan);
  if (rc) return die(rc, "dc_partition_create_global");
  if ((rc = dc_apply_batch(h, inf.k_local, bu.data(), bv.data(), bi.data())))
    return die(rc, "dc_apply_batch");
  std::printf("**Time taken to transfer graph data from CPU to GPU: %.3f milliseconds**\n",
              us_since(t) / 1e3);
  t = clk::now();
  dc_stats st;
  if ((rc = dc_recolor(h, DC_RULE_INDEPENDENT, &st))) return die(rc, "dc_recolor");
  const double recolor_ms = us_since(t) / 1e3;
  print_trace(h, st);
  std::printf("**Time taken for processing affected neighbors: %.3f milliseconds**\n", recolor_ms);
  std::printf("****Total Time for Vertex Color Update: %.3f milliseconds****\n", recolor_ms);
  std::vector<int32_t> local((size_t)inf.n_owned + inf.n_ghost + 1);
  if ((rc = dc_colors_get(h, local.data()))) return die(rc, "dc_colors_get");
  int64_t mono = 0;
  int32_t maxc = 0;
  if ((rc = dc_validate(h, &mono, &maxc))) return die(rc, "dc_validate");
  std::printf("rank %d: highest color id used: %d\n", rank, maxc);
  if (mono) std::printf("##Error found: %lld monochromatic edges\n", (long long)mono);
  if (out_path) {
    FILE* f = std::fopen(out_path, "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", out_path);
      return 1;
    }
    for (int32_t i = 0; i < inf.n_owned; ++i) std::fprintf(f, "%d %d\n", l2g[(size_t)i], local[(size_t)i]);
    std::fclose(f);
  }
  dc_graph_destroy(h);
  return mono ? 3 : 0;
}

int main(int argc, char** argv) {
  int npos = 0;
  for (int i = 1; i < argc && argv[i][0] != '-'; ++i) ++npos;
  if (npos == 5) return single_gpu(argc, argv);
  if (npos != 3) {
    std::fprintf(stderr,
                 "usage: %s <graph> <nodes> <edges> <colors> <changes> [--rule independent|ldf] "
                 "[--driver plain|fused] [--out file]\n"
                 "       %s <graph> <colors> <changes> [--rank r] [--dir d] [--driver plain|fused] "
                 "[--out file]   (one rank of a partitioned graph)\n",
                 argv[0], argv[0]);
    return 2;
  }
  const char* er = env_rank();
  int32_t rank = er ? (int32_t)std::atoi(er) : 0, driver = DC_DRIVER_PLAIN;
  std::string dir = ".";
  const char* out_path = nullptr;
  for (int i = 4; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--rank") && i + 1 < argc) {
      rank = (int32_t)std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "--dir") && i + 1 < argc) {
      dir = argv[++i];
    } else if (!std::strcmp(argv[i], "--driver") && i + 1 < argc) {
      driver = !std::strcmp(argv[++i], "fused") ? DC_DRIVER_FUSED : DC_DRIVER_PLAIN;
    } else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) {
      out_path = argv[++i];
    } else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    }
  }
  return multi_gpu_rank(argv[1], argv[2], argv[3], rank, dir, driver, out_path);
}
