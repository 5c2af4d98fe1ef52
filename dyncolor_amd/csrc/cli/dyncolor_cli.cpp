// dyncolor_cli -- command-line drop-in for the reference program.
//
//   dyncolor_cli <graph file> <nodes> <edges> <color file> <change file>
//                [--rule independent|ldf] [--driver plain|fused] [--out colors.txt]
//
// Same positional arguments and text formats as the reference
// (HI/Color_main.cu:21-38, HB/Color_main.cu:21-38): graph lines "a b w"
// (read_graphEdges, HI/all_structure_undir.cuh:156-181), color lines
// "node color" (read_Input_Color, :130-149), change lines "a b w inst" with
// inst 1 = insert, 0 = delete (readin_changes, :78-123).  Edge weights are
// read and ignored, as in the reference.  Prints the reference's phase timing
// lines; --out writes the final coloring as "node color" lines (the reference
// computes it but never writes it, HI/Color_main.cu:211-216).
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

int main(int argc, char** argv) {
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
