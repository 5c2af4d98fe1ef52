/*
 * dyncolor.h -- C ABI of the MI355X (gfx950) dynamic-recoloring library.
 *
 * The reference (ArindamK92/DynColor) has no library API: each variant is one
 * CUDA executable whose main() reads a graph, a coloring and one change file,
 * then runs a fixed kernel sequence.  Each entry point below replaces one of
 * those host steps; the citation names the reference code it stands in for
 * (paths relative to the reference repo root):
 *   HI = DynColorHeuristicIndependent/DynColorHeuristicIndependent/
 *   HB = DynColorHeuristicBased/DynColorHeuristicBased/
 *   MG = Multi-GPU/
 *
 * Conventions: vertex IDs and colors are int32; every function returns DC_OK
 * (0) or a negative DC_ERR_* code and never aborts or throws; dc_last_error()
 * gives a message for the last failure on the calling thread.  A dc_graph is
 * used by one host thread at a time.  Calls return once their results are
 * valid on the device.
 */
#ifndef DYNCOLOR_H
#define DYNCOLOR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DC_OK 0
#define DC_ERR_INVALID (-1)   /* bad argument: null pointer, negative size, bad rule */
#define DC_ERR_VERTEX (-2)    /* vertex ID out of range, self loop, ghost-ghost edge */
#define DC_ERR_HIP (-3)       /* HIP runtime failure (no device, launch error)      */
#define DC_ERR_NOMEM (-4)     /* device or host allocation failed                  */
#define DC_ERR_ROUNDS (-5)    /* recolor loop hit max_rounds without converging     */
#define DC_ERR_STATE (-6)     /* call out of order (recolor with no staged batch)   */
#define DC_ERR_ADMIT (-7)     /* serving admission rejected (memory budget)         */

/* Rule families.  INDEPENDENT: heuristic-independent rules of HI/ (color total
 * order, vertex-ID tie-break).  LDF: heuristic-based rules of HB/ (largest
 * degree first, lower ID first on ties). */
#define DC_RULE_INDEPENDENT 0
#define DC_RULE_LDF 1

/* Recolor drivers.  PLAIN: one kernel launch per phase, host-driven round loop
 * (the drop-in for the reference's while(*change) loop).  FUSED: one persistent
 * cooperative launch runs every round of a batch with device-side barriers. */
#define DC_DRIVER_PLAIN 0
#define DC_DRIVER_FUSED 1
/* GRAPH: the plain driver's launches, each host block of loop rounds captured
 * once as a HIP graph and replayed (per graph handle, while the round range,
 * the grid and the loop's device state are the same); profiled batches use the
 * plain launches. */
#define DC_DRIVER_GRAPH 2

typedef struct dc_graph dc_graph;
typedef struct dc_batch dc_batch;

typedef struct {
  float slack_frac;    /* spare row capacity as a fraction of degree (default 0.25) */
  int32_t slack_min;   /* minimum spare slots per row (default 4)                     */
  int32_t max_rounds;  /* recolor loop cap (default 1000000)                          */
  int32_t driver;      /* DC_DRIVER_PLAIN (default) or DC_DRIVER_FUSED                */
  int32_t device;      /* HIP device ordinal (default 0)                              */
  int32_t spare_slots; /* spare tail for rows that outgrow their capacity; 0 = half
                          the row slots, at least 2^20 (default 0)                    */
  int32_t exact_rounds; /* 1: the independent-rule loop runs every round the reference
                          runs; 0 (default): it stops once the colours are settled
                          (a round without a change after which every owned vertex
                          holds the mex of its neighbours' colours: the reference's
                          remaining rounds change no colour).  Colours are the same
                          either way; only the trace and round counts differ.         */
  int32_t list_limit;  /* independent-rule loop rounds with at most this many vertices
                          in L take their sets from worklists instead of scanning the
                          per-vertex state arrays; 0 (default): automatic (at most
                          131072, and fewer on small graphs); -1: never; > 0 also lists
                          R in every round that is not a pull round.  Results do not
                          depend on it.                                             */
  int32_t pull_mode;   /* independent-rule loop, how a round's candidate counts are
                          brought up to date: 0 (default) pushes count changes from
                          the rows that changed state and recounts every count by a
                          pull over all rows in rounds whose changed rows hold more
                          than heavy_slots slots and whose candidate set more than
                          doubled; 1: every round pulls; -1: no round pulls.
                          Results do not depend on it.                               */
  int32_t group_deletions; /* batches with at least this many deletion lines tombstone
                          them grouped by scanned row (lines sorted by row; each row
                          read once for all its lines); 0 (default): 65536; -1:
                          never.  Results do not depend on it.                       */
  int64_t heavy_slots; /* the pull threshold in row slots; 0 (default): 1/8 of the
                          laid-out row slots, at least 65536                         */
} dc_options;

/* Per-batch counters.  Byte counts follow the algorithmic model: each visited
 * row costs its list entry, row bounds, every slot of the row and one color
 * gather per live entry (plus a priority gather for LDF). */
typedef struct {
  int64_t updates;          /* insert + delete lines in the batch             */
  int64_t insertions, deletions;
  int32_t phases;           /* entries in the phase trace                     */
  int32_t loop_rounds;      /* iterations of the recolor loop                 */
  int64_t evaluated;        /* vertex evaluations (targets + candidates)      */
  int64_t changed;          /* color changes                                  */
  int64_t frontier_sum;     /* sum of |L| over loop rounds                    */
  int64_t slots_scanned;    /* row slots read by find + recolor + targets     */
  int64_t live_scanned;     /* live entries among them                        */
  int64_t recolor_bytes;    /* algorithmic bytes of the recolor kernels       */
  double ms_apply;          /* device time of the batch apply (hipEvent)      */
  double ms_recolor;        /* device time of rounds + loop (hipEvent)        */
  double ms_recolor_kernel; /* summed duration of the recolor kernels (profiling on) */
  int32_t recolor_launches; /* recolor kernel launches in the batch           */
  int32_t regrows;          /* batches whose rows outgrew their capacity      */
  double ms_count_pass;     /* summed duration of the independent-rule loop's
                               count-pass launches (profiling on, plain driver);
                               for the ldf rule: of the loop's rounds (four
                               launches each)                                 */
  int64_t count_pass_bytes; /* algorithmic bytes of those launches (ldf rule:
                               the reference's bytes for those rounds)        */
  int32_t count_pass_launches;  /* launches (ldf rule: rounds)                */
  int32_t loop_settled;     /* 1 if the loop stopped at settled colours       */
  int32_t pull_rounds;      /* loop rounds whose counts were recounted by pull */
  int32_t heavy_push_rounds;  /* loop rounds that pushed count changes from more
                               than heavy_slots changed row slots             */
  int32_t heavy_push_listed;  /* ... and took their next candidates from lists */
  int32_t graph_captures;   /* round blocks captured as HIP graphs in this batch
                               (graph driver; the others were replays)         */
  /* Work terms of the reference's algorithm for this batch (profiling on;
   * zero otherwise): the reference rescans every row of L_r and C_r each loop
   * round (HI/gpuFunctions_undir.cuh:229-284) and both endpoint rows per
   * deletion (HI/gpuFunctions_undir.cuh:118-142). */
  int64_t lc_vertices;      /* sum over loop rounds of |L_r u C_r|            */
  int64_t lc_slots;         /* sum over loop rounds of the row slots of L_r u C_r */
  int64_t del_row_slots;    /* sum over deletion lines of both rows' slots    */
  double ms_round_a;        /* summed duration of the independent-rule loop's
                               phase-A launches (profiling on, plain driver);
                               with ms_count_pass: the loop rounds' kernel time */
  int64_t moved_slots;      /* ldf rule: row slots of the vertices whose colour
                               the loop changed                               */
  int32_t slots_rewritten;  /* ldf rule: 1 if the neighbours' slot colours and
                               histograms were rewritten for every row after the
                               loop (moved rows holding more than a sixth of
                               the laid-out slots), 0 if by the moved rows    */
  int32_t reserved1;
} dc_stats;

/* Work-split parameters of the independent-rule loop and the batch apply
 * (results never depend on them; the tuner in dyncolor_amd/tune.py picks them
 * per device and workload by timing).  No counterpart in the reference, whose
 * launch shapes are fixed (numberOfBlocks = 32 x SMs, HI/Color_main.cu:54). */
typedef struct {
  int32_t loop_blocks;   /* workgroups of a loop round (1..4096, default 2048)           */
  int32_t light_blocks;  /* workgroups of a light round (0 = two per CU)                  */
  int32_t light_c;       /* a round after one with fewer members is light (default 4096) */
  int32_t list_c;        /* rounds with at most this many vertices in L take their sets
                            from worklists (default 131072; also capped at 1/64 of the
                            words a scan passes)                                          */
  int64_t list_slots;    /* ... and list R when D's rows hold at most this many slots
                            (default 2^22)                                                */
  int32_t heavy_div;     /* pull threshold = laid-out row slots / heavy_div, at least 65536
                            (default 8; dc_options.heavy_slots overrides)                 */
  int32_t group_deletions; /* deletion lines from which the batch apply groups them by
                            scanned row (0: dc_options.group_deletions decides)          */
  int32_t ldf_loop;      /* heuristic-based (ldf) loop: 0 = its sets kept as counts that
                            move with the rows entering or leaving L and the changed
                            colours (default), 1 = every round re-derives C from the rows
                            of L (the reference's form); same colours and traces          */
  int32_t reserved0;
} dc_tuning;

void dc_tuning_default(dc_tuning* t);
/* Set / read a graph's tuning (validated: DC_ERR_INVALID when out of range). */
int dc_set_tuning(dc_graph* g, const dc_tuning* t);
int dc_get_tuning(const dc_graph* g, dc_tuning* t);

/* ---- graph lifetime ------------------------------------------------------ */
void dc_options_default(dc_options* opt);

/* Build a device-resident undirected graph (both directions stored, spare
 * capacity per row for insertions) from m edges (u[i], v[i]).  Replaces
 * read_graphEdges (HI/all_structure_undir.cuh:156-181; the degree counter of
 * HB/all_structure_undir.cuh:158-186) and the CSR flatten + upload of
 * transfer_data_to_GPU (HI/supportingFunctions.cu:19-67).  opt may be NULL. */
int dc_graph_create(int32_t n, int64_t m, const int32_t* u, const int32_t* v,
                    const dc_options* opt, dc_graph** out);

/* One partition's share: vertices [0, n_owned) are owned, [n_owned,
 * n_owned + n_ghost) are ghost copies of other partitions' vertices whose colors
 * (ghost_colors, length n_ghost) stay frozen.  Edges are in these local IDs and
 * must have at least one owned endpoint.  Replaces the partition ingest of
 * MG/all_structure_undir.cuh:229-312 and MG/Supporting.h:68-211. */
int dc_partition_create(int32_t n_owned, int32_t n_ghost, int64_t m,
                        const int32_t* u, const int32_t* v,
                        const int32_t* ghost_colors, const dc_options* opt,
                        dc_graph** out);

/* ---- partition ingest in global IDs ---------------------------------------
 * The multi-GPU variant's inputs, for one rank: owned_ids = the rank's
 * partition file (one global ID per line; line i becomes local ID i,
 * read_Partition_Vertices, MG/Supporting.h:29-46), part_of = the part of
 * every global vertex (partitionAll.txt, read_PartitionID_AllVertices,
 * MG/Supporting.h:49-63; may be NULL), the whole graph's edges and the whole
 * change file in global IDs.  The plan keeps what the rank holds, as
 * read_graphEdges and readin_changes do (MG/all_structure_undir.cuh:229-312,
 * 78-192): edges with an owned endpoint (internal ones, then cross ones as
 * (owned, other)), the other parts' vertices they name as ghosts with local
 * IDs n_owned, n_owned + 1, ... in order of first appearance (graph lines,
 * then change lines), and the change lines with an owned endpoint in file
 * order (cross lines as (owned, ghost)); the rest is dropped.  Host only: no
 * device is touched until dc_partition_create_global. */
typedef struct dc_partition_plan dc_partition_plan;

typedef struct {
  int32_t n_global, rank;
  int32_t n_owned, n_ghost;
  int32_t n_border;     /* owned vertices with a ghost neighbour (edges or inserted lines) */
  int32_t n_parts;      /* 1 + the highest part in part_of (0 without it)           */
  int64_t m_local;      /* edges kept = m_internal + m_cross                         */
  int64_t m_internal, m_cross, m_dropped;
  int64_t k_local;      /* change lines kept = k_del + k_ins + k_xdel + k_xins       */
  int64_t k_del, k_ins, k_xdel, k_xins, k_dropped;
} dc_partition_info;

int dc_partition_plan_create(int32_t n_global, int32_t rank, int32_t n_owned,
                             const int32_t* owned_ids, const int32_t* part_of, int64_t m,
                             const int32_t* u, const int32_t* v, int64_t k, const int32_t* cu,
                             const int32_t* cv, const uint8_t* cins, dc_partition_plan** out);
int dc_partition_plan_info(const dc_partition_plan* p, dc_partition_info* info);
/* Copy out (any pointer may be NULL): local_to_global[n_owned + n_ghost],
 * local edges u, v [m_local], local change lines bu, bv, bins [k_local]. */
int dc_partition_plan_arrays(const dc_partition_plan* p, int32_t* local_to_global, int32_t* u,
                             int32_t* v, int32_t* bu, int32_t* bv, uint8_t* bins);
/* Border vertices per part: per_part[q] = owned vertices with a neighbour in
 * part q (the vertices whose colours the reference's MPI exchange sends to
 * rank q, bvPerPart of MG/kernel.cu:127,351-354).  Needs part_of. */
int dc_partition_plan_border(const dc_partition_plan* p, int64_t* per_part, int32_t nparts);
/* The rank's device graph from the plan: dc_partition_create with the plan's
 * local edges, owned and ghost colours looked up in colors_global (one per
 * global vertex; a negative entry for a vertex the rank holds is an error, as
 * read_Input_Color, MG/all_structure_undir.cuh:199-222, would leave it
 * unset).  Apply the plan's change lines with dc_apply_batch. */
int dc_partition_create_global(const dc_partition_plan* p, const int32_t* colors_global,
                               const dc_options* opt, dc_graph** out);
int dc_partition_plan_destroy(dc_partition_plan* p);

int dc_graph_destroy(dc_graph* g);

int dc_graph_info(const dc_graph* g, int32_t* n, int32_t* n_owned,
                  int64_t* live_entries, int64_t* capacity);

/* ---- colors --------------------------------------------------------------- */
/* Upload a coloring for all n vertices (ghost entries included).  Replaces
 * read_Input_Color (HI/all_structure_undir.cuh:130-149). */
int dc_colors_set(dc_graph* g, const int32_t* colors);
int dc_colors_get(dc_graph* g, int32_t* colors);

/* Deterministic first-fit greedy coloring on the GPU: vertex-ID order for
 * INDEPENDENT, priority order for LDF; identical to the sequential greedy.
 * Stands in for the external static coloring the reference reads from a file
 * (MG/README.txt:21-23). Ghost colors are left untouched. */
int dc_static_color(dc_graph* g, int32_t rule);

/* ---- update batches ------------------------------------------------------- */
/* Apply one batch of k lines (is_insert[i] = 1 insert, 0 delete; order as in a
 * change file) to the device graph and stage it for dc_recolor.  Replaces
 * readin_changes (HI/all_structure_undir.cuh:78-123, HB/all_structure_undir.cuh:
 * 78-129, MG/all_structure_undir.cuh:78-192) and the tombstone launch
 * deleteEdgeFromAdj (HI/supportingFunctions.cu:69-94).  Zero-insert and
 * zero-delete batches are legal. */
int dc_apply_batch(dc_graph* g, int64_t k, const int32_t* u, const int32_t* v,
                   const uint8_t* is_insert);

/* Device-resident batch: validated and copied to HBM once, applied later
 * without a host copy (the form the throughput benchmarks use). */
int dc_batch_create(dc_graph* g, int64_t k, const int32_t* u, const int32_t* v,
                    const uint8_t* is_insert, dc_batch** out);
int dc_batch_destroy(dc_batch* b);
int dc_apply_device_batch(dc_graph* g, const dc_batch* b);

/* Run the staged batch's rounds and the recolor loop.  Replaces the kernel
 * sequence of HI/Color_main.cu:146-202 (INDEPENDENT), HB/Color_main.cu:160-224
 * (LDF) and MG/kernel.cu:296-523 (partition share, INDEPENDENT only).
 * st may be NULL. */
int dc_recolor(dc_graph* g, int32_t rule, dc_stats* st);

/* apply + recolor in one call */
int dc_process_batch(dc_graph* g, int32_t rule, int64_t k, const int32_t* u,
                     const int32_t* v, const uint8_t* is_insert, dc_stats* st);

/* ---- checks ---------------------------------------------------------------- */
/* Count monochromatic live edges (each undirected edge once) and the highest
 * owned color.  Replaces the single-thread validate kernel
 * (HI/gpuFunctions_undir.cuh:289-305) and the host max-color scan
 * (HI/Color_main.cu:218-226). */
int dc_validate(dc_graph* g, int64_t* monochromatic, int32_t* max_color);

/* Phase trace of the last dc_recolor: 4 int64 per phase (kind, a, b, c).
 * kinds: 1 deletion round (targets, changed), 2 insertion round (targets,
 * changed), 3 LDF change-edge round (endpoints, changed), 4 cross-deletion
 * round, 5 cross-insertion round, 10 loop round (|L|, |C|, changed), 11 the
 * loop stopped at settled colours (rounds run, 0, 0): the reference runs
 * more rounds, none of which changes a color, 12 parts (dc_set_parts) found
 * settled after that many rounds left the loop (rounds run, parts, 0): the
 * later loop rounds count only the other parts' vertices. */
int dc_trace_get(dc_graph* g, int64_t* trace, int32_t cap, int32_t* n_phases);

/* Live neighbours of v (multi-edges repeated), for tests and tools. */
int dc_neighbors(dc_graph* g, int32_t v, int32_t* out, int64_t cap, int64_t* deg);

/* Current LDF priority counter (one +-1 per change line, like
 * HB/all_structure_undir.cuh:111-119). */
int dc_prop_get(dc_graph* g, int32_t* prop);

/* Compact tombstoned row slots (keeps spare capacity). */
int dc_compact(dc_graph* g);

/* dc_compact without any device allocation: the slot map of the compaction
 * (4 bytes per row slot, which dc_compact allocates on every call) lives in a
 * static device array of 2^23 entries held by each GPU the library runs on, so
 * graphs of up to 2^23 - 1 row slots (dc_graph_info's capacity, spare tail
 * included) are accepted and larger ones rejected with DC_ERR_INVALID.  Calls
 * that overlap on one GPU share the array and race; the caller serialises
 * them.  No counterpart in the reference, which rebuilds its CSR from text on
 * every run (HB/supportingFunctions.cu:27-45): it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid allocation). */
int dc_compact_noalloc(dc_graph* g);

/* Count row slots whose cached neighbour colour or reverse-slot link is
 * stale (0 when the device layout is consistent); for tests.  No counterpart
 * in the reference, whose rows hold only {neighbour, flag}
 * (HI/all_structure_undir.cuh:26-29). */
int dc_check_consistency(dc_graph* g, int64_t* bad_slots);

/* Enable per-kernel hipEvent timing of the recolor kernels (fills
 * dc_stats.ms_recolor_kernel). */
int dc_set_profiling(dc_graph* g, int32_t on);

/* Select the recolor driver (DC_DRIVER_PLAIN / DC_DRIVER_FUSED) and, for the
 * fused driver, the resident workgroups per compute unit (0 = keep).  The
 * reference has one fixed driver (its while(*change) loop,
 * HI/Color_main.cu:181-202); the fused driver runs the same rounds inside one
 * cooperative launch. */
int dc_set_driver(dc_graph* g, int32_t driver, int32_t blocks_per_cu);

/* Declare the graph a union of disjoint parts (an arena of independent
 * graphs): part p holds the vertices offsets[p] .. offsets[p+1]-1, with
 * offsets[0] = 0, offsets[nparts] = the vertex count, and no edge between
 * two parts (checked: DC_ERR_INVALID).  The independent-rule loop (plain
 * driver, exact_rounds = 0) then also checks each part on its own: a part
 * whose colours are settled leaves the loop while the others go on (trace
 * kind 12); colours are the same either way.  A later batch inserting an edge
 * between two parts ends this.  nparts = 0 clears it.  No counterpart in the
 * reference, which colours one graph per process (HI/Color_main.cu:31-247);
 * it serves the batched and served workloads (dc_server_create sets it). */
int dc_set_parts(dc_graph* g, int32_t nparts, const int32_t* offsets);

int dc_synchronize(dc_graph* g);
const char* dc_last_error(void);
int dc_device_count(int32_t* count);
/* Name (NUL-terminated, at most cap bytes), architecture and compute units of
 * a HIP device, for keying tuned parameters per device. */
int dc_device_info(int32_t device, char* name, int32_t cap, char* arch, int32_t arch_cap,
                   int32_t* compute_units);

/* ---- serving: continuous batching over resident graphs --------------------
 * No counterpart in the reference (it recolors one batch per process run,
 * HI/Color_main.cu:31-247).  The resident graphs live in nlanes lanes: lane l
 * is the arena dc_graph arenas[l] (the disjoint union of its graphs, each
 * lane with its own stream and host worker thread); graph g owns arena IDs
 * [first[g], first[g] + size[g]) of lane lane_of[g].  The server uses each
 * arena from its worker only while dc_server_run runs.  Requests are replayed
 * against the wall clock: request i targets graph gid[i], carries lines
 * [roff[i], roff[i+1]) of (u, v, is_insert) in that graph's local IDs and
 * arrives arrival[i] seconds after the start (non-decreasing).  Requests are
 * admitted in arrival order while the device bytes in flight
 * (dc_server_request_bytes of each admitted, unfinished request) fit
 * budget_bytes, also while steps run (a request that does not fit waits for
 * a step to finish; one larger than the whole budget is rejected, latency
 * -1).  Each lane repeatedly takes all its admitted requests, coalesces them
 * per graph into one batch and recolors its arena in one step; lanes run
 * concurrently on the device.  latency[i] = completion - arrival (seconds);
 * step_of_req[i] = the step (numbered in completion order over all lanes)
 * that served request i, or NULL.  budget_bytes = 0: a quarter of the free
 * device memory at creation. */
typedef struct dc_server dc_server;

typedef struct {
  int32_t steps;
  int32_t lanes;
  int64_t served, rejected, lines, max_step_lines;
  double elapsed_s;
  double busy_s;              /* summed over lanes: wall time inside steps */
  int64_t budget_bytes;
  int64_t max_inflight_bytes; /* most device bytes admitted at once         */
  int64_t admit_waits;        /* requests that waited for bytes to free up  */
  int64_t reserved[1];
} dc_serve_stats;

int dc_server_create(int32_t nlanes, dc_graph* const* arenas, int32_t ngraphs,
                     const int32_t* lane_of, const int32_t* first, const int32_t* size,
                     int64_t budget_bytes, int32_t rule, dc_server** out);
int dc_server_destroy(dc_server* s);
int dc_server_run(dc_server* s, int64_t nreq, const int32_t* gid, const int64_t* roff,
                  const int32_t* u, const int32_t* v, const uint8_t* is_insert,
                  const double* arrival, double* latency, int32_t* step_of_req,
                  dc_serve_stats* st);
/* Batched step over lanes: lane l applies the device batch batches[l] to the
 * arena arenas[l] (distinct handles) and recolors it with `rule`; the lanes
 * run concurrently (one host thread and the arena's own stream each) and the
 * call returns when every lane's colors are final.  stats: nlanes entries (or
 * NULL).  On a failure the first failing lane's code is returned and
 * dc_last_error names the lane.  No counterpart in the reference (one graph
 * per process, HI/Color_main.cu:31-247); it serves the batched workload:
 * graphs split over a few arenas overlap on the device where one arena's loop
 * would wait for its slowest graph. */
int dc_lanes_step(int32_t nlanes, dc_graph* const* arenas, const dc_batch* const* batches,
                  int32_t rule, dc_stats* stats);

/* The device bytes a request of `lines` lines, `insertions` of them
 * insertions, counts against the budget (staged lines, change-list entries,
 * row slots its insertions may add); -1 for bad sizes. */
int64_t dc_server_request_bytes(int64_t lines, int64_t insertions);

/* ---- synthetic inputs (host, deterministic, splitmix64-seeded) ------------ */
/* Erdos-Renyi G(n, m): m distinct undirected edges, no self loops. */
int dc_gen_er(int32_t n, int64_t m, uint64_t seed, int32_t* u, int32_t* v);

/* R-MAT: 2^scale vertices, edge_factor * 2^scale samples with quadrant
 * probabilities a, b, c (d = 1-a-b-c); self loops and duplicates removed.
 * u and v need room for edge_factor * 2^scale entries (every sample); the
 * unique edge count is returned through *m_out. */
int dc_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                uint64_t seed, int32_t* u, int32_t* v, int64_t* m_out);

/* Update-batch stream over a live edge set: nb batches of bs lines each;
 * round(bs * ins_frac) insertions of uniform non-edges (distinct, no self
 * loops) and the rest deletions of distinct existing edges, each batch drawn
 * from the graph as left by the previous ones.  Outputs nb*bs lines. */
int dc_gen_batches(int32_t n, int64_t m, const int32_t* u, const int32_t* v,
                   int32_t nb, int64_t bs, double ins_frac, uint64_t seed,
                   int32_t* bu, int32_t* bv, uint8_t* bins);

#ifdef __cplusplus
}
#endif
#endif
