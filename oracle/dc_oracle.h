/*
 * dc_oracle.h -- CPU oracle for the DynColor recolor path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded restatement of
 * the reference's recoloring rules, used by tests/ as the checker and by
 * bench.py as the timed CPU baseline.  Nothing in the product library links,
 * loads or calls it.
 *
 * Semantics: every reference kernel launch is one *round*; all reads in a round
 * see the colors/prev values of the round start ("snapshot"), all writes commit
 * at round end.  Each write to a vertex t is a function of (t, snapshot) only,
 * so the result is independent of evaluation order.
 *
 * Rules restated (reference paths relative to the reference repo root):
 *   independent (rule 0): DynColorHeuristicIndependent/DynColorHeuristicIndependent/
 *       gpuFunctions_undir.cuh:147-284, Color_main.cu:146-202,
 *       all_structure_undir.cuh:78-123, supportingFunctions.cu:80-93
 *   ldf         (rule 1): DynColorHeuristicBased/DynColorHeuristicBased/
 *       gpuFunctions_undir.cuh:104-129,222-314, Color_main.cu:160-224,
 *       all_structure_undir.cuh:78-129,158-186
 *   partition share (n_owned < n, rule 0): Multi-GPU/gpuFunctions_undir.cuh:118-283,
 *       Multi-GPU/kernel.cu:296-523, Multi-GPU/all_structure_undir.cuh:78-192
 *
 * Parity pins (tests/test_oracle.py): mex against brute force over all subsets of
 * {0..11}; CSR apply against a set model; ldf fixed point against textbook
 * greedy in priority order; both rules against an independently written Python
 * simulator on tiny graphs (all batches of size <= 2 enumerated); validity
 * (zero monochromatic live edges) on every result; hand-derived expected
 * colors for the reference's sample graph (tests/golden/).
 */
#ifndef DC_ORACLE_H
#define DC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_RULE_INDEPENDENT = 0, ORC_RULE_LDF = 1 };

/* phase kinds in the trace (4 int64 per phase: kind, a, b, c) */
enum {
  ORC_PH_DEL = 1,       /* a = distinct targets, b = changed            */
  ORC_PH_INS = 2,       /* a = distinct targets, b = changed            */
  ORC_PH_CHANGE = 3,    /* ldf change-edge round: a = endpoints, b = changed */
  ORC_PH_XDEL = 4,      /* partition cross deletions                    */
  ORC_PH_XINS = 5,      /* partition cross insertions                   */
  ORC_PH_LOOP = 10      /* a = |L|, b = |C|, c = changed               */
};

enum {
  ORC_OK = 0,
  ORC_EINVAL = -1,
  ORC_ECAP = -2,        /* round cap hit */
  ORC_ENOMEM = -3
};

typedef struct orc_graph orc_graph;

/* Vertices [0, n_owned) are owned; [n_owned, n) are frozen ghosts (their
 * colors never change and their rows are empty).  n_owned == n for a plain
 * graph.  Edges between two ghosts are rejected. */
orc_graph* orc_graph_create(int32_t n, int32_t n_owned, int64_t m,
                            const int32_t* u, const int32_t* v);
void orc_graph_destroy(orc_graph* g);

/* Stage one update batch (lines in order; ins[i]=1 insert, 0 delete), applying
 * it to the adjacency exactly as the reference's reader + tombstone kernel
 * do.  `color` is the batch-start coloring (used for the insertion filter). */
int orc_apply_batch(orc_graph* g, const int32_t* color, int64_t k,
                    const int32_t* u, const int32_t* v, const uint8_t* ins);

/* Run the staged batch's rounds and the recolor loop.  `color` and `prev` are
 * length n and updated in place.  `trace` (may be NULL) receives up to
 * trace_cap phases of 4 int64 each; the number of phases is returned through
 * *n_phases.  Returns ORC_OK, or ORC_ECAP if max_loop loop rounds were not
 * enough. */
int orc_recolor(orc_graph* g, int rule, int32_t* color, int32_t* prev,
                int32_t max_loop, int64_t* trace, int32_t trace_cap,
                int32_t* n_phases);

/* orc_recolor for the independent rule with each round's work split over
 * `threads` OpenMP threads (the all-cores CPU baseline): the same colours,
 * prev values and trace as orc_recolor. */
int orc_recolor_par(orc_graph* g, int rule, int32_t* color, int32_t* prev,
                    int32_t max_loop, int64_t* trace, int32_t trace_cap,
                    int32_t* n_phases, int32_t threads);

/* Textbook first-fit greedy: independent -> vertex-ID order, ldf -> priority
 * order (prop desc, ID asc).  Ghost colors (frozen) are taken from `color`. */
void orc_static_greedy(orc_graph* g, int rule, int32_t* color);

/* number of monochromatic live (v,w) entries, each undirected edge once */
int64_t orc_validate(const orc_graph* g, const int32_t* color, int32_t* max_color);

int32_t orc_mex(const int32_t* c, int64_t k);
int64_t orc_live_degree(const orc_graph* g, int32_t v);
/* live neighbours of v (with multiplicity) into out (cap entries); returns count */
int64_t orc_neighbors(const orc_graph* g, int32_t v, int32_t* out, int64_t cap);
void orc_get_prop(const orc_graph* g, int32_t* prop);

#ifdef __cplusplus
}
#endif
#endif
