/*
 * oracle.h -- CPU restatement of SYgraph's traversal algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (sygraph_amd/, include/)
 * includes, links or calls this code.  It is the checker that the HIP path is
 * compared against, and (cpu_baseline.c) the CPU baseline that bench.py times.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Graphs are CSR with int64 row offsets and int32 column ids (ascending within
 * a row), the layout the HIP library uses for every workload.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based PRNG (SplitMix64 finaliser stepped by the golden gamma) -- */
uint64_t or_mix64(uint64_t z);
uint64_t or_rand64(uint64_t seed, uint64_t ctr);

/* ---- synthetic graph generators (undirected edge tuples) ------------------- */
/* RMAT: nedges tuples over 2^scale vertices, quadrant probabilities a,b,c
 * (d = 1-a-b-c), optional seeded bijective relabel of the ids. */
void or_rmat_pairs(int scale, int64_t nedges, double a, double b, double c, uint64_t seed, int relabel, uint32_t* u,
                   uint32_t* v);
uint32_t or_rmat_relabel(uint32_t x, int scale, uint64_t seed);
/* rows x cols 4-neighbour grid; each undirected edge deleted with prob p.
 * Arrays must hold 2*rows*cols entries; returns the number kept. */
int64_t or_grid_pairs(int rows, int cols, double p_delete, uint64_t seed, uint32_t* u, uint32_t* v);
/* Erdos-Renyi G(n, m) with uniform endpoints. */
void or_er_pairs(int64_t n, int64_t m, uint64_t seed, uint32_t* u, uint32_t* v);

/* Symmetrise, drop self loops, sort and deduplicate.  offsets has n+1 entries,
 * cols must hold 2*npairs entries.  Returns the number of arcs m. */
int64_t or_build_csr(int64_t n, int64_t npairs, const uint32_t* u, const uint32_t* v, int64_t* offsets, int32_t* cols);
/* Weight of arc (u,v): one uint32 in [1,255] per undirected edge, keyed on the unordered pair. */
uint32_t or_edge_weight(uint32_t u, uint32_t v, uint64_t seed);
void or_fill_weights(int64_t n, const int64_t* offsets, const int32_t* cols, uint64_t seed, uint32_t* w);
/* Order-sensitive, parallel-reducible digest: sum_i mix64(i*K ^ x_i). */
uint64_t or_digest_i64(const int64_t* x, int64_t n);
uint64_t or_digest_i32(const int32_t* x, int64_t n);

/* ---- algorithms ----------------------------------------------------------- */
/* Level-synchronous BFS, unreached = n+1 (examples/bfs/bfs.cpp:17-56). Returns #levels
 * (number of non-empty frontiers, i.e. BFSRunDetails::iterations). */
int32_t or_bfs(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t src, int32_t* depth);
/* Dijkstra, unreached = n+1 (examples/sssp/sssp.cpp:24-66). */
void or_sssp(int64_t n, const int64_t* offsets, const int32_t* cols, const uint32_t* w, int32_t src, uint32_t* dist);
/* Connected components, label = max vertex id in the component (fixpoint of algorithms/cc.hpp:103-169). */
void or_cc(int64_t n, const int64_t* offsets, const int32_t* cols, uint32_t* labels);
/* Triangle count on the (degree,id)-oriented DAG (true count; fixes algorithms/tc.hpp:104). */
uint64_t or_tc(int64_t n, const int64_t* offsets, const int32_t* cols);
/* Brandes betweenness, FP64 with exact power-of-two per-level sigma scaling,
 * summed over nsrc sources, source excluded, no /2 (algorithms/bc.hpp:118-200). */
void or_bc(int64_t n, const int64_t* offsets, const int32_t* cols, const int64_t* in_offsets, const int32_t* in_cols,
           int32_t nsrc, const int32_t* srcs, double* bc);

/* Pull-level scan cost for the byte model: for every vertex u with depth[u] > level
 * (unvisited at that level) and degree > 0, k_u = 1 + position of the first neighbour
 * with depth == level, or deg(u) if none.  Returns sum k_u; *n_scanned = #such u. */
int64_t or_pull_scan_count(int64_t n, const int64_t* offsets, const int32_t* cols, const int32_t* depth, int32_t level,
                           const int32_t* hint, int64_t* n_scanned, int64_t* n_missed);
void or_pull_hint(int64_t n, const int64_t* in_offsets, const int32_t* in_cols, const int64_t* out_offsets,
                  int32_t* hint);

/* ---- CPU baseline: OpenMP restatement of the reference's plain method ------- */
/* 2-level bitmap frontier, level-synchronous push (mode 0), pull (1) or the
 * reference hybrid heuristic (2) with alpha/beta (algorithms/bfs.hpp:153-238,
 * 284-350).  Returns #levels. nthreads <= 0 uses the OpenMP default.  dirs
 * (may be NULL) receives the direction of each step (0 push, 1 pull). */
int32_t or_bfs_plain_omp(int64_t n, const int64_t* offsets, const int32_t* cols, int32_t src, int mode, float alpha,
                         float beta, int nthreads, int32_t* depth, int32_t* dirs, int32_t dirs_cap);
/* Frontier Bellman-Ford with atomic min (algorithms/sssp.hpp:152-210), OpenMP. */
int32_t or_sssp_plain_omp(int64_t n, const int64_t* offsets, const int32_t* cols, const uint32_t* w, int32_t src,
                          int nthreads, uint32_t* dist);
int or_omp_max_threads(void);

#ifdef __cplusplus
}
#endif
