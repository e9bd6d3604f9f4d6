/*
 * sygraph_hip.h -- C ABI of libsygraph_hip, the MI355X (gfx950) runtime and
 * traversal kernels behind the SYgraph operator API.
 *
 * The reference is a header-only C++20/SYCL library with no ABI of its own;
 * every entry point below replaces a host-side operation of that library that
 * a C / ctypes / FFI caller would otherwise have to reach through templates.
 * Each declaration cites the reference interface it stands in for
 * (paths relative to the reference repository root).
 *
 * Conventions
 *  - Every function returns an int status: SG_OK (0) or a negative SG_ERR_*.
 *    sg_last_error() returns a thread-local message for the last failure.
 *    No exception crosses this boundary (reference: std::runtime_error throws,
 *    e.g. operators/advance/advance.hpp:63, algorithms/bfs.hpp:155).
 *  - Vertex ids are 32-bit; edge offsets are 32- or 64-bit (chosen per graph);
 *    edge weights are 4-byte (uint32 or float32).
 *  - All work of one device handle is ordered on that handle's HIP stream.
 *    Calls return after completion unless their name ends in _async.
 *  - Pointers documented as "device" must come from sg_alloc (or another HIP
 *    allocation on the same device); "host" pointers are ordinary memory.
 */
#ifndef SYGRAPH_HIP_H
#define SYGRAPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SG_ABI_VERSION 1

enum sg_status {
  SG_OK = 0,
  SG_ERR_INVALID = -1,     /* bad argument */
  SG_ERR_HIP = -2,         /* HIP runtime error */
  SG_ERR_NOMEM = -3,       /* allocation failed */
  SG_ERR_UNSUPPORTED = -4, /* valid but not implemented combination */
  SG_ERR_NODEVICE = -5,    /* no HIP device present */
  SG_ERR_STATE = -6        /* object not initialised / wrong state */
};

typedef struct sg_device_s* sg_device_t;
typedef struct sg_graph_s* sg_graph_t;
typedef struct sg_frontier_s* sg_frontier_t;
typedef struct sg_event_s* sg_event_t;
typedef struct sg_policy_s* sg_policy_t;

/* ------------------------------------------------------------------ errors */
/* Thread-local text of the last error (reference: the what() of the
 * std::runtime_error its host code throws). Never NULL. */
const char* sg_last_error(void);
int sg_abi_version(void);

/* ------------------------------------------------------------------ device */
typedef struct sg_device_info {
  char name[256];          /* e.g. "AMD Instinct MI355X" */
  char arch[64];           /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int32_t compute_units;   /* utils/device.hpp:31-34 getNumComputeUnits */
  int32_t wave_size;       /* utils/device.hpp:21-24 getSubgroupSize */
  int32_t max_workgroup;   /* utils/device.hpp:26-29 getWorkgroupSize */
  int32_t max_waves_per_wg;/* utils/device.hpp:16-19 getMaxNumSubgroups */
  int64_t lds_bytes;       /* LDS per workgroup */
  int64_t global_mem_bytes;
  int64_t l2_bytes;
  int32_t clock_khz;
  int32_t memory_clock_khz;
  int32_t memory_bus_bits;
  int32_t device_ordinal;
} sg_device_info;

/* Number of visible HIP devices (0 when none; never an error). */
int sg_device_count(int* count);
/* Open device `ordinal` with its own stream. Replaces constructing a
 * sycl::queue{sycl::gpu_selector_v} (examples/bfs/bfs.cpp:120-124,
 * tests/include/test_utils.hpp:49-61). */
int sg_device_open(int ordinal, sg_device_t* out);
int sg_device_close(sg_device_t dev);
/* Device queries (utils/device.hpp:16-34). */
int sg_device_info_get(sg_device_t dev, sg_device_info* info);
/* The handle's hipStream_t, for header templates that launch functor kernels
 * on the same stream (the queue a GraphCSR carries, graph_csr.hpp:391). */
int sg_device_stream(sg_device_t dev, void** stream);
/* queue.wait() (used throughout, e.g. mlb_frontier.hpp:170). */
int sg_device_synchronize(sg_device_t dev);
/* Bound on a grid barrier's wait inside the device-resident level loops
 * (default 30 s); a launch that waits longer ends with an error.  No
 * counterpart in the reference (its level loops run on the host). */
int sg_device_set_wait_limit(sg_device_t dev, double seconds);

/* ------------------------------------------------------------------ memory */
enum sg_space { SG_SPACE_HOST = 0, SG_SPACE_DEVICE = 1, SG_SPACE_SHARED = 2 };
/* utils/memory.hpp:12-39 memoryAlloc<T, space> (host = pinned, device,
 * shared = managed). */
int sg_alloc(sg_device_t dev, size_t bytes, int space, void** out);
/* utils/memory.hpp:41-46 releaseUSM. NULL is a no-op. */
int sg_free(sg_device_t dev, void* ptr);
/* queue.copy (e.g. graph_csr.hpp:401-403); direction inferred by HIP. */
int sg_memcpy(sg_device_t dev, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte accesses); dst,
 * src and bytes must be multiples of 16.  No counterpart in the reference:
 * it is the copy kernel against which HBM bandwidth and the FETCH_SIZE /
 * WRITE_SIZE counters are calibrated. */
int sg_copy_device(sg_device_t dev, void* dst, const void* src, size_t bytes);
/* Random gather of count 4-byte words, dst[i] = src[h(i, seed) mod src_words]
 * (h: splitmix64).  No counterpart in the reference: the random-probe
 * reference point for the same counters, where a 4-byte read brings in a
 * whole line. */
int sg_gather_device(sg_device_t dev, uint32_t* dst, const uint32_t* src, uint64_t src_words, size_t count,
                     uint64_t seed);
/* queue.prefetch of shared (managed) memory to the device, used by
 * memoryAlloc<T, space::shared> under ENABLE_PREFETCH (utils/memory.hpp:24-26).
 * A hint: memory the runtime cannot migrate is left where it is (SG_OK). */
int sg_prefetch(sg_device_t dev, const void* ptr, size_t bytes);
/* queue.fill for 1/4/8-byte patterns (e.g. algorithms/bfs.hpp:62-66). */
int sg_fill(sg_device_t dev, void* dst, uint64_t pattern, int pattern_bytes, size_t count);

/* ------------------------------------------------------------------ events */
/* sygraph::Event::getRuntime (sycl/event.hpp:36-41) as start/stop hipEvents. */
int sg_event_create(sg_device_t dev, sg_event_t* out);
int sg_event_destroy(sg_event_t ev);
int sg_event_record(sg_event_t ev);
int sg_event_synchronize(sg_event_t ev);
int sg_event_elapsed_ms(sg_event_t start, sg_event_t stop, float* ms);

/* ------------------------------------------------------------------- graph */
enum sg_graph_flags {
  SG_GRAPH_DIRECTED = 1,       /* graph/properties.hpp:29-32 Properties::directed */
  SG_GRAPH_WEIGHTED = 2,       /* Properties::weighted */
  SG_GRAPH_INPUT_DEVICE = 4,   /* input arrays are device pointers */
  SG_GRAPH_WEIGHT_FLOAT = 8,   /* weights are float32 (else uint32) */
  SG_GRAPH_SPACE_SHARED = 16,  /* store in managed memory (GRAPH_LOCATION=2) */
  SG_GRAPH_SPACE_HOST = 32     /* store in pinned host memory (GRAPH_LOCATION=0) */
};
/* Upload a CSR graph (graph/impls/graph_csr.hpp:170-178,394-426 +
 * graph/build.hpp:31-41).  offsets has n+1 entries of offset_bytes (4 or 8),
 * cols has m int32 ids, weights has m 4-byte values or is NULL.  For a
 * directed graph the inverse CSR is built on the device (counting sort
 * transpose); an undirected graph aliases it, as graph_csr.hpp:423-425. */
int sg_graph_create(sg_device_t dev, int64_t n, int64_t m, const void* offsets, int offset_bytes, const int32_t* cols,
                    const void* weights, int flags, sg_graph_t* out);
int sg_graph_destroy(sg_graph_t g);

typedef struct sg_graph_view {
  int64_t n, m;
  int32_t offset_bytes;
  int32_t flags;
  void* offsets;          /* device, n+1 entries */
  int32_t* cols;          /* device, m entries */
  void* weights;          /* device, m entries or NULL */
  void* inv_offsets;      /* inverse CSR (== offsets when undirected) */
  int32_t* inv_cols;
  void* inv_weights;
} sg_graph_view;
/* GraphCSR::getDeviceGraph / getInverseDeviceGraph (graph_csr.hpp:206-208). */
int sg_graph_get_view(sg_graph_t g, sg_graph_view* view);
/* Build the per-graph helpers that BFS otherwise builds on its first call:
 * the bitmap of vertices without in-arcs and the pull hint (per vertex, its
 * in-neighbour of largest out-degree).  ms (may be NULL) receives their
 * device time, 0 when already built.  No counterpart in the reference (its
 * pull reads every in-list from the start). */
int sg_graph_prepare(sg_graph_t g, float* ms);

/* Synthetic graphs, generated and assembled on the device (symmetrise, drop
 * self loops, sort, deduplicate; int64 offsets).  No reference counterpart:
 * the reference reads graphs from files (io/read_csr.hpp); the bench and the
 * tests need the seeded generators described in the benchmark configuration.
 *  rmat: 2^scale vertices, edge_factor*2^scale tuples, quadrant probabilities
 *        a,b,c (d=1-a-b-c), optional seeded id relabel;
 *  grid: rows x cols 4-neighbour lattice, each edge deleted with p_delete;
 *  er:   G(n, m_pairs) with uniform endpoints.
 * weight_seed != 0 attaches uint32 weights in [1,255], one per undirected edge. */
int sg_graph_generate_rmat(sg_device_t dev, int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                           int relabel, uint64_t weight_seed, sg_graph_t* out);
int sg_graph_generate_grid(sg_device_t dev, int rows, int cols, double p_delete, uint64_t seed, uint64_t weight_seed,
                           sg_graph_t* out);
int sg_graph_generate_er(sg_device_t dev, int64_t n, int64_t m_pairs, uint64_t seed, uint64_t weight_seed,
                         sg_graph_t* out);
/* Copy the CSR back to host buffers (any may be NULL). offsets get n+1 int64. */
int sg_graph_download(sg_graph_t g, int64_t* offsets, int32_t* cols, uint32_t* weights);
/* Order-sensitive digests sum_i mix64(i*K ^ x_i) of offsets and cols, computed
 * on the device (used to pin generated graphs without moving them). */
int sg_graph_digest(sg_graph_t g, uint64_t* offsets_digest, uint64_t* cols_digest);
/* The same order-sensitive digest of any device array of count 4- or 8-byte
 * elements (no reference counterpart: the autotuner compares outputs of
 * tuned runs without copying them to the host). */
int sg_digest(sg_device_t dev, const void* ptr, int64_t count, int elem_bytes, uint64_t* digest);
/* Sum of degrees of vertices whose depth != n+1 (algorithms/bfs.hpp:88-99,
 * the TEPS numerator). depth is a device int32 array. */
int sg_graph_reached_edges(sg_graph_t g, const int32_t* depth, int64_t* edges, int64_t* vertices);

/* ---------------------------------------------------------------- frontier */
/* FrontierMLB<T, 2> (frontier/impls/mlb_frontier.hpp:155-449): level 0 holds
 * one bit per element in 64-bit words, level 1 one bit per non-empty level-0
 * word.  Elements are < n. */
int sg_frontier_create(sg_device_t dev, int64_t n, sg_frontier_t* out);
int sg_frontier_destroy(sg_frontier_t f);
typedef struct sg_frontier_view {
  int64_t n;
  int64_t words0, words1;
  uint64_t* l0;        /* device */
  uint64_t* l1;        /* device */
  int32_t* list;       /* device compaction output (capacity n) */
  uint32_t* list_size; /* device counter written by compaction */
} sg_frontier_view;
/* MLBDevice accessors (mlb_frontier.hpp:58-153). */
int sg_frontier_get_view(sg_frontier_t f, sg_frontier_view* view);
/* check / insert / remove of one element (mlb_frontier.hpp:248-285). */
int sg_frontier_check(sg_frontier_t f, int64_t idx, int* present);
int sg_frontier_insert(sg_frontier_t f, int64_t idx);
int sg_frontier_remove(sg_frontier_t f, int64_t idx);
/* clear (mlb_frontier.hpp:367-380). */
int sg_frontier_clear(sg_frontier_t f);
int sg_frontier_clear_async(sg_frontier_t f);
/* empty / size (mlb_frontier.hpp:216-246, 287-311). */
int sg_frontier_empty(sg_frontier_t f, int* empty);
int sg_frontier_size(sg_frontier_t f, int64_t* size);
/* this |= other / this &= other; both rebuild the level-1 summary
 * (mlb_frontier.hpp:313-337; the reference's merge leaves level 1 stale). */
int sg_frontier_merge(sg_frontier_t f, sg_frontier_t other);
int sg_frontier_intersect(sg_frontier_t f, sg_frontier_t other);
/* out = a & b in one pass (operators/intersection/bitmap_intersect_impl.hpp:21-30). */
int sg_frontier_and(sg_frontier_t out, sg_frontier_t a, sg_frontier_t b);
/* computeActiveFrontier (mlb_frontier.hpp:388-442) as an ordered
 * set-scan-compact.  mode 0: element ids of set bits (ascending) into
 * view.list; mode 1: level-0 word ids that contain a set bit; mode 2 (pull):
 * level-0 word ids that are not all ones.  *count may be NULL (device count
 * stays in view.list_size). */
int sg_frontier_compact(sg_frontier_t f, int mode, int64_t* count);
int sg_frontier_compact_async(sg_frontier_t f, int mode);
/* saveState / loadState: copy both levels to / from host words
 * (mlb_frontier.hpp:339-365). Buffers hold words0 + words1 uint64. */
int sg_frontier_save(sg_frontier_t f, uint64_t* host_words);
int sg_frontier_load(sg_frontier_t f, const uint64_t* host_words);
/* Swap the storage of two frontiers of equal size (frontier.hpp:92-99). */
int sg_frontier_swap(sg_frontier_t a, sg_frontier_t b);

/* -------------------------------------------------------------- algorithms */
enum sg_direction_policy {
  SG_POLICY_PUSH = 0,    /* bfs_direction::push */
  SG_POLICY_PULL = 1,    /* bfs_direction::pull */
  SG_POLICY_HYBRID = 2,  /* bfs_direction::hybrid, reference alpha/beta rule */
  SG_POLICY_LEARNED = 3  /* per-level MLP choice of direction and balancer */
};
enum sg_load_balancer {
  SG_LB_WORKITEM = 0,    /* load_balancer::workitem_mapped (thread per vertex) */
  SG_LB_SUBGROUP = 1,    /* load_balancer::subgroup_mapped (wave per vertex) */
  SG_LB_WORKGROUP = 2,   /* load_balancer::workgroup_mapped (edge-balanced merge path) */
  SG_LB_BUCKETING = 3    /* load_balancer::bucketing (thread/wave/workgroup bins) */
};

#define SG_MAX_LEVELS_RECORDED 4096
typedef struct sg_bfs_options {
  int32_t policy;          /* sg_direction_policy */
  int32_t load_balancer;   /* sg_load_balancer: the work mapping of every level under the push, pull and
                              hybrid policies.  Push levels: BUCKETING (and WORKITEM) = lane / wave /
                              workgroup bins; WORKGROUP = the arcs of the lane and wave bins laid end to
                              end (edge-balanced); SUBGROUP = large lists split over waves instead of
                              workgroups.  Pull levels: lane-serial head of 16 / 4 / 64 in-arcs
                              (BUCKETING / WORKGROUP / SUBGROUP) before the wave joins. */
  float alpha, beta;       /* hybrid thresholds (algorithms/bfs.hpp:153) */
  int32_t record_levels;   /* fill per-level stats in sg_bfs_stats */
  int32_t sync_per_level;  /* 1: the host waits for every level before enqueuing the next (default 0:
                              one level ahead) */
  sg_policy_t learned;     /* SG_POLICY_LEARNED: the trained policy choosing direction and work mapping
                              per level */
  /* tuning knobs (0 = default); the depths never depend on them */
  int32_t t_medium;        /* degree from which a push vertex gets a wave (default 64) */
  int32_t t_large;         /* degree from which its arcs are split into chunks (default 1024) */
  int32_t chunk;           /* arcs per chunk of a large vertex (default 1024, at least 64) */
  int32_t blocks_per_cu;   /* workgroups per compute unit of the traversal kernel (default 4) */
  int32_t burst_arcs;      /* push levels whose frontier has at most this many arcs run back to back
                              inside one single-workgroup launch (default 2048, at most 65536;
                              negative: never) */
  int32_t grid_burst_arcs; /* push levels whose frontier has more arcs than burst_arcs but at most this
                              many run back to back inside one cooperative launch of the whole grid,
                              one grid barrier per level, once two steps in a row have left such a
                              level (default 262144, at most 2^24; negative: never) */
  int32_t grid_burst_workgroups; /* 512-thread workgroups of that launch (default 32; at most 1024, and
                                     never more than fit on the device at once) */
  int32_t learned_gate;    /* SG_POLICY_LEARNED: a level whose frontier has fewer arcs than the graph's
                              arcs / learned_gate pushes with the bucketing mapping without asking the
                              network (default 1024; negative: the network decides every level) */
} sg_bfs_options;

typedef struct sg_level_stat {
  int32_t direction;       /* 0 push, 1 pull */
  int32_t load_balancer;
  int64_t frontier_vertices; /* |F| entering the level */
  int64_t frontier_edges;    /* sum deg over F */
  int64_t discovered;        /* vertices discovered at this level */
  int64_t edges_inspected;   /* bitmap probes (push: sum deg F; pull: one hint probe per scanned
                                vertex plus k_u for each hint miss) */
  int64_t vertices_scanned;  /* pull: unvisited vertices with degree > 0 */
  int64_t hint_misses;       /* pull: scanned vertices whose hint in-neighbour was not in the frontier */
  int64_t unexplored_edges;  /* sum deg over unvisited after the level */
  int64_t max_degree;        /* largest degree among the discovered vertices */
  int32_t action;            /* direction * 3 + work-mapping variant (0 bucketing, 1 edge-balanced /
                                short pull head, 2 wave-split / long pull head) of this level */
  float kernel_ms;           /* device time of the level's traversal kernel (record_levels; 0 for the
                                levels of a burst) */
  float convert_ms;          /* device time of what follows it in the same step: the count, scan,
                                decision and queue write, and the burst of small levels if one ran */
  int32_t burst;             /* 1: the level ran inside a single-workgroup burst of small push levels,
                                timed in convert_ms of the level before it; the level at which a burst
                                stops is counted and decided by the next step, whose times it carries */
} sg_level_stat;

typedef struct sg_bfs_stats {
  int32_t levels;          /* BFSRunDetails::iterations (algorithms/bfs.hpp:31-35) */
  int32_t n_push, n_pull;
  int64_t reached_vertices;
  int64_t reached_edges;   /* TEPS numerator */
  float kernel_ms;         /* device time of the traversal (hipEvents) */
  sg_level_stat level[SG_MAX_LEVELS_RECORDED];
} sg_bfs_stats;

void sg_bfs_options_default(sg_bfs_options* opt);
/* BFS::run (algorithms/bfs.hpp:153-238): depth (device int32[n]) receives
 * hop counts, unreached = n+1 (bfs.hpp:62-67).  stats may be NULL. */
int sg_bfs(sg_graph_t g, int32_t source, const sg_bfs_options* opt, int32_t* depth, sg_bfs_stats* stats);

typedef struct sg_bfs_many_stats {
  int64_t reached_vertices; /* summed over the sources */
  int64_t reached_edges;    /* TEPS numerator, summed over the sources */
  float kernel_ms;          /* device time of all the traversals (hipEvents) */
} sg_bfs_many_stats;
/* sg_bfs from each of nsrc sources (host int32[nsrc]) in turn: row i of
 * depth (device int32[nsrc][n]) receives the hop counts from sources[i].  No
 * counterpart in the reference, whose callers run BFS::run once per source
 * (examples/bfs/bfs.cpp:102-124, scripts/benchmark.sh); the traversals are
 * sg_bfs's, enqueued back to back without a host synchronisation between
 * them.  options.record_levels is ignored; stats may be NULL. */
int sg_bfs_many(sg_graph_t g, int32_t nsrc, const int32_t* sources, const sg_bfs_options* opt, int32_t* depth,
                sg_bfs_many_stats* stats);

/* ---- learned per-level policy (no counterpart in the reference, whose only
 * policy is the alpha/beta rule of algorithms/bfs.hpp:325-350) -------------
 * An 8 -> 32 -> 32 -> 6 ReLU MLP in FP32 (1,542 parameters, layout W1[32][8]
 * b1[32] W2[32][32] b2[32] W3[6][32] b3[6]) scoring action = direction * 3 +
 * work-mapping variant from eight features of the frontier entering a level.
 * Features, scores and labels live in device memory. */
#define SG_POLICY_PARAMS 1542
#define SG_POLICY_FEATURES 8
#define SG_POLICY_ACTIONS 6
int sg_policy_create(sg_device_t dev, sg_policy_t* out); /* parameters start at zero */
int sg_policy_destroy(sg_policy_t p);
int sg_policy_set_params(sg_policy_t p, const float* host_params); /* also resets the optimizer */
int sg_policy_get_params(sg_policy_t p, float* host_params);
/* scores[batch][6] for features[batch][8] */
int sg_policy_forward(sg_policy_t p, const float* features, int32_t batch, float* scores);
/* gradient (device float[1542]) of the mean softmax cross-entropy of labels[batch];
 * loss (host float, may be NULL) receives the mean loss */
int sg_policy_gradient(sg_policy_t p, const float* features, const int32_t* labels, int32_t batch, float* grad,
                       float* loss);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss */
int sg_policy_train_step(sg_policy_t p, const float* features, const int32_t* labels, int32_t batch, float lr,
                         float* loss);

typedef struct sg_sssp_stats {
  int32_t rounds;            /* frontier rounds until no distance changed */
  int64_t reached_vertices;  /* dist != n+1 */
  int64_t reached_edges;     /* TEPS numerator (algorithms/sssp.hpp:66-77) */
  int64_t relaxations;       /* arcs relaxed over all rounds (light arcs in light rounds, heavy arcs in a
                                bucket's heavy round) */
  float kernel_ms;           /* device time of the whole run */
  int64_t round_vertices[SG_MAX_LEVELS_RECORDED]; /* |F| per round */
  int64_t round_edges[SG_MAX_LEVELS_RECORDED];    /* arcs relaxed per round: sum deg over F (Bellman-Ford),
                                                     or over F of the light / heavy arcs (buckets) */
  float relax_ms;            /* sg_sssp_options.record_relax: device time of the relax launches of the
                                relaxing rounds (hipEvents around each), else 0 */
  int32_t relax_launches;    /* those launches */
} sg_sssp_stats;
/* SSSP::run (algorithms/sssp.hpp:152-210): frontier Bellman-Ford with atomic
 * min; dist (device, 4-byte, same type as the weights) unreached = n+1.
 * stats may be NULL. */
int sg_sssp(sg_graph_t g, int32_t source, int32_t load_balancer, void* dist, sg_sssp_stats* stats);

typedef struct sg_sssp_options {
  int32_t load_balancer; /* sg_load_balancer (the bucketing bins in every case) */
  double delta;          /* bucket width of the near/far split: 0 = automatic (half the mean arc
                            weight on a graph of at least 2^25 arcs and 8 arcs per vertex; none
                            otherwise: on smaller graphs an iteration's launches cost more than the
                            relaxations buckets save, on sparse long-diameter ones such as grids
                            buckets multiply the iterations), negative or infinite =
                            none, i.e. every dirty vertex is relaxed each round (the reference's
                            frontier Bellman-Ford, sssp.hpp:152-210) */
  int32_t heavy_pull;    /* direction of a bucket's heavy round on an undirected graph: 0 = automatic
                            (pull when the heavy arcs of the vertices not yet settled are fewer than
                            those of the bucket's settled set), 1 = always pull, -1 = always push.
                            Pull: every vertex not yet settled takes the minimum over its heavy arcs
                            from the settled set; push: every heavy arc of the settled set is
                            relaxed with an atomic minimum.  Directed graphs always push. */
  int32_t burst_arcs;    /* rounds of the reference's schedule (no buckets) with at most this many
                            arcs run back to back inside one workgroup while they stay that small:
                            0 = default (4,096), negative = never, at most 8,192 */
  int32_t record_relax;  /* 1: time every relax launch on its own (two events per round; stats must
                            not be NULL): stats->relax_ms / relax_launches */
} sg_sssp_options;
void sg_sssp_options_default(sg_sssp_options* opt);
/* sg_sssp with options; the distances do not depend on delta. */
int sg_sssp_ex(sg_graph_t g, int32_t source, const sg_sssp_options* opt, void* dist, sg_sssp_stats* stats);
/* CC::run (algorithms/cc.hpp:103-169): labels (device uint32[n]) converge to
 * the largest vertex id that reaches each vertex along out-arcs: the largest
 * id of its component when the graph stores both directions of every arc
 * (union-find), the reference's label propagation otherwise (directed
 * graphs, and undirected ones whose CSR lacks some reverse arc; symmetry is
 * checked once per graph). */
int sg_cc(sg_graph_t g, uint32_t* labels, int32_t* rounds);
/* sg_cc without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same labels as sg_cc;
 * every buffer besides labels comes from a static array of the library
 * (32 MiB per device, held until the process ends), so graphs of at most
 * 2^27 vertices are accepted (SG_ERR_INVALID beyond).  Calls that overlap on
 * one device share that array and race; calls one after another are safe. */
int sg_cc_noalloc(sg_graph_t g, uint32_t* labels, int32_t* rounds);
/* TC::run + getNumTriangles (algorithms/tc.hpp:67-138): exact count on the
 * (degree, id)-oriented DAG; adjacency lists must be sorted. per_vertex
 * (device uint64[n] or NULL) gets the triangles attributed to each DAG source. */
int sg_tc(sg_graph_t g, uint64_t* count, uint64_t* per_vertex);
/* BC::run (algorithms/bc.hpp:118-200) for many sources: bc (device
 * float[n]) += sum_s delta_s(v), v != s, no /2.  Per source: fused BFS for
 * depths, vertices grouped by level, level-by-level pulls of sigma (FP64,
 * each level rescaled by a power of two) and of delta (FP64); the FP64 total
 * is added to bc once.  Sources are processed one after another. */
int sg_bc(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, float* bc);
/* Batched BC (no reference counterpart): the sums of sg_bc computed 64 sources
 * at a time, one bit-lane per source, each batch's whole level loop (forward
 * multi-source BFS with path counts, then dependencies deepest level first)
 * in one cooperative launch.  Path counts are FP32 mantissas with an int16
 * exponent per (vertex, source), dependencies FP32, so any diameter works;
 * equal to sg_bc up to FP32 rounding; added into bc (device float[n]).
 * SG_ERR_UNSUPPORTED if a path count reaches 2^32000. */
int sg_bc_batched(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, float* bc);
/* Multi-source BFS: depth is device int32[nsrc][n] (row i = sources_host[i]),
 * unreached = n+1.  Sources run 64 per batch, one bit-lane each: every vertex
 * holds 64-bit visited and level masks, so one pass over the arcs advances
 * all 64 traversals by a level; a batch's level loop runs in one cooperative
 * launch with a grid barrier per level.  No counterpart in the reference
 * (its BFS takes one source). */
int sg_bfs_multi(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, int32_t* depth);
/* Direction rule of each level of a multi-source batch.  policy: -1 pull when
 * the level holds more than n/20 vertices (the default); SG_POLICY_PUSH /
 * SG_POLICY_PULL always; SG_POLICY_HYBRID the reference's alpha / beta rule
 * (algorithms/bfs.hpp:341-350) on the batch's totals (frontier arcs against
 * the arcs of vertices some lane has not reached yet); SG_POLICY_LEARNED the
 * learned policy's direction on the same totals (its work-mapping part is not
 * used: a batch has one push and one pull mapping).  The depths do not
 * depend on it. */
typedef struct sg_msbfs_options {
  int32_t policy;
  float alpha, beta;
  sg_policy_t learned;
} sg_msbfs_options;
void sg_msbfs_options_default(sg_msbfs_options* opt);
/* sg_bfs_multi with a direction rule (opt may be NULL: the default). */
int sg_bfs_multi_ex(sg_graph_t g, int32_t nsrc, const int32_t* sources_host, const sg_msbfs_options* opt,
                    int32_t* depth);

#ifdef __cplusplus
}
#endif
#endif /* SYGRAPH_HIP_H */
