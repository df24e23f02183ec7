/*
 * gmp.h -- C ABI of the gfx950 vertex-centric graph engine (libgmp.so).
 *
 * The engine replaces GraphMP's shared-memory engine (cap-ntu/Graphee, GraphMP/system/):
 * a destination-interval CSR held in HBM replaces the per-iteration shard reads, and
 * hand-written HIP kernels replace the OpenMP per-shard gather.  Every entry point takes
 * plain host pointers and sizes; device memory never crosses this boundary.  All entry
 * points return GMP_OK (0) or a GMP_E* code and never throw; gmp_last_error() gives the
 * calling thread's last message.  An engine handle is not thread-safe: callers serialise
 * calls on one handle (the engine runs everything on one HIP stream per handle).
 */
#ifndef GMP_H_
#define GMP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMP_ABI_VERSION 4

/* status codes */
#define GMP_OK 0
#define GMP_EINVAL 1       /* bad argument (null pointer, size out of range) */
#define GMP_EIO 2          /* file missing / short read / write failure */
#define GMP_EFORMAT 3      /* malformed .npy or shard header */
#define GMP_EHIP 4         /* HIP runtime error */
#define GMP_ENOMEM 5       /* device or host allocation failed */
#define GMP_ESTATE 6       /* call out of order (e.g. run before a graph is loaded) */
#define GMP_EUNSUPPORTED 7 /* feature not available for this graph */

/* GraphMP-compatible apps for gmp_run_app: the per-vertex rules of
 * GraphMP/app/pagerank.cpp:42-59 (float), sssp.cpp:40-57 (int32, +1 per edge) and
 * cc.cpp:40-56 (int32 min label), plus the weighted SSSP generalisation (uint32,
 * saturating add of the per-edge weight). */
#define GMP_APP_PAGERANK 1
#define GMP_APP_SSSP 2
#define GMP_APP_CC 3
#define GMP_APP_SSSP_W 4

/* gmp_sssp methods */
#define GMP_SSSP_BF 0      /* Bellman-Ford: frontier pushes while the frontier is small
                              (GMP_OPT_PUSH_LOG), sweeps over the source strips that skip
                              unchanged strips while it is large; frontier relaxation alone
                              when strips are off */
#define GMP_SSSP_DELTA 1   /* fixed-delta stepping */
#define GMP_SSSP_LEARNED 2 /* delta chosen per bucket epoch by the trained MLP */

/* gmp_pr_params.flags */
#define GMP_PR_NO_KAHAN 1u   /* plain FP32 gather sum instead of the compensated one */
#define GMP_PR_COMPRESSED 2u /* gather from the delta/varint neighbour lists */
#define GMP_PR_STRIP_IDS 4u  /* source-strip gather reads each edge's 4-byte source id, as the
                                in-CSR stores it, in place of its 2-byte offset in the strip
                                (same ranks, bit for bit; not with GMP_PR_COMPRESSED) */

typedef struct gmp_engine gmp_engine;

typedef struct gmp_stats {
  int32_t iterations;  /* iterations executed */
  int32_t launches;    /* launches of the dominant kernel */
  double ms_total;     /* device time of the run, HIP events on the engine stream */
  double ms_kernel;    /* device time of the dominant kernel, summed over its launches */
  uint64_t edges;      /* edges gathered or relaxed */
  uint64_t vertices;   /* vertex updates applied */
  double residual;     /* PageRank: L1 change of the last iteration (FP64 reduction) */
  double dangling;     /* PageRank: dangling mass entering the last iteration (FP64) */
  uint64_t changed;    /* vertices changed in the last iteration */
  uint64_t bytes;      /* algorithmic bytes of the run (uncompressed yardstick) */
  uint64_t index_bytes; /* PageRank: bytes of the column-id stream read per iteration */
  int32_t epochs;      /* SSSP: delta-stepping bucket epochs */
  int32_t policy;      /* SSSP: 0 Bellman-Ford, 1 fixed delta, 2 learned delta, 3 trace exploration */
} gmp_stats;

typedef struct gmp_rmat_params {
  uint32_t scale;       /* V = 2^scale */
  uint32_t edge_factor; /* samples = edge_factor * V */
  uint64_t seed;
  double a, b, c;       /* quadrant probabilities, d = 1 - a - b - c */
  uint32_t undirected;  /* 1: add the reverse of every edge (symmetrised graph) */
  uint32_t weighted;    /* 1: attach uint32 weights in [1,255] */
  uint32_t num_shards;  /* destination-interval shards, equal in-edges per shard */
  uint32_t flags;       /* GMP_GRAPH_OFF64: keep 64-bit row offsets even when E < 2^32 */
} gmp_rmat_params;

#define GMP_GRAPH_OFF64 1u

typedef struct gmp_pr_params {
  float damping;       /* 0.85 */
  int32_t max_iter;    /* iteration cap */
  double tol;          /* > 0: stop once the L1 change of an iteration is below tol */
  int64_t seed_vertex; /* -1: uniform teleport; else personalised to this vertex */
  uint32_t flags;      /* GMP_PR_* */
  uint32_t reserved;
} gmp_pr_params;

typedef struct gmp_sssp_params {
  int32_t method;      /* GMP_SSSP_* */
  uint32_t delta;      /* bucket width for GMP_SSSP_DELTA (0 -> 64) */
  uint32_t unit_weights; /* 1: every edge weighs 1 (ignores stored weights) */
  uint32_t reserved;
} gmp_sssp_params;

/* ---- engine lifetime ---------------------------------------------------------------
 * Replace init_workers()/GraphPS<T>::init (GraphMP/system/Global.h:336-362,
 * GraphMP/system/GraphPS.h:77-151): one engine per GPU, no MPI. */
const char *gmp_last_error(void);
int gmp_abi_version(void);
int gmp_create(int32_t device, gmp_engine **out);
int gmp_destroy(gmp_engine *h);

/* ---- tunables -----------------------------------------------------------------------
 * GraphMP fixes its knobs at compile time (GraphMP/system/Global.h:45-58); the engine takes
 * them at run time.  Changing GMP_OPT_REORDER drops the work layout (rebuilt on next use). */
#define GMP_OPT_REORDER 1    /* 1: degree-ordered work layout (default), 0: as loaded */
#define GMP_OPT_STREAM_CAP 2 /* gather plan: edges per low-degree task, 64..4096 */
#define GMP_OPT_ROW_CAP 3    /* gather plan: rows per low-degree task, 1..4096 */
#define GMP_OPT_CHUNK 4      /* gather plan: edges per hub chunk, 64..2^20 */
#define GMP_OPT_WG_PER_CU 5  /* gather grid: workgroups per compute unit, 1..64 */
#define GMP_OPT_VEC_MIN 6    /* gather plan: rows above this many edges are reduced wave-wide */
#define GMP_OPT_STRIP_LOG 7  /* source strips of 2^k ids, 8..15 (default 15); 0 = PageRank row
                                gather, frontier CC and Bellman-Ford */
#define GMP_OPT_BIN_LOG 8    /* PageRank strips: destination bins of up to 2^k rows, 8..14 (default 14) */
#define GMP_OPT_BIN_TASK_LOG 9 /* PageRank strips: messages per bin task 2^k, 8..22 (default 20) */
/* Semi-external graphs, GraphMP's sliding window (GraphMP/system/GraphPS.h:256-386) with its
 * compressed edge cache (GraphMP/system/Global.h:129-255).  Read when a graph is loaded. */
#define GMP_OPT_EDGE_BUDGET_MB 10 /* 0 (default): graphs load whole into device memory.  n > 0: a
                                     graph whose column ids and weights take more than n MiB loads
                                     semi-externally -- they stay in pinned host memory, windows
                                     of destination rows stream through a ring of device slots
                                     and the windows that fit into n MiB stay cached on the
                                     device; gmp_pagerank, gmp_run_app, gmp_sssp and gmp_cc run
                                     over the windows, the other runs return GMP_EUNSUPPORTED */
#define GMP_OPT_WINDOW_MB 11      /* semi-external: id bytes per window, 1..4096 (default 64) */
#define GMP_OPT_WINDOW_PACKED 12  /* semi-external: 1 (default) windows hold delta + varint packed
                                     ids (decoded in the gather), 0 plain uint32 ids */
#define GMP_OPT_PUSH_LOG 13        /* CC and Bellman-Ford: frontiers of at most V >> k vertices are
                                     pushed over the out-edges, larger ones swept over the source
                                     strips (GraphPS's sparse/dense switch); 0 = sweeps only,
                                     0..31 (default 10).  Results do not depend on it. */
#define GMP_OPT_CC_UNION_FIND 14   /* 1 (default): gmp_cc on a symmetric graph (undirected R-MAT,
                                     edge lists loaded with GMP_EDGES_SYMMETRIZE) by sampled
                                     union-find; 0: min-label sweeps and pushes.  Same labels. */
#define GMP_OPT_DELTA_PULL_LOG 15  /* delta-stepping SSSP (fixed and learned): a round whose near
                                     frontier has more than E >> k out-edges relaxes them by a pull
                                     over the source strips (only the frontier's sources offer)
                                     instead of pushes; 0 = pushes only, 0..31 (default 3).  The
                                     relaxations, and the distances, are the same. */
#define GMP_OPT_LEARNED_FINISH 16  /* GMP_SSSP_LEARNED and gmp_sssp_trace: 0 (default) the eight classes
                                     are bucket widths 8 .. 1024; 1: class 7 instead finishes the
                                     solve as Bellman-Ford (pushes and strip sweeps) from the live
                                     far pile.  Distances are the same. */
/* Batched runs (gmp_batch_*); written by gmp_batch_autotune.  Results are the same for every
 * value (SSSP distances bit for bit; PPR ranks within rounding of the summation order). */
#define GMP_OPT_BATCH_SSSP_UNROLL 17  /* batched SSSP: edges per lane in flight, 2, 4 (default) or 8 */
#define GMP_OPT_BATCH_SSSP_GRID 18    /* batched SSSP: relax workgroups per CU, 1..64 (default 4) */
#define GMP_OPT_BATCH_PPR_PIECE 19    /* batched PPR: rows with more in-edges than this are cut into
                                        pieces of this many, 64..65536 (default 1024) */
#define GMP_OPT_BATCH_PPR_RUN_ROWS 20 /* batched PPR: rows per wave slot in a row run, 1..1024 (32) */
#define GMP_OPT_BATCH_PPR_RUN_EDGES 21 /* batched PPR: edges per wave slot in a row run, 64..65536 (2048) */
#define GMP_OPT_BATCH_PPR_LOADS 22    /* batched PPR: source loads in flight per lane, 4, 8 (default) or 16 */
#define GMP_OPT_BATCH_PPR_GRID 23     /* batched PPR: gather workgroups per CU, 1..64 (default 8) */
#define GMP_OPT_MLP_NET 24  /* the network the gmp_mlp_* calls act on: 0 (default) the delta-stepping
                              schedule of GMP_SSSP_LEARNED, 1 the semi-external window predictor of
                              GMP_OPT_WINDOW_PREFETCH */
#define GMP_OPT_WINDOW_PREFETCH 25 /* semi-external min sweeps (gmp_sssp, gmp_cc): 1 = the trained
                              window predictor's shard outputs pick the windows whose copies start
                              before the next sweep's activation is known, and those windows run
                              first; 0 (default) = exact activation only.  Results are the same. */
#define GMP_OPT_STRIP_TASK_LOG 26 /* strip gathers (PageRank, sweeps, pull rounds): 2^k wave tasks of
                              8192 edges per workgroup task, each task staging its strip once;
                              0..12 (default 6).  Results do not depend on it. */
#define GMP_OPT_SELECTIVE 27 /* 1 (default): sweeps (CC, Bellman-Ford, delta-stepping pulls, and the
                              semi-external windows of every min app) skip the strips and windows
                              none of whose sources changed in the sweep before -- GraphMP's
                              selective scheduling (GraphMP/system/GraphPS.h:318-368), exact; 0:
                              every strip and window every sweep.  Results are the same. */
#define GMP_OPT_BATCH_SSSP_SCAN_LOG 28 /* batched SSSP: a round whose frontier holds more than
                              (group vertices) >> k entries takes it from the mask array in
                              group-vertex order, one group at a time, instead of from the queue;
                              0..31 (default 8), 0 = always the queue */
#define GMP_OPT_BATCH_SSSP_HUB_LOG 29 /* batched SSSP: frontier entries with more than 2^k out-edges
                              are relaxed by a second launch in pieces of 2^k edges spread over the
                              grid; 6..24 (default 11), 0 = never */
int gmp_set_option(gmp_engine *h, int32_t key, int64_t value);
int gmp_get_option(gmp_engine *h, int32_t key, int64_t *value);

/* Autotuner (no reference counterpart: GraphMP's knobs are compile-time constants): times the
 * PageRank gather on the loaded graph for candidate bin thresholds, tile sizes, chunk sizes
 * and workgroup counts (coordinate descent, `iters` iterations per trial), then Bellman-Ford
 * for candidate GMP_OPT_PUSH_LOG values and delta = 64 stepping for candidate
 * GMP_OPT_DELTA_PULL_LOG values, applies the best and reports it.  Results of every app are
 * unchanged by any choice. */
typedef struct gmp_tune_result {
  uint32_t stream_cap, row_cap, chunk, vec_min, wg_per_cu, trials;
  double gather_ms_before, gather_ms_after; /* mean PageRank iteration time, first and best */
  uint32_t strip_log, bin_log, bin_task_log, push_log;
  uint32_t delta_pull_log, strip_task_log;
} gmp_tune_result;
int gmp_autotune(gmp_engine *h, int32_t iters, gmp_tune_result *out);
/* Tunables as a flat JSON object, written after tuning and read at start-up. */
int gmp_tune_save(gmp_engine *h, const char *path);
int gmp_tune_load(gmp_engine *h, const char *path);

/* ---- graph loading -----------------------------------------------------------------
 * In-CSR = destination rows with ascending, deduplicated source ids, the layout of
 * GraphMP's shards (GraphMP/tools/format_spark.py:137-157, GraphMP/system/Comp.h:21-50).
 * row_off has V+1 entries (global offsets); shard_bounds has P+1 entries (NULL: the
 * equal-in-edge partitioner of format_spark.py:85-108 with P shards, P=0 -> 1).
 * weights may be NULL.  Repeated sources in a row are accepted as parallel edges: each copy
 * counts in out-degrees and sums, and the lightest copy bounds a distance. */
int gmp_load_csr(gmp_engine *h, uint32_t V, uint64_t E, const uint64_t *row_off,
                 const uint32_t *col_idx, const uint32_t *weights, uint32_t P,
                 const uint32_t *shard_bounds);
/* Reads dir/{p}.edge.npy for p in [0,P) (cnpy rules, GraphMP/system/cnpy.h:73-173, shard
 * header [edge_num_raw, len(indices), len(indptr), start, end], Comp.h:37-42), optional
 * dir/{p}.weight.npy (uint32 aligned with indices).  Replaces load_edge
 * (GraphMP/system/Global.h:129-255): the shards are read once and stay in HBM. */
int gmp_load_shards(gmp_engine *h, const char *dir, uint32_t V, uint32_t P);
/* Reads a text edge list -- one "src dst [weight]" per line, non-negative integers separated by
 * spaces or tabs; empty lines and lines starting with '#' or '%' are skipped -- and loads it the
 * way GraphMP/tools/format_spark.py shards it: destination rows with ascending, deduplicated
 * sources, self-loops kept (:137-157), and P partitions of equal raw in-edge count, duplicates
 * counted (:54-108).  V = 0: the largest id + 1.  Duplicate edges keep their least weight. */
#define GMP_EDGES_WEIGHTED 1u   /* third column: uint32 weight */
#define GMP_EDGES_SYMMETRIZE 2u /* add the reverse of every edge (CC input) */
int gmp_load_edge_list(gmp_engine *h, const char *path, uint32_t V, uint32_t P, uint32_t flags);
/* Builds the R-MAT graph on the device (definition in oracle/rmat_ref.py's docstring). */
int gmp_generate_rmat(gmp_engine *h, const gmp_rmat_params *p);
int gmp_graph_info(gmp_engine *h, uint32_t *V, uint64_t *E, uint32_t *P, uint32_t *weighted);
/* Copies the device graph back; any output may be NULL.  row_off: V+1, col_idx/weights: E,
 * shard_bounds: P+1, out_degree: V. */
/* Residency of the loaded graph.  windows == 0: resident in device memory.  Counters run from
 * the load: bytes copied from host to device for windows, windows gathered, windows skipped
 * because no source of theirs changed (GraphMP's bloom-filter activation, GraphPS.h:318-351,
 * here exact at the granularity of source blocks). */
typedef struct gmp_window_stats {
  uint32_t windows;        /* windows of consecutive destination rows (0: resident graph) */
  uint32_t cached;         /* windows kept in device memory */
  uint32_t ring_slots;     /* device slots the other windows stream through */
  uint32_t packed;         /* 1: windows hold varint-packed ids */
  uint64_t host_bytes;     /* edge bytes in pinned host memory (ids + weights) */
  uint64_t device_bytes;   /* device bytes for edges: cache + ring */
  uint64_t streamed_bytes; /* host-to-device bytes copied for windows */
  uint64_t windows_run;
  uint64_t windows_skipped;
  uint64_t prefetched;            /* windows copied ahead by the learned prefetch */
  uint64_t prefetch_wasted_bytes; /* bytes of prefetched windows that then did not run */
} gmp_window_stats;
int gmp_windows(gmp_engine *h, gmp_window_stats *out);
int gmp_graph_download(gmp_engine *h, uint64_t *row_off, uint32_t *col_idx, uint32_t *weights,
                       uint32_t *shard_bounds, uint32_t *out_degree);
/* Seeded source pick (definition in oracle/rmat_ref.py): first vertex with out-degree > 0. */
int gmp_pick_source(gmp_engine *h, uint64_t seed, uint32_t salt, uint32_t *vertex);

/* ---- GraphMP-compatible run --------------------------------------------------------
 * GraphPS<T>::run (GraphMP/system/GraphPS.h:171-387) with the app's _comp
 * (GraphPS.h:25-34): msg_init is init_vertex()'s _VertexMsg (4-byte elements), the run is
 * synchronous (MsgNew written from Msg, swapped each step) and stops when no vertex
 * changed and step > 1, or after max_iter steps.  vertex_out is _VertexOut (pagerank
 * only, zeros read as 1).  msg_out / value_out receive _VertexMsg / _VertexValue. */
int gmp_run_app(gmp_engine *h, int32_t app, int32_t max_iter, const void *msg_init,
                const int32_t *vertex_out, void *msg_out, void *value_out, gmp_stats *st);

/* ---- engine apps -------------------------------------------------------------------- */
int gmp_pagerank(gmp_engine *h, const gmp_pr_params *p, float *rank_out, gmp_stats *st);
/* No counterpart in the reference.  For callers that cannot allocate device memory (work under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): gmp_pagerank's
 * standard PageRank (same parameters and formula; GMP_PR_COMPRESSED is refused) over a
 * destination-row CSR the caller holds in DEVICE memory -- d_row_off uint64[V+1], d_col_idx
 * uint32[E], d_out_degree uint32[V] -- with the ranks written to the caller's device buffer
 * d_rank_out float[V].  It calls no hipMalloc/hipFree: its working memory is a fixed per-GPU
 * device array sized for GMP_NOALLOC_MAX_VERTICES vertices (48 MiB), resolved by gmp_create and
 * kept until the process ends; larger V is refused with GMP_EINVAL.  Calls on one GPU share that
 * array: calls that overlap on one GPU (two engines or threads) race.  Runs on the engine's stream;
 * returns after the last iteration finished (with tol > 0 it waits for every iteration's residual). */
#define GMP_NOALLOC_MAX_VERTICES 4194304u
int gmp_pagerank_noalloc(gmp_engine *h, uint32_t V, uint64_t E, const uint64_t *d_row_off,
                         const uint32_t *d_col_idx, const uint32_t *d_out_degree, const gmp_pr_params *p,
                         float *d_rank_out, gmp_stats *st);
int gmp_sssp(gmp_engine *h, uint32_t source, const gmp_sssp_params *p, uint32_t *dist_out,
             gmp_stats *st);
/* Connected components of a symmetric graph: label = minimum vertex id in the component
 * (the fixed point of GraphMP/app/cc.cpp's min-label rule).  Synchronous min-label sweeps over
 * the source strips (st->iterations = sweeps, st->edges = edges gathered), or frontier
 * relaxation when strips are off; the labels are the same. */
int gmp_cc(gmp_engine *h, uint32_t *label_out, gmp_stats *st);

/* ---- learned scheduling ---------------------------------------------------------------
 * No reference counterpart.  A 16 -> 64 -> 72 MLP (8 bucket-width logits for delta in
 * {8, 16, ..., 1024} + 64 shard-activity logits; 5768 FP32 parameters laid out W1[16][64],
 * b1[64], W2[64][72], b2[72]) chooses the width of each delta-stepping bucket epoch for
 * GMP_SSSP_LEARNED.  It never changes which edges are relaxed: distances are identical to
 * every other method.  Until it is trained or set, GMP_SSSP_LEARNED runs fixed delta.
 *
 * gmp_sssp_trace runs SSSP while recording a row at explored bucket epochs (the first four,
 * then every power of two): the 16 features, and for every class the device time and
 * relaxations of finishing the run from that state with one epoch of that class followed by
 * width 64; the best class (fastest finish), which the run then uses, and the shards (64ths of
 * the vertex range) that joined the frontier in that epoch. */
#define GMP_MLP_IN 16
#define GMP_MLP_HIDDEN 64
#define GMP_MLP_CLASSES 8
#define GMP_MLP_SHARDS 64
#define GMP_MLP_PARAMS 5768
typedef struct gmp_trace_row {
  float feat[GMP_MLP_IN];
  int32_t best;         /* best width class: delta = 8 << best */
  uint32_t epoch;
  uint64_t shard_mask;  /* bit j: a vertex of shard j joined the near frontier */
  float ms[GMP_MLP_CLASSES];      /* time to finish: one epoch of each class, then the widest */
  float relaxed[GMP_MLP_CLASSES]; /* relaxations to finish with each class */
} gmp_trace_row;
int gmp_sssp_trace(gmp_engine *h, uint32_t source, const gmp_sssp_params *p, gmp_trace_row *rows,
                   uint32_t cap, uint32_t *nrows, uint32_t *dist_out, gmp_stats *st);
/* Per-class counts of the bucket epochs the last GMP_SSSP_LEARNED run of this engine chose
 * (counts[c], c < n <= GMP_MLP_CLASSES). */
int gmp_sssp_actions(gmp_engine *h, uint32_t *counts, uint32_t n);
/* The window predictor's trace (no reference counterpart), on a semi-external graph: min sweeps
 * over the windows -- kind 0 labels (CC), 1 unit-weight or 2 weighted distances from source --
 * recording one row per sweep: the predictor's 16 features of that sweep (best = -1: no class)
 * and, as shard_mask, the 64ths of the row range holding a window active one sweep later (the
 * windows the prefetch should have copied).  Train with GMP_OPT_MLP_NET = 1. */
int gmp_window_trace(gmp_engine *h, int32_t kind, uint32_t source, gmp_trace_row *rows, uint32_t cap,
                     uint32_t *nrows, gmp_stats *st);
int gmp_mlp_init(gmp_engine *h, uint64_t seed);
int gmp_mlp_set(gmp_engine *h, const float *params);
int gmp_mlp_get(gmp_engine *h, float *params);
/* Y[n][72] = network(X[n][16]) on the GPU. */
int gmp_mlp_forward(gmp_engine *h, const float *X, uint32_t n, float *Y);
/* Mean loss and its gradient (5768 floats) over a batch, computed by the GPU kernels. */
int gmp_mlp_loss_grad(gmp_engine *h, const float *X, const int32_t *labels, const uint64_t *masks,
                      uint32_t n, float *loss, float *grad);
/* `epochs` full-batch Adam steps (learning rate lr) on the GPU; returns the last loss. */
int gmp_mlp_train(gmp_engine *h, const float *X, const int32_t *labels, const uint64_t *masks,
                  uint32_t n, int32_t epochs, float lr, float *loss);

/* ---- batched multi-instance runs ----------------------------------------------------
 * No reference counterpart (GraphMP runs one app on one graph per process).  A batch keeps
 * many graphs resident at once (independent of the graph loaded above); instances name a
 * graph and a source/seed.  Each kind runs as persistent launches replayed from a HIP graph.
 * dist_out / rank_out receive the instances' vectors back to back (instance i has as many
 * entries as its graph has vertices), in instance order. */
int gmp_batch_generate(gmp_engine *h, uint32_t ngraphs, const gmp_rmat_params *params);
int gmp_batch_info(gmp_engine *h, uint32_t graph, uint32_t *V, uint64_t *E);
/* In-CSR of one batch graph with local ids (row_off: V+1, col_idx/weights: E). */
int gmp_batch_download(gmp_engine *h, uint32_t graph, uint64_t *row_off, uint32_t *col_idx, uint32_t *weights);
int gmp_batch_sssp(gmp_engine *h, uint32_t n, const uint32_t *graph, const uint32_t *source, uint32_t *dist_out,
                   gmp_stats *st);
int gmp_batch_ppr(gmp_engine *h, uint32_t n, const uint32_t *graph, const uint32_t *seed, float damping,
                  int32_t iters, float *rank_out, gmp_stats *st);
/* SSSP instances (ns) and personalised-PageRank instances (np) in one replayed HIP graph whose two
 * branches run at the same time on the device: two SSSP rounds and two PPR iterations per replay,
 * replayed until the frontier is empty and `iters` iterations ran.  Same outputs as
 * gmp_batch_sssp and gmp_batch_ppr; both stats carry the combined run's time.  Either count may
 * be 0. */
int gmp_batch_run(gmp_engine *h, uint32_t ns, const uint32_t *sssp_graph, const uint32_t *source, uint32_t np,
                  const uint32_t *ppr_graph, const uint32_t *seed, float damping, int32_t iters,
                  uint32_t *dist_out, float *rank_out, gmp_stats *st_sssp, gmp_stats *st_ppr);
/* Autotuner for the batched kernels (no reference counterpart): coordinate descent over the
 * GMP_OPT_BATCH_* values, timing gmp_batch_sssp on the given SSSP instances and a few
 * iterations of gmp_batch_ppr on the given PPR instances; applies and reports the best. */
typedef struct gmp_batch_tune_result {
  uint32_t sssp_unroll, sssp_grid_per_cu, ppr_piece, ppr_run_rows, ppr_run_edges, ppr_loads, ppr_grid_per_cu,
      trials;
  double sssp_ms_before, sssp_ms_after; /* one batched SSSP run, first and best */
  double ppr_ms_before, ppr_ms_after;   /* batched PPR, per iteration, first and best */
  uint32_t sssp_scan_log, sssp_hub_log;
} gmp_batch_tune_result;
int gmp_batch_autotune(gmp_engine *h, uint32_t ns, const uint32_t *sssp_graph, const uint32_t *source,
                       uint32_t np, const uint32_t *ppr_graph, const uint32_t *seed, float damping, int32_t iters,
                       gmp_batch_tune_result *out);

/* ---- host-side format utilities (no GPU needed) -----------------------------------
 * Equal-in-edge partitioner, format_spark.py:85-108. */
int gmp_partition_bounds(const uint64_t *indeg, uint32_t V, uint32_t P, uint32_t *bounds);
/* 1-D little-endian .npy reader under cnpy's rules (cnpy.h:73-126); *n_out = element
 * count, word size must be 4.  buf may be NULL to query the size. */
int gmp_npy_read_u32(const char *path, uint32_t *buf, uint64_t cap, uint64_t *n_out);
int gmp_npy_write_i32(const char *path, const int32_t *data, uint64_t n);
/* Writes dir/{p}.edge.npy for every shard (empty shards included), dir/{p}.weight.npy when
 * weights != NULL, and dir/vertexout.npy when vertex_out != NULL. */
int gmp_write_shards(const char *dir, uint32_t V, const uint64_t *row_off,
                     const uint32_t *col_idx, const uint32_t *weights, uint32_t P,
                     const uint32_t *bounds, const int32_t *vertex_out);
/* format_spark.py natively: the edge list of gmp_load_edge_list written as GraphMP shards --
 * dir/{p}.edge.npy with the header [raw edge lines, nnz, rows + 1, start, end], dir/{p}.weight.npy
 * with GMP_EDGES_WEIGHTED, dir/vertexout.npy (out-degrees of the deduplicated graph).  *V_out
 * (may be NULL) receives V.  No GPU is used. */
int gmp_convert_edge_list(const char *path, const char *dir, uint32_t V, uint32_t P, uint32_t flags,
                          uint32_t *V_out);

#ifdef __cplusplus
}
#endif
#endif /* GMP_H_ */
