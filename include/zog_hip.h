/*
 * zog_hip.h -- C-ABI of the MI355X (gfx950) backend for Zog's `:soa`
 * ArrayGraph resource: unweighted Brandes betweenness and PageRank.
 *
 * The entry points are what the Zigler NIF bodies in lib/zog/resource_graph.ex
 * call in place of the Zig CPU code.  Plain pointers and sizes only; the
 * caller owns every host buffer.  Status codes are 0 on success and negative
 * on error (zg_strerror names them).
 *
 * Threading: every call on one handle takes that handle's lock, so concurrent
 * BEAM dirty-scheduler calls on one resource are safe (they serialise).  Calls
 * on different handles run concurrently, each on its handle's own HIP stream.
 */
#ifndef ZOG_HIP_H
#define ZOG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define ZG_OK 0
#define ZG_ERR_INVALID_ARG (-1)
#define ZG_ERR_NODE_NOT_FOUND (-2)   /* ArrayGraph.Error.NodeNotFound, array_graph.zig:77-80 */
#define ZG_ERR_OOM (-3)              /* host allocation failed (OutOfMemory) */
#define ZG_ERR_GPU_OOM (-4)          /* hipMalloc failed */
#define ZG_ERR_HIP (-5)              /* any other HIP runtime error */
#define ZG_ERR_NO_DEVICE (-6)        /* no usable gfx950 device */
#define ZG_ERR_TOO_LARGE (-7)        /* more than 2^32-1 arcs or nodes */
#define ZG_ERR_INVALID_HANDLE (-8)   /* NULL or destroyed handle */
#define ZG_ERR_IO (-9)               /* a file could not be opened or read */
#define ZG_ERR_SCRATCH (-10)         /* more than the fixed scratch of zg_pagerank_device_noalloc holds */

typedef struct zg_graph zg_graph;

/* Per-run information for Brandes (the reference returns none). */
typedef struct {
    uint64_t n_sources;       /* sources processed */
    uint64_t arcs_traversed;  /* sum over sources of arcs scanned by the forward
                                 sweep (arcs whose tail is reachable) */
    uint64_t nodes_reached;   /* sum over sources of reached vertices */
    uint32_t max_levels;      /* deepest BFS over all sources, in levels
                                 (zg_bc_f64: most arcs on a shortest path, + 1) */
    uint32_t batches;         /* batches of the multi-source sweep */
    uint32_t push_levels;     /* forward levels run from the frontier (push;
                                 zg_bc_f64: distance rounds launched) */
    uint32_t pull_levels;     /* forward levels run over predecessor chunks (pull) */
    double kernel_ms;         /* device time of the sweep (HIP events) */
    uint64_t arcs_relaxed;    /* zg_bc_f64: arcs scanned by the distance rounds,
                                 a vertex's arcs once per round it is queued
                                 in (0 for the unweighted sweeps) */
    /* Work of the batched (64-source group) sweep, for its byte model; 0 from
     * the other engines.  A row visit is a (group, level, vertex) with the
     * vertex on that level for at least one of the group's sources, levels
     * >= 1; row_arcs sums the out-degrees of the visited rows; tree_arcs is
     * the number of (source, arc v->w) pairs with depth(w) = depth(v) + 1
     * and depth(v) >= 1 (the arcs the dependency pass follows), counted only
     * when ZG_TUNE_COUNT_TREE_ARCS is on. */
    uint64_t row_visits;
    uint64_t row_arcs;
    uint64_t tree_arcs;
    /* Load balance of the batched sweep: source slots swept (64 per group,
     * empty slots included; compare n_sources), and the groups' vertex spans
     * over the span capacity of their launches (1 = every group of every
     * batch as large as the batch's largest). */
    uint64_t group_slots;
    double span_balance;
} zg_bc_info;

/* Per-run information for PageRank (and the eigenvector, Katz and alpha
 * iterations below). */
typedef struct {
    uint64_t iterations;      /* power iterations performed */
    double last_l1;           /* L1 change of the last iteration, reduced in FP64
                                 (eigenvector: the L2 change) */
    double residual_l1;       /* ||G(r) - r||_1 of the returned vector,
                                 evaluated in FP64 from the stored FP32 ranks */
    double error_bound_l1;    /* residual_l1 / (1 - damping): bound on the
                                 L1 distance of the returned vector to the
                                 exact fixed point */
    double kernel_ms;         /* device time of the iterations (HIP events) */
    int32_t precision;        /* ZG_PR_FP32, ZG_PR_FP64 or ZG_PR_MIXED */
    int32_t outcome;          /* zg_eigenvector: ZG_EIG_* (0 otherwise) */
    uint64_t fp64_iterations; /* iterations run in FP64 by the mixed mode */
} zg_pr_info;

/* PageRank precision modes. */
#define ZG_PR_FP32 32     /* FP32 values, Kahan row sums, FP64 reductions */
#define ZG_PR_FP64 64     /* FP64 throughout (same data layout) */
#define ZG_PR_MIXED 96    /* FP32 Kahan down to L1 change 1e-6, then FP64 from
                             that iterate to the tolerance (for tolerances below
                             the FP32 rounding floor) */
#define ZG_PR_SKIP_RESIDUAL 0x100  /* or-ed into precision: skip the FP64
                                      residual evaluation */

/* Replaces `new/5` + `buildGraph` (lib/zog/resource_graph.ex:188-200,
 * :120-129) for the `:soa` backend: node ids 0..n_cap-1, arcs from[i]->to[i]
 * in array order (parallel arcs and self-loops kept, weights not needed by
 * either path).  live_mask (nullable = all live) marks tombstoned ids the way
 * ArrayGraph.removeNode leaves them (array_graph.zig:133-157).  Builds the
 * successor CSR (reverse insertion order, array_graph.zig:94-104,166-180) and
 * the predecessor CSR (ascending source id, duplicates kept,
 * centrality.zig:881-899) once and keeps both resident in HBM.  An id >=
 * n_cap returns ZG_ERR_NODE_NOT_FOUND (addEdge, array_graph.zig:89). */
int zg_graph_create(uint32_t n_cap, const uint8_t *live_mask, const uint32_t *from,
                    const uint32_t *to, uint64_t m, zg_graph **out);

/* Same as zg_graph_create, keeping the arc weights: weight[i] (f64) is the
 * weight of arc i, as buildGraph stores it with addEdge(f, t, w)
 * (lib/zog/resource_graph.ex:120-129; weight is the `weight` list new/5
 * takes, :188).  Only zg_bc_f64 reads the weights; the other entry points
 * behave as on a zg_graph_create handle. */
int zg_graph_create_weighted(uint32_t n_cap, const uint8_t *live_mask, const uint32_t *from,
                             const uint32_t *to, const double *weight, uint64_t m, zg_graph **out);

/* Arc count from which zg_graph_create* build the CSR on the GPU (stable
 * radix sorts; default 2^18; below it, host loops).  Process-wide; the
 * arrays are the same either way.  No reference counterpart. */
int zg_set_build_threshold(uint64_t min_arcs);

/* Copies the resident successor CSR (out_ptr[n_cap+1], out_col[arcs]: each
 * row in the reference's iteration order, newest arc first), predecessor CSR
 * (in_ptr, in_col: each row ascending by source, parallel arcs repeated) and
 * the PageRank ids (pr_id[n_cap]) to host buffers; any pointer may be NULL.
 * Returns 1 if the CSR was built on the GPU, 0 if on the host, <0 on error.
 * No reference counterpart (parity checks, callers that want the arrays). */
int zg_graph_csr(zg_graph *g, uint32_t *out_ptr, uint32_t *out_col, uint32_t *in_ptr, uint32_t *in_col,
                 uint32_t *pr_id);

/* Work tables derived from the CSR at load (on the device), for
 * zg_graph_table.  Each is an array of u32 words:
 *   ZG_TABLE_ICH        predecessor chunks: vertex[n], first arc[n], end[n]
 *   ZG_TABLE_ICH_FIRST  first predecessor chunk of each vertex [n_cap + 1]
 *   ZG_TABLE_OCH        successor chunks: vertex[n], first arc[n], end[n]
 *   ZG_TABLE_OCH_FIRST  first successor chunk of each vertex [n_cap + 1]
 *   ZG_TABLE_OHUB       first partial-sum slot of each vertex with several
 *                       successor chunks, else 0xFFFFFFFF [n_cap]
 *   ZG_TABLE_PR_ROWS    PageRank rows in bin order: short, medium (4-, 16-
 *                       and 64-lane bins) [n_short + n_medium]
 *   ZG_TABLE_PR_DESC    {PageRank id, first binned arc, in-degree,
 *                       out-degree} of those rows [4 (n_short + n_medium)]
 *   ZG_TABLE_PR_LARGE   long rows: PageRank ids[k], out-degrees[k], first
 *                       segment[k + 1]
 *   ZG_TABLE_PR_SEG     long-row segments: row[s], first arc[s], end[s]
 *   ZG_TABLE_PR_COL     predecessor columns as PageRank ids, rows in bin
 *                       order [arcs]
 *   ZG_TABLE_PR_BINS    row counts: short, 4-, 16-, 64-lane, long [5] */
#define ZG_TABLE_ICH 0
#define ZG_TABLE_ICH_FIRST 1
#define ZG_TABLE_OCH 2
#define ZG_TABLE_OCH_FIRST 3
#define ZG_TABLE_OHUB 4
#define ZG_TABLE_PR_ROWS 5
#define ZG_TABLE_PR_DESC 6
#define ZG_TABLE_PR_LARGE 7
#define ZG_TABLE_PR_SEG 8
#define ZG_TABLE_PR_COL 9
#define ZG_TABLE_PR_BINS 10

/* Copies work table `which` (ZG_TABLE_*) to out[0..cap) and its length in
 * words to *n (out may be NULL to ask for the length only; cap < *n is
 * ZG_ERR_INVALID_ARG).  No reference counterpart: for tests and diagnostics
 * of the load. */
int zg_graph_table(zg_graph *g, int which, uint32_t *out, uint64_t cap, uint64_t *n);

/* Replaces the parsing half of `nif_read_edgelist` (resource_graph.ex:
 * 990-1084) with `addParsedEdges` (:953-988): reads an edge list (lines of
 * `src dst [weight]`, separators any of space, tab, comma; `#` comments; a
 * `source target` header; weight 1.0 when absent or not a number) with
 * integer ids (integer_labels != 0; node count = largest id + 1; lines whose
 * ids do not parse as u32 are skipped) or labels numbered by first
 * appearance, and expands it into arcs: file order when directed, else one
 * arc pair per unordered pair (the first in file order, by a stable sort on
 * (min, max); one arc for a self-loop).  The arcs then go to
 * zg_graph_create_weighted.  Host memory only.  ZG_ERR_IO when the file
 * cannot be read. */
typedef struct zg_edgelist zg_edgelist;
int zg_edgelist_read(const char *path, int directed, int integer_labels, zg_edgelist **out);
/* The expanded arcs and the node count; pointers stay valid until free. */
int zg_edgelist_arcs(const zg_edgelist *el, const uint32_t **from, const uint32_t **to, const double **weight,
                     uint64_t *n_arcs, uint32_t *n_nodes);
/* Labels in id order: label i is blob[offsets[i] .. offsets[i+1]) (none
 * with integer labels). */
int zg_edgelist_labels(const zg_edgelist *el, const char **blob, const uint64_t **offsets, uint32_t *n_labels);
void zg_edgelist_free(zg_edgelist *el);

/* Batch of independent graphs in one resource (no reference counterpart; the
 * reference builds one resource per graph with new/5): graph i has
 * node_counts[i] nodes and arc_counts[i] arcs whose local ids are in from/to,
 * concatenated in graph order.  Ids are placed one graph after another, so
 * results of zg_bc_unweighted over all sources (sources = NULL) are every
 * graph's own BC, concatenated: paths never cross graphs.  The sweep limits
 * each source group's work to the graphs its sources belong to.  PageRank on
 * a batch of more than one graph returns ZG_ERR_INVALID_ARG. */
int zg_graph_create_batch(uint32_t n_graphs, const uint32_t *node_counts, const uint64_t *arc_counts,
                          const uint32_t *from, const uint32_t *to, zg_graph **out);

/* Number of graphs in the resource (1 for zg_graph_create). */
int zg_graph_components(const zg_graph *g, uint32_t *n_components);

/* Replaces the resource destructor (resource_graph.ex:109-118); the NIF's
 * `nif_destroy` (:202-204) only releases the creation reference, so the
 * dtor is where this is called.  NULL is a no-op. */
void zg_graph_destroy(zg_graph *g);

/* Node capacity V (nodeCapacity, array_graph.zig:206-208), live node count
 * (nodeCount, :196-198) and live arc count; any out pointer may be NULL. */
int zg_graph_info(const zg_graph *g, uint32_t *n_cap, uint32_t *n_live, uint64_t *n_arcs);

/* Replaces zog.centrality.betweennessUnweighted on an ArrayGraph
 * (priv/zog/src/centrality.zig:694-798) as called by
 * `nif_betweenness_unweighted` (resource_graph.ex:206-215) followed by
 * extractScores (:166-174): unnormalised directed-sum BC over all live
 * sources into out_bc[n_cap] (0.0 for non-live ids).  sources == NULL means
 * every live node (the reference); otherwise only the listed sources are
 * summed (an extension for sampled BC; a non-live source is
 * ZG_ERR_NODE_NOT_FOUND).  The x0.5 undirected scaling stays with the
 * caller, as in maybe_scale_undirected (resource_graph.ex:2633-2637).
 * info may be NULL. */
int zg_bc_unweighted(zg_graph *g, const uint32_t *sources, uint64_t n_sources, double *out_bc,
                     zg_bc_info *info);

/* Replaces zog.centrality.betweennessF64 on an ArrayGraph(void, f64)
 * (priv/zog/src/centrality.zig:1355-1357 -> betweenness, :529-667) as
 * called by `nif_betweenness_f64` (resource_graph.ex:217-226) followed by
 * extractScores: single-source shortest paths under the arc weights (sums
 * added in path order, a path is shortest when its sum equals the distance
 * exactly), path counts over every such arc, unnormalised directed-sum
 * dependencies into out_bc[n_cap] (0.0 for non-live ids).  The handle must
 * come from zg_graph_create_weighted (else ZG_ERR_INVALID_ARG when it has
 * arcs).  sources / n_sources as zg_bc_unweighted (NULL = every live node,
 * the reference).  info may be NULL.
 * A live arc with a NaN, negative or infinite weight is ZG_ERR_INVALID_ARG
 * (the reference's Dijkstra assumes finite non-negative weights and returns
 * order-dependent values otherwise).  Weight 0 is accepted; where arcs of
 * weight 0 join vertices at equal distance, the reference's values depend on
 * its heap's pop order, while this path counts every tight arc in DAG order
 * and leaves vertices on a zero-weight cycle without paths.  Path counts are
 * FP64 (exact below 2^53; the reference's are u64). */
int zg_bc_f64(zg_graph *g, const uint32_t *sources, uint64_t n_sources, double *out_bc, zg_bc_info *info);

/* Same computation with device-resident source list and output (both in HBM,
 * d_sources may be NULL for all live nodes) on the caller's HIP stream
 * (hipStream_t passed as void*, NULL = the handle's stream); returns after
 * the stream has drained.  For callers that keep results on the GPU. */
int zg_bc_unweighted_device(zg_graph *g, const uint32_t *d_sources, uint64_t n_sources,
                            double *d_out_bc, void *stream, zg_bc_info *info);

/* Replaces zog.centrality.pagerank on an ArrayGraph
 * (priv/zog/src/centrality.zig:871-965, options :16-20) as called by the
 * `pagerank/4` NIF (resource_graph.ex:250-264) followed by extractScores:
 * out[n_cap] (0.0 for non-live ids).  Same iteration, dangling rule, strict
 * `l1 < tolerance` stop after the swap, n=0 and n=1 special cases.
 * precision is ZG_PR_FP32, ZG_PR_FP64 or ZG_PR_MIXED.  info may be NULL. */
int zg_pagerank(zg_graph *g, double damping, uint64_t max_iterations, double tolerance,
                int precision, double *out, zg_pr_info *info);

/* Same computation with the output written to device memory d_out[n_cap]
 * (FP64) on the caller's stream (NULL = the handle's stream). */
int zg_pagerank_device(zg_graph *g, double damping, uint64_t max_iterations, double tolerance,
                       int precision, double *d_out, void *stream, zg_pr_info *info);

/* Allocation-free form of zg_pagerank_device: no counterpart in the
 * reference; for callers that cannot allocate device memory (work under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same computation and result, bit for bit; every buffer the iteration needs
 * beyond the resident graph, the caller's d_out and stream comes from one
 * fixed device array of 64 MiB that the library holds on each device it runs
 * on until the process ends: its rank vectors take 16 bytes per node
 * capacity in FP32 (ZG_PR_FP32 | ZG_PR_SKIP_RESIDUAL) and 48 otherwise
 * (FP64 vectors for the FP64 phases and the residual check): up to about
 * 1.39 million nodes, or just under 4.19 million in FP32 without the
 * residual check.  A larger graph returns ZG_ERR_SCRATCH.  Creates no events either: info->kernel_ms
 * is 0 and no per-kernel profile is taken.  Calls on one device take turns
 * and share the array, which keeps the last call's contents (every call
 * rewrites what it reads); the call returns after the stream has drained. */
int zg_pagerank_device_noalloc(zg_graph *g, double damping, uint64_t max_iterations, double tolerance,
                               int precision, double *d_out, void *stream, zg_pr_info *info);

/* How zg_eigenvector's iteration ended (zg_pr_info.outcome). */
#define ZG_EIG_CONVERGED 0    /* L2 change below tolerance: the new iterate */
#define ZG_EIG_OSCILLATING 1  /* iterate within tolerance of the one two steps
                                 back: the normalised average of the last two */
#define ZG_EIG_ZERO 2         /* the iterate vanished: 1/sqrt(n) everywhere */
#define ZG_EIG_CAP 3          /* max_iterations reached: the last iterate */
#define ZG_EIG_TRIVIAL 4      /* at most one live node: 1.0 */

/* Replaces the `eigenvector/3` NIF (resource_graph.ex:266-275) over
 * zog.centrality.eigenvector (priv/zog/src/centrality.zig:1050-1201) on an
 * ArrayGraph, followed by extractScores: out[n_cap] (0.0 for non-live ids).
 * Power iteration over the predecessor lists (utils.zig:52-73) in FP64 from
 * 1 + i/1000 for the i-th live id, L2-normalised each step, with the
 * reference's zero-vector, oscillation (compared with the iterate two steps
 * back), convergence and iteration-cap rules in that order.  info may be
 * NULL. */
int zg_eigenvector(zg_graph *g, uint64_t max_iterations, double tolerance, double *out, zg_pr_info *info);

/* Replaces the `katz/5` NIF (resource_graph.ex:277-286) over
 * zog.centrality.katz (centrality.zig:1212-1267): x <- alpha * sum over
 * in(v) of x + beta from x = beta, FP64, stop after the step whose L1 change
 * is below tolerance.  out[n_cap]. */
int zg_katz(zg_graph *g, double alpha, double beta, uint64_t max_iterations, double tolerance, double *out,
            zg_pr_info *info);

/* Replaces the `alpha_centrality/5` NIF (resource_graph.ex:288-297) over
 * zog.centrality.alphaCentrality (centrality.zig:1278-1333): as zg_katz with
 * x <- alpha * sum from x = initial.  out[n_cap]. */
int zg_alpha_centrality(zg_graph *g, double alpha, double initial, uint64_t max_iterations, double tolerance,
                        double *out, zg_pr_info *info);

/* Measurement only (no reference counterpart): the device time of one
 * plain gather pass sum_e x[col[e]] over the resident predecessor columns
 * (one thread per arc; precision 32 or 64; averaged over `reps` passes after
 * a warm-up), the floor for any pull iteration over these columns.  With
 * ZG_FLOOR_LDS_COPY or'ed into precision the pass runs in the PageRank
 * iteration kernel's launch shape and reads the most-gathered values from
 * the same LDS copy: the floor of that layout. */
#define ZG_FLOOR_LDS_COPY 0x100
int zg_gather_floor(zg_graph *g, int precision, int reps, double *us_per_pass);

/* Per-source BFS export for parity checks (no reference counterpart; the
 * reference keeps these workspaces private, centrality.zig:697-703): for each
 * of n_sources sources, depth_out[i*n_cap + v] (0xFFFFFFFF = unreached) and
 * sigma_out[i*n_cap + v] = stored path count (FP64, exact integer below 2^53)
 * with shift_out[i*max_levels + L] = the power-of-two shift k_L applied to
 * level L+1's stored counts (true count at level L = stored * 2^(k_0+..+k_{L-1})).
 * levels_out[i] = number of BFS levels.  shift_out may be NULL; max_levels
 * bounds its row length (ZG_ERR_INVALID_ARG if a BFS is deeper). */
int zg_bc_paths(zg_graph *g, const uint32_t *sources, uint64_t n_sources, uint32_t *depth_out,
                double *sigma_out, int32_t *shift_out, uint32_t max_levels, uint32_t *levels_out);

/* Sets the number of 64-source groups swept per batch (1..64, default 4 =
 * 256 sources per launch).  Returns the previous value, or <0. */
int zg_set_batch_groups(zg_graph *g, int groups);

/* Tuning parameters of the Brandes sweep (no reference counterpart; results
 * are the same up to FP64 summation order).
 *   ZG_TUNE_PUSH_ALPHA: a forward level runs from its frontier (push) when the
 *   frontier's successor arcs times alpha are below the graph's arc count,
 *   otherwise over all predecessor chunks (pull).  Default 8; 0 = always
 *   pull; a tiny positive value = push whenever the frontier is non-empty.
 * Returns 0 or <0. */
#define ZG_TUNE_PUSH_ALPHA 1
/*   ZG_TUNE_ENGINE: ZG_ENGINE_AUTO (default: the per-source queue engine for
 *   graphs whose BFS from the hub core does not settle within 64 levels, the
 *   64-lane engine otherwise), ZG_ENGINE_LANES or ZG_ENGINE_QUEUE. */
#define ZG_TUNE_ENGINE 2
#define ZG_ENGINE_AUTO 0
#define ZG_ENGINE_LANES 1
#define ZG_ENGINE_QUEUE 2
/*   ZG_TUNE_PULL_GRID: workgroups per 64-source group for a pull level (they
 *   stride over the predecessor chunks); 0 = one wavefront per chunk.
 *   Default 4096. */
#define ZG_TUNE_PULL_GRID 3
/*   ZG_TUNE_GRAPH: 1 (default) replays the forward levels from a captured
 *   HIP graph (the level index lives on the device), 0 launches every level
 *   kernel from the host.  Per-kernel profiling always launches directly. */
#define ZG_TUNE_GRAPH 4
/*   ZG_TUNE_PR_CACHE: PageRank values of the most-gathered vertices each
 *   iteration block copies into LDS (default, when they serve at least 10%
 *   of the gathers: 16 KiB for 256-thread blocks; for larger blocks as many
 *   as let a full CU's blocks -- 2048 threads -- share its LDS, each keeping
 *   1 KiB; a larger value, up to the device's per-block limit, gives fewer
 *   blocks per CU; 0 = none).  ZG_TUNE_PR_GRID: PageRank iteration blocks (0 = default:
 *   2048 256-thread blocks, or as many larger blocks as fit the CUs at once).
 *   ZG_TUNE_PR_BLOCK: threads per PageRank iteration block, 256, 512 or
 *   1024 (0 = default: 1024 when the LDS copy is used or some row has more
 *   than 2048 predecessors -- a whole block sums such a row -- else 256).
 *   Sums of rows up to 2048 predecessors are the same under every setting;
 *   a longer row's sum follows the block size, and the FP64 sums of the
 *   dangling mass and of the change are added per block, so another shape
 *   can move the last bits of the ranks (a given setting repeats bit for
 *   bit). */
#define ZG_TUNE_PR_CACHE 5
#define ZG_TUNE_PR_GRID 6
#define ZG_TUNE_PR_BLOCK 12
/*   ZG_TUNE_WEIGHTED_BATCH: sources per batch of zg_bc_f64, a multiple of 64
 *   up to 2048 (0 = default: up to 2048, as half the free device memory
 *   allows). */
#define ZG_TUNE_WEIGHTED_BATCH 7
/*   ZG_TUNE_WEIGHTED_DELTA: zg_bc_f64 relaxes, per source, only vertices
 *   whose distance is within a bound that rises in steps of this value once
 *   nothing under it is left (0 = default: 0.3 x mean arc weight / mean
 *   out-degree; +inf = no bound).  Only the amount of work changes. */
#define ZG_TUNE_WEIGHTED_DELTA 8
/*   ZG_TUNE_WEIGHTED_GRID: workgroups of zg_bc_f64's per-round kernels
 *   (0 = default: when batch sources x node capacity >= 2^28, 12288, with
 *   32768 for the tight-arc count and sigma peel; else 4096; the far-pile
 *   split uses at most 4096). */
#define ZG_TUNE_WEIGHTED_GRID 9
/*   ZG_TUNE_COUNT_TREE_ARCS: 1 makes the 64-lane engine count per-lane
 *   tree arcs (zg_bc_info.tree_arcs) in its backward pass (one atomic per
 *   work-list entry; for measurement runs), 0 (default) does not. */
#define ZG_TUNE_COUNT_TREE_ARCS 10
/*   ZG_TUNE_QUEUE_ORDER: 1 (default) runs the per-source queue engine on a
 *   copy of the CSR renumbered in the order of one breadth-first search, so
 *   the neighbour records a level touches share lines; 0 keeps the ids.
 *   Rows keep their arc order: results are the same bit for bit. */
#define ZG_TUNE_QUEUE_ORDER 11
/*   ZG_TUNE_LANE_ORDER: 1 (default) runs the 64-lane engine on a copy of the
 *   resource whose vertices are renumbered by decreasing degree within each
 *   component (built at the engine's first use, kept with the handle; rows
 *   keep their arc order), so the rows and lane masks the sweep gathers most
 *   share lines; 0 runs it on the resource's own ids.  Scores agree to FP64
 *   summation order (the source groups can differ). */
#define ZG_TUNE_LANE_ORDER 13
int zg_set_tuning(zg_graph *g, int param, double value);

/* Source grouping for the batched sweep (no reference counterpart; the
 * result is the same sum, only FP64 summation order changes).  mode 0 keeps
 * the caller's order; mode 1 (default) sorts the sources by their distance to
 * the graph's hub core (the highest in-degree vertices) so the 64 sources of
 * a group reach most vertices at the same BFS level; mode 2 sorts them by
 * their discovery rank in a breadth-first search from the hub core over
 * predecessor arcs, so a group holds sources that are close to each other.
 * Returns the previous mode, or <0. */
int zg_set_source_order(zg_graph *g, int mode);

/* Kernel-level profile of the most recent Brandes / PageRank call (no
 * reference counterpart; for the benchmark's roofline figures).  While
 * profiling is on, HIP events are recorded on the launch stream around every
 * kernel launch, and the per-kernel device time is summed.  PageRank counts
 * only launches that ran an iteration (not the no-op launches queued past
 * convergence).  zg_bc_f64 reports its distance rounds (relaxation and
 * far-pile splits) as ZG_PROF_BC_FORWARD,
 * its dependency rounds as ZG_PROF_BC_BACKWARD and its round bookkeeping as
 * ZG_PROF_BC_SETUP. */
#define ZG_PROF_BC_FORWARD 0
#define ZG_PROF_BC_FINALIZE 1
#define ZG_PROF_BC_BACKWARD 2
#define ZG_PROF_BC_HUB 3
#define ZG_PROF_BC_SETUP 4       /* per-batch resets, sources, final reduction */
#define ZG_PROF_PR_ITER 5
#define ZG_PROF_BC_PUSH 6
#define ZG_PROF_BC_DAG 7         /* zg_bc_f64: tight-arc counts and path-count rounds */
#define ZG_PROF_SLOTS 8
typedef struct {
    double ms[ZG_PROF_SLOTS];
    uint64_t launches[ZG_PROF_SLOTS];
} zg_profile;

/* Turns per-kernel profiling on (1) or off (0) for this handle. */
int zg_set_profiling(zg_graph *g, int on);

/* Copies the profile of the most recent call(s) since profiling was turned
 * on or last read, then clears it. */
int zg_get_profile(zg_graph *g, zg_profile *out);

/* ---- Sampled Brandes and learned pivot selection -------------------------
 * None of the entry points below has a counterpart in the reference: they
 * serve sampled betweenness (weighted sources) and the node-scoring network
 * that proposes the sampled sources. */

/* Weighted multi-source Brandes: out_bc[v] = sum_i weights[i] *
 * delta_{sources[i]}(v) over v != sources[i] (weights NULL = all 1, as
 * zg_bc_unweighted; repeated sources count repeatedly), and the per-source
 * dependency totals out_totals[i] = sum_v delta_{sources[i]}(v) =
 * sum over t reachable from sources[i], t != sources[i], of (dist - 1)
 * (every shortest s-t path has dist - 1 interior vertices).  sources NULL =
 * every live vertex in id order (weights must then be NULL).  Either output
 * may be NULL, not both.  Host buffers; ZG_ERR_NODE_NOT_FOUND as
 * zg_bc_unweighted, ZG_ERR_INVALID_ARG for a non-finite weight. */
int zg_bc_weighted(zg_graph *g, const uint32_t *sources, const double *weights, uint64_t n_sources,
                   double *out_bc, double *out_totals, zg_bc_info *info);

/* Folding of single-successor sources for sampled BC (no reference
 * counterpart).  A live vertex of class 1 has exactly one successor other
 * than itself (parallel arcs counted once), class 0 none, class 2 several.
 * A class-1 vertex whose successor has class 2 is folded into that
 * successor: its own Brandes sweep is the successor's, shifted by one arc.
 * Out (each may be NULL, [n_cap] each): class, the class-1 successor
 * (0xFFFFFFFF otherwise), the vertex each is folded into (0xFFFFFFFF when
 * not folded), and how many vertices are folded into each. */
int zg_fold_map(zg_graph *g, uint8_t *class_out, uint32_t *succ_out, uint32_t *folded_into_out,
                uint32_t *count_out);

/* Weighted multi-source Brandes over groups: each source m stands for
 * itself and the sources folded into it (zg_fold_map), so out_bc[v] =
 * sum_i weights[i] * (delta_{m_i}(v) + sum over s folded into m_i of
 * delta_s(v)), m_i itself included (the folded sources' paths all pass
 * through it) -- computed exactly from m_i's own sweep, with per-target
 * weights in its dependency pass.  With weights 1 and every live vertex
 * that is not folded as a source, out_bc equals zg_bc_unweighted over all
 * live sources.  A folded vertex in the list is ZG_ERR_INVALID_ARG (its
 * group counts it); weights NULL = all 1.  out_totals[i] is the dependency
 * total of m_i's own sweep, as zg_bc_weighted; either output may be NULL,
 * not both.  Runs on the 64-lane engine. */
int zg_bc_folded(zg_graph *g, const uint32_t *sources, const double *weights, uint64_t n_sources, double *out_bc,
                 double *out_totals, zg_bc_info *info);

/* Replaces `nif_node_degrees` (resource_graph.ex:562-599, Elixir
 * node_degrees/1 :2130-2136): out[v] = live out-arcs of v (arcs to
 * tombstoned ids skipped; 0 for a tombstoned v, whose arcs removeNode
 * deleted) for v < n_cap.  Read from the handle's host copy of the degrees. */
int zg_node_degrees(zg_graph *g, uint32_t *out);

/* Core numbers of the graph taken as undirected and simple (self-loops
 * dropped, each neighbour once), out[v] for v < n_cap (0 for tombstones),
 * as coreNumbers in priv/zog/src/connectivity.zig:5-101 (the
 * `nif_core_numbers` NIF, resource_graph.ex:622-628). */
int zg_core_numbers(zg_graph *g, uint32_t *out);

/* Node features, ZG_NFEAT floats per node into d_features[n_cap][8] (device
 * memory): log1p(in+out degree), log1p(in degree), log1p(out degree),
 * log1p(sum of the successors' out-degrees), log(n * PageRank) (damping
 * 0.85, tol 1e-6, FP32), log1p(core number), each standardised over the live
 * nodes; column 6 = 1, column 7 = 0 (rows of tombstones are 0).  One graph
 * per handle (ZG_ERR_INVALID_ARG on a batch handle of several graphs). */
#define ZG_NFEAT 8
int zg_node_features(zg_graph *g, float *d_features, void *stream);

/* Node-scoring network ZG_NFEAT -> 128 -> 128 -> 1 with ReLU hidden layers,
 * BF16 operands on the matrix cores and FP32 accumulation.  Parameters are
 * ZG_MLP_PARAMS floats: W1[8][128], b1[128], W2[128][128], b2[128],
 * w3[128], b3 (x @ W convention).  The model lives on the device that is
 * current at zg_mlp_create. */
#define ZG_MLP_HIDDEN 128
#define ZG_MLP_PARAMS 17793
#define ZG_MLP_MAX_BATCH 65536
typedef struct zg_mlp zg_mlp;
int zg_mlp_create(uint64_t seed, zg_mlp **out);
void zg_mlp_destroy(zg_mlp *m);
int zg_mlp_get_params(zg_mlp *m, float *out);
/* Sets the parameters and resets the Adam state. */
int zg_mlp_set_params(zg_mlp *m, const float *in);
/* d_y[i] = network(d_x[i]) for n rows of d_x[n][ZG_NFEAT] (device memory);
 * returns when d_y is written. */
int zg_mlp_forward(zg_mlp *m, const float *d_x, uint64_t n, float *d_y, void *stream);
/* Gradient of the mean squared error (1/n) sum (y_i - target_i)^2 over a
 * batch of n <= ZG_MLP_MAX_BATCH rows (device memory), kept for
 * zg_mlp_adam; optionally copied to grad_out[ZG_MLP_PARAMS] and the loss to
 * *loss_out (host memory).  Returns when done, as zg_mlp_adam. */
int zg_mlp_gradient(zg_mlp *m, const float *d_x, const float *d_target, uint64_t n, float *grad_out,
                    double *loss_out, void *stream);
/* One Adam step with the last gradient (bias-corrected moments). */
int zg_mlp_adam(zg_mlp *m, float lr, float beta1, float beta2, float eps, void *stream);

/* Human-readable name of a status code. */
const char *zg_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif /* ZOG_HIP_H */
