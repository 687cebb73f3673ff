// Shared definitions of the gfx950 backend: the resident graph handle, HIP
// error plumbing and wave-level helpers (64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/zog_hip.h"

#define ZG_NONE 0xFFFFFFFFu

// A vertex's predecessor / successor list is cut into chunks of at most this
// many arcs; one wavefront sweeps one chunk, so a hub is shared by several
// wavefronts (hub split).
#define ZG_CHUNK 256

// Path counts are kept as FP64 with a per-(source, level) power-of-two shift:
// once the largest binary exponent at a level passes this threshold, the
// next level's counts are stored divided by 2^exponent (exact).
#define ZG_SHIFT_T 512

#define ZG_MAX_GROUPS 64

// The finalize step runs as a fixed grid: each block owns a contiguous range
// of predecessor chunks, so it reserves list space with one atomic per level.
#define ZG_FIN_BLOCKS 256
#define ZG_FIN_THREADS 1024
// Per-(group, finalize block) statistics of the lane engine: per-lane arcs
// scanned, lanes reached, row visits (a vertex on a level of some lane), arcs
// of the visited rows.
#define ZG_BC_STATS 4
// Counters per group of the lane engine's tree-arc count.
#define ZG_TREE_SLOTS 256

// Error bits a Brandes kernel reports to the host.
#define ZG_BC_ERR_SOURCE 1u     // a source id is out of range or not live
#define ZG_BC_ERR_CAPACITY 2u   // a work list would overflow its buffer

// Returns a ZG_ERR_* code for a HIP error.
static inline int zg_hip_code(hipError_t e)
{
    if (e == hipSuccess) return ZG_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return ZG_ERR_GPU_OOM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return ZG_ERR_NO_DEVICE;
    return ZG_ERR_HIP;
}

#define ZG_TRY(expr)                                   \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return zg_hip_code(_e);  \
    } while (0)

#define ZG_TRY_RC(expr)                 \
    do {                                \
        int _rc = (expr);               \
        if (_rc != ZG_OK) return _rc;   \
    } while (0)

template <typename T>
static inline int zg_dmalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return zg_hip_code(e);
    }
    return ZG_OK;
}

template <typename T>
static inline void zg_dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// ---------------------------------------------------------------------------
// Brandes batched-sweep workspace (sized for `groups` 64-source groups).
// Layouts, per group g (lane = source slot within the group):
//   sig[g][v][lane]    FP64 stored path counts (row of 64 = 512 B per vertex)
//   cc [k][g][v][lane] FP64 (1 + delta) / sigma of finished vertices, k =
//                      level parity (a level reads k = (d+1)&1, writes d&1)
//   vis[g][v], fr[k][g][v]  u64 lane masks (visited, frontier buffers)
//   flist/fmask[g][i]       per-level successor-chunk work lists
//   hlist/hmask[g][i]       per-level lists of multi-chunk vertices
//   emax[L][g][lane]        largest exponent seen at level L (shift k_L)
struct zg_bc_ws {
    int groups = 0;
    uint32_t vloc = 0;                   // per-group vertex capacity of the state arrays
    bool with_depth = false;
    double *sig = nullptr, *cc = nullptr, *hubacc = nullptr, *bcg = nullptr;
    uint32_t *dep = nullptr;
    uint32_t *maxdep = nullptr;          // [G][V] deepest level of v over the group's lanes
    uint64_t *lvl = nullptr;             // [lmcap][G][V] lanes that reached v at each level
    uint32_t lmcap = 0;                  // levels the lane table holds (0: none)
    uint64_t *vis = nullptr, *fr = nullptr;
    unsigned long long *stats = nullptr; // [G][ZG_FIN_BLOCKS][ZG_BC_STATS] arcs, reached, row visits, row arcs
    unsigned long long *tree = nullptr;  // [G][ZG_TREE_SLOTS] per-lane tree arcs (counted on request)
    uint32_t *touched = nullptr;         // [G][V] vertices reached by a push level
    uint32_t *flist = nullptr;           // [G][fcap]
    uint64_t *fmask = nullptr;
    uint32_t *hlist = nullptr;           // [G][hcap]
    uint64_t *hmask = nullptr;
    int *emax = nullptr;                 // [lcap][G][64]
    uint32_t *foff = nullptr, *hoff = nullptr;  // [lcap][G] level offsets in flist / hlist
    uint32_t *modes = nullptr;           // [lcap][G] direction taken at each level
    uint32_t lcap = 0;
    uint32_t *src = nullptr;             // [G][64]
    double *wts = nullptr;               // [G][64] per-source weights of a weighted sweep
    unsigned long long *dsum = nullptr;  // [G][64] per-source dependency totals (integers)
    unsigned long long *rsum = nullptr, *fsum = nullptr;  // [G][64] folded sweep: reached, folded members reached
    uint32_t cmax = 0;                   // most predecessor chunks in one group's span
    void *level_args = nullptr;          // device arguments of the graph-replayed level kernels
    void *h_level_args = nullptr;        // pinned staging copy of them
    int *cur = nullptr;                  // device level counter of the replayed levels
    hipGraphExec_t graph = nullptr;      // ZG_GRAPH_LEVELS captured levels
    uint32_t graph_dims[5] = {0, 0, 0, 0, 0};  // push, pull, sparse/dense finalize grids, groups
    struct Counters {
        uint32_t fcount, hcount;
        uint32_t err;               // ZG_BC_ERR_* bits
        uint32_t tcount;            // vertices first reached by the current push level
        unsigned long long farcs[2];  // successor arcs of a level's vertices, by level parity
        uint32_t mode;              // ZG_MODE_* of the current level
        uint32_t pad;
    };
    Counters *cnt = nullptr;             // [G]
    Counters *h_cnt = nullptr;           // pinned host mirror
    size_t fcap = 0, hcap = 0;
    uint32_t last_levels = 0;            // levels used by the previous batch
    // per group: lane-table rows [0, lvl_span[g]) may hold bits (of levels
    // below last_levels); every other row of the table is zero
    uint32_t lvl_span[ZG_MAX_GROUPS] = {};
};

// Per-(source, vertex) record of the queue engine: depth, path-count
// exponent and one value, the path-count mantissa until the dependency sweep
// replaces it with the dependency term; one 16-byte line segment, so a
// neighbour's depth, exponent and value arrive in one load.
struct alignas(16) zg_qrec {
    uint32_t depth;
    uint32_t exp;
    double val;
};

// Per-source queue engine workspace (brandes_queue.hip), S sources per batch.
struct zg_bcq_ws {
    uint32_t S = 0, lcap = 0;
    zg_qrec *rec = nullptr;              // [S][V]
    uint32_t *queue = nullptr;           // [S][V]
    double *dlt = nullptr;               // [S][V] weighted dependencies of the batch
    uint32_t *lstart = nullptr;          // [lcap][S]
    uint32_t *cpre = nullptr;            // [lcap][S + 1] per-level prefix of 64-entry chunks, total last
    uint32_t *qlen = nullptr;            // [S]
    uint32_t *newcnt = nullptr;          // [lcap]
    uint32_t *src = nullptr;             // [S]
    double *wts = nullptr;               // [S] per-source weights of a weighted sweep
    double *bc = nullptr;                // [V]
    unsigned long long *stats = nullptr; // [2]
};

// Per-(source, vertex) record of the weighted engine: everything a lowering,
// a tight-arc check, a path-count step or a dependency step touches for one
// vertex sits in one 32-byte line segment.
struct alignas(32) zg_wrec {
    unsigned long long dist;  // distance bits (all ones = unreached)
    double sig;               // path count
    double cc;                // dependency term (1 + delta) / sigma
    uint32_t aux;             // relax round that queued the vertex last
    uint32_t flag;            // far-pile flag; 2 for a reached vertex once the relax rounds end (all piles
                              // empty), then 2 + its tight in-arcs not yet processed
};

// Weighted (f64 arc weight) engine workspace (brandes_weighted.hip), S
// sources per batch.
struct zg_wbc_ws {
    uint32_t S = 0, lcap = 0;
    zg_wrec *rec = nullptr;              // [S][V]
    uint32_t *fr = nullptr;              // [2][S][V]
    uint32_t *fcnt = nullptr;            // [2][S]
    uint32_t *reach = nullptr;           // [S][V]
    uint32_t *rcnt = nullptr;            // [S]
    uint32_t *order = nullptr;           // [S][V]
    uint32_t *qlen = nullptr;            // [S]
    uint32_t *lstart = nullptr;          // [lcap][S]
    uint32_t *tot = nullptr;             // [lcap]
    uint32_t *tbits = nullptr;           // [S][(E + 31) / 32 + 1] tight-arc bits
    uint32_t *far = nullptr;             // [2 piles][2 buffers][S][V] far piles
    uint32_t *farcnt = nullptr;          // [2][2][S]
    uint32_t *fsel = nullptr, *split = nullptr;  // [2][S]
    unsigned long long *thr = nullptr, *wlim = nullptr;  // [S]
    unsigned long long *fmin = nullptr;  // [2][S]
    uint32_t *hcnt = nullptr, *hub_v = nullptr, *hub_a = nullptr;  // hub pieces of a relax round
    uint32_t *src = nullptr;             // [S]
    double *bc = nullptr;                // [V]
    unsigned long long *stats = nullptr; // [3]
};

// PageRank workspace: double-buffered ranks and gathered values.
struct zg_pr_ws {
    int prec = 0;              // 32 or 64
    void *r = nullptr;         // [2][V] float or double
    void *xh = nullptr;        // [2][V]
    double *part = nullptr;    // [nblocks][2] block partial sums (l1, dangling)
    float2 *seg32 = nullptr;   // large-row segment partials (FP32 Kahan pairs)
    double *seg64 = nullptr;
    void *state = nullptr;     // device iteration state
    void *h_state = nullptr;   // pinned host mirror
};

// ZG_TUNE_PR_CACHE's default: as many values as the default block shape holds
#define ZG_PR_XCACHE_DEFAULT (1u << 30)

struct zg_graph {
    std::mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t cap_stream = nullptr;    // stream the level graph is captured on (never launched on)
    bool use_graph = true;               // replay captured level chunks (ZG_TUNE_GRAPH)
    bool count_tree = false;             // lane engine counts per-lane tree arcs (ZG_TUNE_COUNT_TREE_ARCS)
    uint32_t V = 0, n_live = 0;
    uint64_t E = 0;
    // resident CSR
    uint32_t *out_ptr = nullptr, *out_col = nullptr;
    uint32_t *in_ptr = nullptr, *in_col = nullptr;
    uint8_t *live = nullptr;
    std::vector<uint8_t> h_live;
    std::vector<uint32_t> h_out_deg;     // out-degrees (lazy, with h_coredist)
    std::vector<uint32_t> h_coredist;    // distance to the hub core (lazy)
    // folding of single-successor sources (zg_fold_map, zg_bc_folded; lazy):
    // class of each vertex (0: no successor but itself, 1: one distinct
    // successor, 2: several), that successor, the vertex it is folded into
    // (its successor when that has several, else ZG_NONE) and the number of
    // sources folded into each vertex
    bool fold_built = false;
    uint32_t *fold_succ1 = nullptr, *fold_count = nullptr;   // device [V]
    std::vector<uint8_t> h_fold_class;
    std::vector<uint32_t> h_fold_first, h_fold_succ1, h_fold_count;
    std::vector<uint32_t> h_bfsrank;     // discovery rank in a search from the hub core (lazy)
    int source_order = 1;
    double push_alpha = 8.0;
    uint32_t n_dangling = 0;
    // disjoint components given at creation (zg_graph_create_batch): vertex
    // offsets, n_comp + 1 entries; a plain graph is one component [0, V)
    std::vector<uint32_t> h_comp_off;
    // work tables (graph_derive.hip), built on the device at load
    // predecessor chunks (forward sweep work items)
    uint32_t n_ich = 0;
    uint32_t *ich_first = nullptr;       // [V+1] first predecessor chunk of each vertex
    uint32_t *ich_v = nullptr, *ich_e0 = nullptr, *ich_e1 = nullptr;
    // successor chunks (backward sweep work items)
    uint32_t n_och = 0, max_och_per_v = 0;
    uint32_t *och_first = nullptr, *och_v = nullptr, *och_e0 = nullptr, *och_e1 = nullptr;
    uint32_t n_ohub = 0;                 // partial-sum slots of the multi-chunk vertices' chunks
    uint32_t *ohub_idx = nullptr;        // [V] first partial-sum slot of a multi-chunk vertex, or ZG_NONE
    // at each component offset h_comp_off[i]: predecessor CSR offset, first
    // predecessor chunk, first successor chunk, multi-chunk vertices before it
    std::vector<uint32_t> h_cb_in, h_cb_ich, h_cb_och, h_cb_hub;
    // per graph c: 1 + its last vertex with an arc (its first vertex if none)
    std::vector<uint32_t> h_comp_act;
    // PageRank row bins
    uint32_t n_small = 0, n_med = 0, n_large = 0, n_seg = 0;
    uint32_t n_m4 = 0, n_m16 = 0, n_m64 = 0;   // n_med = n_m4 + n_m16 + n_m64
    uint32_t *pr_small = nullptr, *pr_med = nullptr;   // pr_med = [m4 | m16 | m64]
    uint4 *pr_desc = nullptr;            // [small | m4 | m16 | m64] rows: {v, first arc, in-deg, out-deg}
    uint32_t *pr_large = nullptr;        // [n_large] rows (PageRank ids)
    uint32_t *pr_large_deg = nullptr;    // [n_large] out-degrees of the large rows
    uint32_t *pr_id = nullptr;           // [V] PageRank id of each vertex (by decreasing out-degree)
    uint32_t *pr_col = nullptr;          // [E] predecessor CSR columns as PageRank ids
    uint32_t *pr_large_seg0 = nullptr;   // [n_large + 1] first segment of each large row
    uint32_t *pr_seg_row = nullptr;      // [n_seg] large-row index of each segment
    uint32_t *pr_seg_e0 = nullptr, *pr_seg_e1 = nullptr;
    // long rows dealt to pr_nlong iteration blocks by arc count (largest
    // first, each to the least loaded block): slot b sums the rows
    // pr_long_plan[pr_long_off[b] .. pr_long_off[b+1]), entries {row, first
    // arc, end arc, 0}
    uint32_t pr_nlong = 0;
    uint4 *pr_long_plan = nullptr;
    uint32_t *pr_long_off = nullptr;
    int groups = 4;
    int engine = 0;                      // ZG_ENGINE_*
    double *out_w = nullptr;             // [E] arc weights in successor-CSR order (weighted resources)
    uint4 *qadj = nullptr;               // [V][4] inline adjacency of the queue engine (on demand)
    // locality order of the queue engine (ZG_TUNE_QUEUE_ORDER; on demand):
    // the internal id of each vertex and the CSR under those ids, every row
    // in its arc order
    bool queue_order = true;
    uint32_t *qperm = nullptr;           // [V]
    uint32_t *q_out_ptr = nullptr, *q_out_col = nullptr, *q_in_ptr = nullptr, *q_in_col = nullptr;
    bool weights_ok = true;              // every live arc's weight is finite and non-negative
    bool built_on_gpu = false;           // the CSR came from zg_build_csr_gpu
    uint32_t wbc_batch = 0;              // weighted sweep: sources per batch (0 = default)
    double wbc_delta = 0.0;              // weighted sweep: bound step of the near frontier (0 = default)
    uint32_t wbc_grid = 0;               // weighted sweep: workgroups per round kernel (0 = default)
    double wmean = 0.0;                  // mean weight of the live arcs
    uint32_t *und_ptr = nullptr, *und_col = nullptr;  // simple undirected CSR (core numbers), built on demand
    uint32_t pull_grid = 4096;
    // the 64-lane engine's view (ZG_TUNE_LANE_ORDER; built on first use): a
    // copy of this resource with every component's vertices renumbered by
    // decreasing degree, rows in their arc order; lane_perm[v] = the view's
    // id of v
    bool lane_order = true;
    zg_graph *lane = nullptr;
    uint32_t *lane_perm = nullptr;       // device [V]
    std::vector<uint32_t> h_lane_perm;
    uint32_t pr_xcache = ZG_PR_XCACHE_DEFAULT;  // PageRank values staged in LDS (0 = none)
    bool pr_xcache_useful = false;       // the cached values carry a useful share of the gathers
    uint32_t pr_grid_cap = 0;            // PageRank iteration blocks (0 = default)
    uint32_t pr_block = 0;               // PageRank iteration block threads: 256, 512, 1024 (0 = default)
    uint32_t dev_cus = 0;                // compute units of the device (lazy)
    uint32_t dev_lds = 0;                // LDS bytes one block may use (lazy)
    bool high_diameter = false;          // set with h_coredist
    // device staging of the host-buffer entry points (sources in, scores
    // out), kept between calls so a call allocates nothing once warm
    double *io_out = nullptr;            // [V]
    uint32_t *io_src = nullptr;          // [io_src_cap]
    uint64_t io_src_cap = 0;
    zg_bc_ws bc;
    zg_bcq_ws bcq;
    zg_wbc_ws wbc;
    zg_pr_ws pr;
    // per-kernel profiling (events around launches, summed on read)
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_marks;   // (slot, index of the start event)
    zg_profile prof{};
};

// Profiling helpers: zg_prof_begin returns the index of the start event (or
// SIZE_MAX when off); zg_prof_end records the end event and remembers the
// pair; zg_prof_collect waits for the stream and folds pairs into g->prof.
size_t zg_prof_begin(zg_graph *g, hipStream_t s);
void zg_prof_end(zg_graph *g, hipStream_t s, int slot, size_t start);
int zg_prof_collect(zg_graph *g, hipStream_t s);
void zg_prof_release(zg_graph *g);

#define ZG_PROF_LAUNCH(slot, ...)                    \
    do {                                             \
        size_t _zp = zg_prof_begin(g, s);            \
        hipLaunchKernelGGL(__VA_ARGS__);             \
        zg_prof_end(g, s, (slot), _zp);              \
    } while (0)

int zg_build_csr_gpu(zg_graph *g, const uint32_t *from, const uint32_t *to, const double *weight, uint64_t m,
                     bool device_arcs = false);
// The 64-lane engine's degree-ordered view of g (g->lane, g->lane_perm,
// g->h_lane_perm), built once; leaves g->lane null when ZG_TUNE_LANE_ORDER
// is off or the device has no room for it.
int zg_lane_view(zg_graph *g, hipStream_t s);
// Chunk lists, hub slots and PageRank row bins from the device CSR and
// PageRank ids (graph_derive.hip).
int zg_derive_tables(zg_graph *g);
// The long-row plan of the PageRank iteration (pagerank.hip) from the long
// rows' first arcs and lengths, on the host (n_large entries).
int zg_pr_long_plan(zg_graph *g, const uint32_t *h_e0, const uint32_t *h_len, hipStream_t s);
// Index i of the first component offset h_comp_off[i] >= v: v's own index
// when v is 0, V or a component's first vertex, else the index of the end
// of the component holding v.
static inline size_t zg_comp_at(const zg_graph *g, uint32_t v)
{
    return (size_t)(std::lower_bound(g->h_comp_off.begin(), g->h_comp_off.end(), v) - g->h_comp_off.begin());
}
// The handle's staging buffers: *out = [V] doubles, *src = [n] u32 (grown
// as needed); the caller holds the handle's lock.
int zg_io_out(zg_graph *g, double **out);
int zg_io_src(zg_graph *g, uint64_t n, uint32_t **src);
int zg_bc_ws_release(zg_graph *g);
int zg_bcq_ws_release(zg_graph *g);
void zg_bcq_order_release(zg_graph *g);  // the queue engine's locality order and inline adjacency
int zg_wbc_ws_release(zg_graph *g);
int zg_pr_ws_release(zg_graph *g);
int zg_bcq_sweep(zg_graph *g, const uint32_t *d_sources, uint64_t n, const double *d_weights, double *h_totals,
                 double *d_out, hipStream_t s, zg_bc_info *inf);

// ---------------------------------------------------------------------------
// Device helpers
#define ZG_DEV __device__ __forceinline__

ZG_DEV uint64_t zg_readlane64(uint64_t x, int j)
{
    uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)x, j);
    uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

ZG_DEV uint32_t zg_readlane32(uint32_t x, int j)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)x, j);
}

// Inclusive prefix sum over the 64 lanes of a wavefront.
ZG_DEV uint32_t zg_wave_incl_scan(uint32_t x, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

ZG_DEV double zg_wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

ZG_DEV unsigned long long zg_wave_sum_u64(unsigned long long x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
