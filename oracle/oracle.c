/*
 * CPU oracle for Zog's `:soa` Brandes betweenness and PageRank.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is
 * compared against and the CPU baseline that bench.py times beside it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it;
 * the product library never links or calls it.
 *
 * It restates, in plain C, the algorithms of the reference's Zig code:
 *   - ArrayGraph storage and successor order: priv/zog/src/models/array_graph.zig
 *     addNode :66-75, addEdge :88-108 (prepend to the per-node linked list),
 *     SuccessorIterator.next :166-180 (skip deleted edges and deleted
 *     targets; parallel arcs all yielded), nodeIds :268-284, hasNode :210-213.
 *   - buildGraph in lib/zog/resource_graph.ex:120-129 (all nodes first, then
 *     addEdge in array order; an out-of-range id is NodeNotFound).
 *   - betweennessUnweighted ArrayGraph branch, priv/zog/src/centrality.zig:694-798.
 *   - pagerank ArrayGraph branch, priv/zog/src/centrality.zig:871-965 with
 *     PageRankOptions :16-20.
 *   - CentralityResult.get default 0.0 (centrality.zig:31-33) and
 *     extractScores (resource_graph.ex:166-174): dense f64[V], 0.0 for
 *     non-live ids.
 *
 * Extensions the reference does not have (each documented where defined):
 *   - a source subset for Brandes (sampled/batched sources),
 *   - a "scaled sigma" mode that stores path counts as FP64 with a per-level
 *     power-of-two shift, so counts beyond 2^64 (and beyond the FP64 range)
 *     stay usable; it reduces to the reference's formula whenever no shift
 *     is taken,
 *   - an "exponent" mode that gives every count its own binary exponent, for
 *     graphs whose counts on one level span more than the FP64 range,
 *   - per-source export of BFS depths and path counts,
 *   - an N-thread source-chunked BC (the pattern of centrality.zig:216-256).
 *   - core numbers, coreNumbers in priv/zog/src/connectivity.zig:5-101
 *     (input to the learned pivot scorer's features).
 *   - weighted Brandes, betweenness (centrality.zig:529-667) as
 *     betweennessF64 instantiates it (:1355-1357: f64 weights, 0.0, a + b,
 *     std.math.order).
 *   - eigenvector (centrality.zig:1050-1201), Katz (:1212-1267) and alpha
 *     centrality (:1278-1333) over collectNodes / buildInNeighbors
 *     (utils.zig:42-73).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ZO_NONE 0xFFFFFFFFu

enum {
    ZO_OK = 0,
    ZO_ERR_NODE_NOT_FOUND = -2,
    ZO_ERR_OOM = -3,
    ZO_ERR_ARG = -4,
};

/* ArrayGraph(void, f64) restated: node SoA (first_edge, is_deleted) and edge
 * SoA (to, next_edge, is_deleted).  Edge data (the f64 weight) is unused by
 * both paths and is not stored. */
typedef struct {
    uint32_t n_cap;
    uint64_t m;
    uint32_t *first_edge;   /* ZO_NONE = empty list */
    uint8_t *node_deleted;
    uint32_t *to;
    uint32_t *next_edge;    /* ZO_NONE = end of list */
    uint8_t *edge_deleted;
    double *weight;         /* edge data (f64); NULL until zo_graph_set_weights */
} zo_graph;

void zo_graph_destroy(zo_graph *g)
{
    if (!g) return;
    free(g->first_edge);
    free(g->node_deleted);
    free(g->to);
    free(g->next_edge);
    free(g->edge_deleted);
    free(g->weight);
    free(g);
}

static int zo_has_node(const zo_graph *g, uint32_t id)
{
    return id < g->n_cap && !g->node_deleted[id];
}

/* buildGraph (resource_graph.ex:120-129) + addEdge (array_graph.zig:88-108).
 * live_mask (nullable) marks tombstoned nodes; the reference reaches that
 * state through removeNode (array_graph.zig:133-157), which also tombstones
 * every arc touching the node, so arcs at a dead endpoint are kept but marked
 * deleted here. */
int zo_graph_create(uint32_t n_cap, const uint8_t *live_mask, const uint32_t *from,
                    const uint32_t *to, uint64_t m, zo_graph **out)
{
    *out = NULL;
    zo_graph *g = (zo_graph *)calloc(1, sizeof(zo_graph));
    if (!g) return ZO_ERR_OOM;
    g->n_cap = n_cap;
    g->m = m;
    g->first_edge = (uint32_t *)malloc(sizeof(uint32_t) * (n_cap ? n_cap : 1));
    g->node_deleted = (uint8_t *)calloc(n_cap ? n_cap : 1, 1);
    g->to = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    g->next_edge = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    g->edge_deleted = (uint8_t *)calloc(m ? m : 1, 1);
    if (!g->first_edge || !g->node_deleted || !g->to || !g->next_edge || !g->edge_deleted) {
        zo_graph_destroy(g);
        return ZO_ERR_OOM;
    }
    for (uint32_t i = 0; i < n_cap; i++) g->first_edge[i] = ZO_NONE;
    for (uint64_t e = 0; e < m; e++) {
        uint32_t f = from[e], t = to[e];
        if (f >= n_cap || t >= n_cap) {
            zo_graph_destroy(g);
            return ZO_ERR_NODE_NOT_FOUND;
        }
        g->to[e] = t;
        g->next_edge[e] = g->first_edge[f];
        g->first_edge[f] = (uint32_t)e;
    }
    if (live_mask) {
        for (uint32_t i = 0; i < n_cap; i++) g->node_deleted[i] = live_mask[i] ? 0 : 1;
        for (uint64_t e = 0; e < m; e++)
            if (g->node_deleted[from[e]] || g->node_deleted[to[e]]) g->edge_deleted[e] = 1;
    }
    *out = g;
    return ZO_OK;
}

/* SuccessorIterator.next (array_graph.zig:166-180): advance `*cur` to the next
 * live arc; returns its target or ZO_NONE at the end of the list. */
static inline uint32_t zo_next_succ(const zo_graph *g, uint32_t *cur)
{
    while (*cur != ZO_NONE) {
        uint32_t e = *cur;
        *cur = g->next_edge[e];
        if (!g->edge_deleted[e]) {
            uint32_t w = g->to[e];
            if (!g->node_deleted[w]) return w;
        }
    }
    return ZO_NONE;
}

uint32_t zo_node_count(const zo_graph *g)
{
    uint32_t n = 0;
    for (uint32_t i = 0; i < g->n_cap; i++) n += g->node_deleted[i] ? 0 : 1;
    return n;
}

/* Successor list of v in the reference's iteration order (reverse insertion),
 * written to out (capacity cap).  Returns the full count. */
uint64_t zo_successors(const zo_graph *g, uint32_t v, uint32_t *out, uint64_t cap)
{
    uint64_t k = 0;
    uint32_t cur = g->first_edge[v];
    uint32_t w;
    while ((w = zo_next_succ(g, &cur)) != ZO_NONE) {
        if (k < cap) out[k] = w;
        k++;
    }
    return k;
}

/* ------------------------------------------------------------------------- */
/* Brandes, unweighted (centrality.zig:694-798)                              */
/* ------------------------------------------------------------------------- */

/* Path-count modes.  ZO_SIGMA_U64 is the reference (`sigma: []usize`,
 * centrality.zig:701, incremented at :763, converted to f64 at :775-779);
 * an overflow wraps as in the ReleaseFast build and is reported.
 * ZO_SIGMA_SCALED keeps sigma as FP64 with a per-(source, level) exponent
 * shift k_L: values at level L+1 are stored as 2^-k_L * sum(level-L values),
 * k_L = e_L if e_L > ZO_SHIFT_T else 0, where e_L is the largest binary
 * exponent (ilogb) among level-L values.  The true count is the stored value
 * times 2^(k_0 + ... + k_{L-1}).  In the backward sweep sigma_u/sigma_w for an
 * arc u(level L)->w(level L+1) is therefore (s'_u / s'_w) * 2^-k_L.  While no
 * value exceeds 2^ZO_SHIFT_T every k is 0 and the arithmetic is exactly the
 * reference's (on FP64 counts). */
#define ZO_SIGMA_U64 0
#define ZO_SIGMA_SCALED 1
#define ZO_SIGMA_EXP 2
#define ZO_SHIFT_T 512
/* ZO_SIGMA_EXP keeps every count as a mantissa in [1, 2) (sigma_f) and its
 * own binary exponent (sexp): sigma = sigma_f * 2^sexp.  Counts are added in
 * the reference's order (:763), aligned to the larger exponent and
 * renormalised; below 2^53 this is exact.  sigma_u/sigma_w in the backward
 * sweep is (sigma_f[u] / sigma_f[w]) * 2^(sexp[u] - sexp[w]).  Unlike the
 * per-level shift, it represents levels whose counts span more than the
 * FP64 range (a 2048 x 2048 grid: 1 to 2^4000 on one level). */

typedef struct {
    /* per-source workspaces (centrality.zig:697-726) */
    uint32_t *dist;
    uint64_t *sigma_u;
    double *sigma_f;
    double *delta;
    uint32_t *queue;    /* BFS queue; queue order is the stack order */
    /* preds[w] as a singly linked list in append order over an arena */
    uint32_t *pred_head, *pred_tail, *pred_node, *pred_next;
    uint64_t pred_cap;
    int32_t *shift;     /* k_L per level (scaled mode) */
    int32_t *sexp;      /* per-vertex exponent (exponent mode) */
} zo_ws;

static void zo_ws_free(zo_ws *w)
{
    free(w->dist); free(w->sigma_u); free(w->sigma_f); free(w->delta); free(w->queue);
    free(w->pred_head); free(w->pred_tail); free(w->pred_node); free(w->pred_next);
    free(w->shift);
    free(w->sexp);
}

static int zo_ws_init(zo_ws *w, const zo_graph *g)
{
    memset(w, 0, sizeof(*w));
    size_t V = g->n_cap ? g->n_cap : 1;
    w->dist = malloc(sizeof(uint32_t) * V);
    w->sigma_u = malloc(sizeof(uint64_t) * V);
    w->sigma_f = malloc(sizeof(double) * V);
    w->delta = malloc(sizeof(double) * V);
    w->queue = malloc(sizeof(uint32_t) * V);
    w->pred_head = malloc(sizeof(uint32_t) * V);
    w->pred_tail = malloc(sizeof(uint32_t) * V);
    w->pred_cap = g->m ? g->m : 1;   /* at most one pred entry per arc */
    w->pred_node = malloc(sizeof(uint32_t) * w->pred_cap);
    w->pred_next = malloc(sizeof(uint32_t) * w->pred_cap);
    w->shift = malloc(sizeof(int32_t) * (V + 1));
    w->sexp = malloc(sizeof(int32_t) * V);
    if (!w->dist || !w->sigma_u || !w->sigma_f || !w->delta || !w->queue || !w->pred_head ||
        !w->pred_tail || !w->pred_node || !w->pred_next || !w->shift || !w->sexp) {
        zo_ws_free(w);
        return ZO_ERR_OOM;
    }
    return ZO_OK;
}

typedef struct {
    uint64_t arcs_scanned;  /* m_s: arcs yielded by successors() in the forward sweep */
    uint32_t reached;       /* n_s: vertices dequeued */
    uint32_t levels;        /* eccentricity + 1 */
    int overflow;           /* u64 path count wrapped */
} zo_src_stats;

/* One source of the ArrayGraph branch: forward BFS (:745-767) then the
 * dependency accumulation (:769-787), adding delta into bc.  Returns stats. */
static zo_src_stats zo_brandes_one(const zo_graph *g, zo_ws *w, uint32_t s, int mode, double *bc)
{
    zo_src_stats st = {0, 0, 0, 0};
    uint32_t V = g->n_cap;
    /* workspace reset (:730-738) */
    for (uint32_t i = 0; i < V; i++) {
        w->dist[i] = ZO_NONE;
        w->sigma_u[i] = 0;
        w->sigma_f[i] = 0.0;
        w->delta[i] = 0.0;
        w->pred_head[i] = ZO_NONE;
        w->pred_tail[i] = ZO_NONE;
        w->sexp[i] = 0;
    }
    uint64_t npred = 0;
    uint32_t qlen = 0, head = 0;
    w->dist[s] = 0;
    w->sigma_u[s] = 1;
    w->sigma_f[s] = 1.0;
    w->queue[qlen++] = s;

    uint32_t cur_level = ZO_NONE;
    int32_t k_cur = 0;           /* shift applied to values of level cur_level+1 */

    while (head < qlen) {
        uint32_t v = w->queue[head];
        uint32_t d_v = w->dist[v];
        if (mode == ZO_SIGMA_SCALED && d_v != cur_level) {
            /* entering level d_v: its sigma values are complete (all of level
             * d_v - 1 has been dequeued); find the shift for level d_v + 1 */
            cur_level = d_v;
            int emax = -100000;
            for (uint32_t i = head; i < qlen && w->dist[w->queue[i]] == d_v; i++) {
                int e = ilogb(w->sigma_f[w->queue[i]]);
                if (e > emax) emax = e;
            }
            k_cur = emax > ZO_SHIFT_T ? emax : 0;
            w->shift[d_v] = k_cur;
        }
        head++;
        uint32_t cur = g->first_edge[v];
        uint32_t t;
        while ((t = zo_next_succ(g, &cur)) != ZO_NONE) {
            st.arcs_scanned++;
            if (w->dist[t] == ZO_NONE) {
                w->dist[t] = d_v + 1;
                w->queue[qlen++] = t;
            }
            if (w->dist[t] == d_v + 1) {
                uint64_t sum;
                if (__builtin_add_overflow(w->sigma_u[t], w->sigma_u[v], &sum)) st.overflow = 1;
                w->sigma_u[t] = sum;
                if (mode == ZO_SIGMA_EXP) {
                    /* (f_t, e_t) += (f_v, e_v), then normalise f_t to [1, 2) */
                    double f = w->sigma_f[t];
                    int32_t e = w->sexp[t];
                    if (f == 0.0) {
                        f = w->sigma_f[v];
                        e = w->sexp[v];
                    } else {
                        const int32_t E = e > w->sexp[v] ? e : w->sexp[v];
                        f = ldexp(f, e - E) + ldexp(w->sigma_f[v], w->sexp[v] - E);
                        e = E;
                    }
                    const int k = ilogb(f);
                    w->sigma_f[t] = ldexp(f, -k);
                    w->sexp[t] = e + k;
                } else {
                    w->sigma_f[t] += ldexp(w->sigma_f[v], -k_cur);
                }
                /* preds[t].append(v) */
                uint32_t p = (uint32_t)npred++;
                w->pred_node[p] = v;
                w->pred_next[p] = ZO_NONE;
                if (w->pred_tail[t] == ZO_NONE) w->pred_head[t] = p;
                else w->pred_next[w->pred_tail[t]] = p;
                w->pred_tail[t] = p;
            }
        }
    }
    st.reached = qlen;
    st.levels = qlen ? w->dist[w->queue[qlen - 1]] + 1 : 0;

    /* dependency accumulation: stack = queue order, popped in reverse */
    for (uint32_t i = qlen; i > 0; i--) {
        uint32_t v = w->queue[i - 1];
        double delta_v = w->delta[v];
        if (mode == ZO_SIGMA_U64) {
            double sigma_v = (double)w->sigma_u[v];
            for (uint32_t p = w->pred_head[v]; p != ZO_NONE; p = w->pred_next[p]) {
                uint32_t u = w->pred_node[p];
                double c = ((double)w->sigma_u[u] / sigma_v) * (1.0 + delta_v);
                w->delta[u] += c;
            }
        } else if (mode == ZO_SIGMA_EXP) {
            for (uint32_t p = w->pred_head[v]; p != ZO_NONE; p = w->pred_next[p]) {
                uint32_t u = w->pred_node[p];
                double ratio = ldexp(w->sigma_f[u] / w->sigma_f[v], w->sexp[u] - w->sexp[v]);
                w->delta[u] += ratio * (1.0 + delta_v);
            }
        } else {
            double sigma_v = w->sigma_f[v];
            for (uint32_t p = w->pred_head[v]; p != ZO_NONE; p = w->pred_next[p]) {
                uint32_t u = w->pred_node[p];
                /* u sits one level above v: ratio carries 2^-k_{dist(u)} */
                double ratio = ldexp(w->sigma_f[u] / sigma_v, -w->shift[w->dist[u]]);
                w->delta[u] += ratio * (1.0 + delta_v);
            }
        }
        if (v != s) bc[v] += delta_v;
    }
    return st;
}

typedef struct {
    const zo_graph *g;
    const uint32_t *sources;
    uint64_t begin, end;
    int mode;
    double *bc;            /* private accumulator, length n_cap */
    uint64_t *arcs_per_source; /* nullable, indexed by source position */
    int overflow;
    int err;
} zo_bc_job;

static void *zo_bc_worker(void *arg)
{
    zo_bc_job *j = (zo_bc_job *)arg;
    zo_ws ws;
    if (zo_ws_init(&ws, j->g) != ZO_OK) {
        j->err = ZO_ERR_OOM;
        return NULL;
    }
    for (uint64_t i = j->begin; i < j->end; i++) {
        zo_src_stats st = zo_brandes_one(j->g, &ws, j->sources[i], j->mode, j->bc);
        if (j->arcs_per_source) j->arcs_per_source[i] = st.arcs_scanned;
        j->overflow |= st.overflow;
    }
    zo_ws_free(&ws);
    return NULL;
}

/* Betweenness over a list of sources (NULL = every live node in id order,
 * centrality.zig:728-729), unnormalised, into out_bc[n_cap] (0.0 for non-live
 * ids).  n_threads > 1 splits the source list into contiguous chunks, each
 * with its own workspace and accumulator; the chunks are added in chunk
 * order, so the result is deterministic for a given thread count.
 * Returns 0, or 1 if a u64 path count wrapped (mode ZO_SIGMA_U64), or <0. */
int zo_bc_unweighted(const zo_graph *g, const uint32_t *sources, uint64_t n_sources, int mode,
                     int n_threads, double *out_bc, uint64_t *arcs_per_source)
{
    uint32_t V = g->n_cap;
    uint32_t *all = NULL;
    if (!sources) {
        all = malloc(sizeof(uint32_t) * (V ? V : 1));
        if (!all) return ZO_ERR_OOM;
        n_sources = 0;
        for (uint32_t i = 0; i < V; i++)
            if (!g->node_deleted[i]) all[n_sources++] = i;
        sources = all;
    } else {
        for (uint64_t i = 0; i < n_sources; i++)
            if (!zo_has_node(g, sources[i])) {
                free(all);
                return ZO_ERR_NODE_NOT_FOUND;
            }
    }
    for (uint32_t i = 0; i < V; i++) out_bc[i] = 0.0;
    if (n_threads < 1) n_threads = 1;
    if ((uint64_t)n_threads > n_sources) n_threads = n_sources ? (int)n_sources : 1;

    zo_bc_job *jobs = calloc((size_t)n_threads, sizeof(zo_bc_job));
    pthread_t *th = calloc((size_t)n_threads, sizeof(pthread_t));
    int rc = ZO_OK;
    if (!jobs || !th) rc = ZO_ERR_OOM;
    uint64_t chunk = n_threads ? (n_sources + n_threads - 1) / n_threads : 0;
    for (int t = 0; t < n_threads && rc == ZO_OK; t++) {
        jobs[t].g = g;
        jobs[t].sources = sources;
        jobs[t].begin = (uint64_t)t * chunk;
        jobs[t].end = jobs[t].begin + chunk > n_sources ? n_sources : jobs[t].begin + chunk;
        if (jobs[t].begin > jobs[t].end) jobs[t].begin = jobs[t].end;
        jobs[t].mode = mode;
        jobs[t].arcs_per_source = arcs_per_source;
        jobs[t].bc = (n_threads == 1) ? out_bc : calloc(V ? V : 1, sizeof(double));
        if (!jobs[t].bc) rc = ZO_ERR_OOM;
    }
    if (rc == ZO_OK) {
        if (n_threads == 1) {
            zo_bc_worker(&jobs[0]);
        } else {
            for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, zo_bc_worker, &jobs[t]);
            for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
            for (int t = 0; t < n_threads; t++)
                for (uint32_t i = 0; i < V; i++) out_bc[i] += jobs[t].bc[i];
        }
        for (int t = 0; t < n_threads; t++) {
            if (jobs[t].err) rc = jobs[t].err;
            if (jobs[t].overflow && rc == ZO_OK) rc = 1;
        }
    }
    if (jobs && n_threads > 1)
        for (int t = 0; t < n_threads; t++) free(jobs[t].bc);
    free(jobs);
    free(th);
    free(all);
    return rc;
}

/* Per-source export of the exponent mode: depth, count mantissa in [1, 2)
 * and exponent per vertex (0 / 0 when unreached). */
int zo_bfs_paths_exp(const zo_graph *g, uint32_t s, uint32_t *depth_out, double *mant_out, int32_t *exp_out)
{
    if (!zo_has_node(g, s)) return ZO_ERR_NODE_NOT_FOUND;
    zo_ws ws;
    if (zo_ws_init(&ws, g) != ZO_OK) return ZO_ERR_OOM;
    double *scratch = calloc(g->n_cap ? g->n_cap : 1, sizeof(double));
    if (!scratch) {
        zo_ws_free(&ws);
        return ZO_ERR_OOM;
    }
    zo_brandes_one(g, &ws, s, ZO_SIGMA_EXP, scratch);
    for (uint32_t i = 0; i < g->n_cap; i++) {
        if (depth_out) depth_out[i] = ws.dist[i];
        if (mant_out) mant_out[i] = ws.sigma_f[i];
        if (exp_out) exp_out[i] = ws.sexp[i];
    }
    free(scratch);
    zo_ws_free(&ws);
    return ZO_OK;
}

/* Per-source BFS export: depth (ZO_NONE if unreached), path counts as u64
 * (reference type, wraps) and as scaled FP64 with the per-level shifts
 * (shift_out[L] for L < levels; nullable), plus stats. */
int zo_bfs_paths(const zo_graph *g, uint32_t s, uint32_t *depth_out, uint64_t *sigma_u_out,
                 double *sigma_f_out, int32_t *shift_out, uint64_t *arcs_scanned,
                 uint32_t *levels, int *overflow)
{
    if (!zo_has_node(g, s)) return ZO_ERR_NODE_NOT_FOUND;
    zo_ws ws;
    if (zo_ws_init(&ws, g) != ZO_OK) return ZO_ERR_OOM;
    double *scratch = calloc(g->n_cap ? g->n_cap : 1, sizeof(double));
    if (!scratch) {
        zo_ws_free(&ws);
        return ZO_ERR_OOM;
    }
    zo_src_stats st = zo_brandes_one(g, &ws, s, ZO_SIGMA_SCALED, scratch);
    for (uint32_t i = 0; i < g->n_cap; i++) {
        if (depth_out) depth_out[i] = ws.dist[i];
        if (sigma_u_out) sigma_u_out[i] = ws.sigma_u[i];
        if (sigma_f_out) sigma_f_out[i] = ws.sigma_f[i];
    }
    if (shift_out)
        for (uint32_t L = 0; L < st.levels; L++) shift_out[L] = ws.shift[L];
    if (arcs_scanned) *arcs_scanned = st.arcs_scanned;
    if (levels) *levels = st.levels;
    if (overflow) *overflow = st.overflow;
    free(scratch);
    zo_ws_free(&ws);
    return ZO_OK;
}

/* ------------------------------------------------------------------------- */
/* PageRank (centrality.zig:871-965)                                          */
/* ------------------------------------------------------------------------- */

/* Row range of one thread of the row-parallel mode: every row's sum is the
 * same loop in the same order as the 1-thread mode, so the new ranks are
 * bit-identical whatever the thread count. */
typedef struct {
    const uint32_t *nodes;
    uint32_t k0, k1;
    const uint64_t *in_ptr, *out_deg;
    const uint32_t *in_col;
    const double *ranks;
    double *new_ranks;
    double damping, n_f, sink_sum;
} zo_pr_rows;

static void zo_pr_rows_run(const zo_pr_rows *j)
{
    for (uint32_t k = j->k0; k < j->k1; k++) {
        uint32_t v = j->nodes[k];
        double rank_sum = 0.0;
        for (uint64_t e = j->in_ptr[v]; e < j->in_ptr[v + 1]; e++) {
            uint32_t u = j->in_col[e];
            uint64_t du = j->out_deg[u];
            if (du > 0) rank_sum += j->ranks[u] / (double)du;
        }
        j->new_ranks[v] = (1.0 - j->damping) / j->n_f + j->damping * (j->sink_sum + rank_sum);
    }
}

static void *zo_pr_rows_worker(void *arg)
{
    zo_pr_rows_run((const zo_pr_rows *)arg);
    return NULL;
}

/* Returns 0 and fills out[n_cap] (0.0 for non-live), *iters_out = iterations
 * performed, *l1_out = L1 change of the last iteration (0 if none).
 * n_threads > 1 computes the rows of an iteration on that many threads
 * (contiguous row ranges); the dangling sum and the L1 change stay one
 * serial loop in node order, so the result equals the 1-thread result bit
 * for bit (an extension: the reference's loop is single-threaded). */
int zo_pagerank_mt(const zo_graph *g, double damping, uint64_t max_iterations, double tolerance,
                   int n_threads, double *out, uint64_t *iters_out, double *l1_out);

int zo_pagerank(const zo_graph *g, double damping, uint64_t max_iterations, double tolerance,
                double *out, uint64_t *iters_out, double *l1_out)
{
    return zo_pagerank_mt(g, damping, max_iterations, tolerance, 1, out, iters_out, l1_out);
}

int zo_pagerank_mt(const zo_graph *g, double damping, uint64_t max_iterations, double tolerance,
                   int n_threads, double *out, uint64_t *iters_out, double *l1_out)
{
    uint32_t V = g->n_cap;
    for (uint32_t i = 0; i < V; i++) out[i] = 0.0;
    if (iters_out) *iters_out = 0;
    if (l1_out) *l1_out = 0.0;
    /* collectNodes (utils.zig:42-49): live ids ascending */
    uint32_t *nodes = malloc(sizeof(uint32_t) * (V ? V : 1));
    if (!nodes) return ZO_ERR_OOM;
    uint32_t n = 0;
    for (uint32_t i = 0; i < V; i++)
        if (!g->node_deleted[i]) nodes[n++] = i;
    if (n == 0) {
        free(nodes);
        return ZO_OK;
    }
    if (n == 1) {
        out[nodes[0]] = 1.0;
        free(nodes);
        return ZO_OK;
    }
    /* in_neighbors (:881-899): for each live `from` in id order, for each
     * successor, append `from` to in[to].  Stored as CSR with a counting pass
     * so each list keeps the append order. */
    uint64_t *in_ptr = calloc((size_t)V + 1, sizeof(uint64_t));
    uint64_t *out_deg = calloc(V, sizeof(uint64_t));
    if (!in_ptr || !out_deg) {
        free(nodes); free(in_ptr); free(out_deg);
        return ZO_ERR_OOM;
    }
    for (uint32_t k = 0; k < n; k++) {
        uint32_t v = nodes[k];
        uint32_t cur = g->first_edge[v], t;
        while ((t = zo_next_succ(g, &cur)) != ZO_NONE) {
            in_ptr[t + 1]++;
            out_deg[v]++;   /* out_degrees (:901-910) */
        }
    }
    for (uint32_t i = 0; i < V; i++) in_ptr[i + 1] += in_ptr[i];
    uint64_t E = in_ptr[V];
    uint32_t *in_col = malloc(sizeof(uint32_t) * (E ? E : 1));
    uint64_t *fill = malloc(sizeof(uint64_t) * V);
    double *ranks = calloc(V, sizeof(double));
    double *new_ranks = calloc(V, sizeof(double));
    if (!in_col || !fill || !ranks || !new_ranks) {
        free(nodes); free(in_ptr); free(out_deg); free(in_col); free(fill); free(ranks); free(new_ranks);
        return ZO_ERR_OOM;
    }
    memcpy(fill, in_ptr, sizeof(uint64_t) * V);
    for (uint32_t k = 0; k < n; k++) {
        uint32_t v = nodes[k];
        uint32_t cur = g->first_edge[v], t;
        while ((t = zo_next_succ(g, &cur)) != ZO_NONE) in_col[fill[t]++] = v;
    }
    double n_f = (double)n;
    double initial_rank = 1.0 / n_f;
    for (uint32_t k = 0; k < n; k++) ranks[nodes[k]] = initial_rank;

    uint64_t it = 0;
    double l1_norm = 0.0;
    while (it < max_iterations) {
        double sink_sum = 0.0;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t v = nodes[k];
            if (out_deg[v] == 0) sink_sum += ranks[v] / n_f;
        }
        zo_pr_rows job = {nodes, 0, n, in_ptr, out_deg, in_col, ranks, new_ranks, damping, n_f, sink_sum};
        if (n_threads <= 1) {
            zo_pr_rows_run(&job);
        } else {
            zo_pr_rows jobs[64];
            pthread_t th[64];
            int joined[64];
            int nt = n_threads > 64 ? 64 : n_threads;
            uint32_t chunk = (n + (uint32_t)nt - 1) / (uint32_t)nt;
            for (int t = 0; t < nt; t++) {
                jobs[t] = job;
                jobs[t].k0 = (uint32_t)t * chunk < n ? (uint32_t)t * chunk : n;
                jobs[t].k1 = jobs[t].k0 + chunk < n ? jobs[t].k0 + chunk : n;
                joined[t] = pthread_create(&th[t], NULL, zo_pr_rows_worker, &jobs[t]) == 0;
                if (!joined[t]) zo_pr_rows_run(&jobs[t]);   /* no thread: run the range here */
            }
            for (int t = 0; t < nt; t++)
                if (joined[t]) pthread_join(th[t], NULL);
        }
        l1_norm = 0.0;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t v = nodes[k];
            l1_norm += fabs(new_ranks[v] - ranks[v]);
        }
        double *tmp = ranks;
        ranks = new_ranks;
        new_ranks = tmp;
        it++;
        if (l1_norm < tolerance) break;
    }
    for (uint32_t k = 0; k < n; k++) out[nodes[k]] = ranks[nodes[k]];
    if (iters_out) *iters_out = it;
    if (l1_out) *l1_out = l1_norm;
    free(nodes); free(in_ptr); free(out_deg); free(in_col); free(fill); free(ranks); free(new_ranks);
    return ZO_OK;
}

/* ------------------------------------------------------------------------- */
/* Core numbers (connectivity.zig:5-101)                                      */
/* ------------------------------------------------------------------------- */

/* out[n_cap]: core number of every live node of the graph taken as
 * undirected and simple (0 for non-live ids).  As the reference: the
 * neighbour lists are built from each live node's successors (self-loops
 * skipped, a pair added once, :27-45), then bucket peeling in increasing
 * degree, a node's core number being the bucket it is popped from and each
 * unprocessed neighbour's degree dropping by one into bucket max(deg, i)
 * (:47-97).  A per-node mark replaces the reference's linear scan for a
 * repeated neighbour; the lists and the peeling order are the same. */
/* Ascending sort of n u64 keys: LSD radix sort, 16-bit digits, a pass
 * skipped when every key has the same digit (any correct sort gives the same
 * array).  Returns 0, or -1 when the buffer cannot be allocated. */
static int zo_sort_u64(uint64_t *keys, uint64_t n)
{
    if (n < 2) return 0;
    uint64_t *tmp = malloc(sizeof(uint64_t) * n);
    uint64_t *cnt = malloc(sizeof(uint64_t) * 65536);
    if (!tmp || !cnt) {
        free(tmp); free(cnt);
        return -1;
    }
    uint64_t *a = keys, *b = tmp;
    for (int sh = 0; sh < 64; sh += 16) {
        memset(cnt, 0, sizeof(uint64_t) * 65536);
        for (uint64_t i = 0; i < n; i++) cnt[(a[i] >> sh) & 0xFFFF]++;
        if (cnt[(a[0] >> sh) & 0xFFFF] == n) continue;
        uint64_t sum = 0;
        for (int d = 0; d < 65536; d++) {
            uint64_t c = cnt[d];
            cnt[d] = sum;
            sum += c;
        }
        for (uint64_t i = 0; i < n; i++) b[cnt[(a[i] >> sh) & 0xFFFF]++] = a[i];
        uint64_t *t = a;
        a = b;
        b = t;
    }
    if (a != keys) memcpy(keys, a, sizeof(uint64_t) * n);
    free(tmp);
    free(cnt);
    return 0;
}

int zo_core_numbers(const zo_graph *g, uint32_t *out)
{
    const uint32_t V = g->n_cap;
    for (uint32_t i = 0; i < V; i++) out[i] = 0;
    if (V == 0) return ZO_OK;
    uint32_t *deg = calloc(V, sizeof(uint32_t));
    uint64_t *cnt = calloc((size_t)V + 1, sizeof(uint64_t));
    if (!deg || !cnt) {
        free(deg); free(cnt);
        return ZO_ERR_OOM;
    }
    /* pass 1: collect the undirected simple pairs (u < v keys, sorted) */
    uint64_t m2 = 0;
    for (uint32_t u = 0; u < V; u++) {
        if (g->node_deleted[u]) continue;
        uint32_t cur = g->first_edge[u], t;
        while ((t = zo_next_succ(g, &cur)) != ZO_NONE) m2 += t != u;
    }
    uint64_t *keys = malloc(sizeof(uint64_t) * (m2 ? m2 : 1));
    if (!keys) {
        free(deg); free(cnt);
        return ZO_ERR_OOM;
    }
    uint64_t k = 0;
    for (uint32_t u = 0; u < V; u++) {
        if (g->node_deleted[u]) continue;
        uint32_t cur = g->first_edge[u], t;
        while ((t = zo_next_succ(g, &cur)) != ZO_NONE)
            if (t != u) keys[k++] = u < t ? ((uint64_t)u << 32 | t) : ((uint64_t)t << 32 | u);
    }
    /* sort + unique (insertion order of the lists does not change core numbers) */
    if (zo_sort_u64(keys, m2) != 0) {
        free(deg); free(cnt); free(keys);
        return ZO_ERR_OOM;
    }
    uint64_t m = 0;
    for (uint64_t i = 0; i < m2; i++)
        if (m == 0 || keys[i] != keys[m - 1]) keys[m++] = keys[i];
    for (uint64_t i = 0; i < m; i++) {
        cnt[(keys[i] >> 32) + 1]++;
        cnt[(uint32_t)keys[i] + 1]++;
    }
    for (uint32_t i = 0; i < V; i++) cnt[i + 1] += cnt[i];
    uint32_t *adj = malloc(sizeof(uint32_t) * (m ? 2 * m : 1));
    uint64_t *fill = malloc(sizeof(uint64_t) * V);
    if (!adj || !fill) {
        free(deg); free(cnt); free(keys); free(adj); free(fill);
        return ZO_ERR_OOM;
    }
    memcpy(fill, cnt, sizeof(uint64_t) * V);
    for (uint64_t i = 0; i < m; i++) {
        const uint32_t u = (uint32_t)(keys[i] >> 32), v = (uint32_t)keys[i];
        adj[fill[u]++] = v;
        adj[fill[v]++] = u;
    }
    free(keys);
    free(fill);
    uint32_t max_deg = 0;
    for (uint32_t v = 0; v < V; v++) {
        deg[v] = (uint32_t)(cnt[v + 1] - cnt[v]);
        if (deg[v] > max_deg) max_deg = deg[v];
    }
    /* buckets as stacks in one array with per-bucket linked storage: every
     * push is an entry (node, next) so repeated pushes of a node are kept */
    uint64_t cap = (uint64_t)V + 2 * m + 1, used = 0;
    uint32_t *e_node = malloc(sizeof(uint32_t) * cap);
    uint64_t *e_next = malloc(sizeof(uint64_t) * cap);
    uint64_t *head = malloc(sizeof(uint64_t) * ((size_t)max_deg + 1));
    uint8_t *done = calloc(V, 1);
    if (!e_node || !e_next || !head || !done) {
        free(deg); free(cnt); free(adj); free(e_node); free(e_next); free(head); free(done);
        return ZO_ERR_OOM;
    }
    const uint64_t NIL = ~(uint64_t)0;
    for (uint32_t i = 0; i <= max_deg; i++) head[i] = NIL;
    for (uint32_t v = 0; v < V; v++) {
        if (g->node_deleted[v]) continue;
        e_node[used] = v;
        e_next[used] = head[deg[v]];
        head[deg[v]] = used++;
    }
    for (uint32_t i = 0; i <= max_deg; i++) {
        while (head[i] != NIL) {
            const uint64_t e = head[i];
            head[i] = e_next[e];
            const uint32_t u = e_node[e];
            if (done[u]) continue;
            out[u] = i;
            done[u] = 1;
            for (uint64_t a = cnt[u]; a < cnt[u + 1]; a++) {
                const uint32_t v = adj[a];
                if (done[v]) continue;
                if (deg[v] > 0) deg[v]--;
                const uint32_t b = deg[v] > i ? deg[v] : i;
                e_node[used] = v;
                e_next[used] = head[b];
                head[b] = used++;
            }
        }
    }
    free(deg); free(cnt); free(adj); free(e_node); free(e_next); free(head); free(done);
    return ZO_OK;
}

/* ------------------------------------------------------------------------- */
/* Weighted Brandes (centrality.zig:529-667, betweennessF64 :1355-1357)       */
/* ------------------------------------------------------------------------- */

/* Edge data of arc e (arc order of zo_graph_create). */
int zo_graph_set_weights(zo_graph *g, const double *w, uint64_t m)
{
    if (m != g->m) return ZO_ERR_ARG;
    double *c = (double *)malloc(sizeof(double) * (m ? m : 1));
    if (!c) return ZO_ERR_OOM;
    for (uint64_t e = 0; e < m; e++) c[e] = w[e];
    free(g->weight);
    g->weight = c;
    return ZO_OK;
}

/* As zo_next_succ, also returning the arc. */
static inline uint32_t zo_next_arc(const zo_graph *g, uint32_t *cur, uint32_t *arc)
{
    while (*cur != ZO_NONE) {
        uint32_t e = *cur;
        *cur = g->next_edge[e];
        if (!g->edge_deleted[e]) {
            uint32_t w = g->to[e];
            if (!g->node_deleted[w]) {
                *arc = e;
                return w;
            }
        }
    }
    return ZO_NONE;
}

/* Binary min-heap of (distance, node) keyed by distance: append + sift up,
 * pop = move the last item to the root + sift down (the smaller child, the
 * right one only when strictly smaller; stop when the moved item is strictly
 * smaller than that child). */
typedef struct {
    double d;
    uint32_t node;
} zo_item;

static void zo_heap_push(zo_item *h, size_t *n, zo_item x)
{
    size_t i = (*n)++;
    while (i > 0) {
        size_t p = (i - 1) >> 1;
        if (!(x.d < h[p].d)) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = x;
}

static zo_item zo_heap_pop(zo_item *h, size_t *n)
{
    zo_item top = h[0];
    zo_item last = h[--(*n)];
    if (*n == 0) return top;
    size_t i = 0;
    for (;;) {
        size_t c = 2 * i + 1;
        if (c >= *n) break;
        if (c + 1 < *n && h[c + 1].d < h[c].d) c++;
        if (last.d < h[c].d) break;
        h[i] = h[c];
        i = c;
    }
    h[i] = last;
    return top;
}

/* Unnormalised weighted BC into out[n_cap] over sources[0..n) (NULL = every
 * live node in id order).  Per source (:592-647): dist unset, sigma 0, delta
 * 0, preds cleared; the source at 0 with sigma 1; pop the closest item, skip
 * it if its distance is above the node's recorded one, push the node on the
 * stack, relax its arcs: strictly shorter -> new distance, sigma copied,
 * preds reset to {v}, pushed; equal -> sigma added, v appended to preds;
 * unset -> as strictly shorter.  Backward (:649-663): pop the stack, for each
 * pred u: delta[u] += (sigma[u] / sigma[v]) * (1 + delta[v]); then
 * bc[v] += delta[v] unless v is the source.  sigma is u64 (the reference's
 * usize).  Returns 0, or <0 (no weights set, out of memory). */
int zo_bc_f64(const zo_graph *g, const uint32_t *sources, uint64_t n_sources, double *out)
{
    const uint32_t V = g->n_cap;
    if (!g->weight) return ZO_ERR_ARG;
    for (uint32_t i = 0; i < V; i++) out[i] = 0.0;
    double *dist = malloc(sizeof(double) * (V ? V : 1));
    uint8_t *has = malloc(V ? V : 1);
    uint64_t *sigma = malloc(sizeof(uint64_t) * (V ? V : 1));
    double *delta = malloc(sizeof(double) * (V ? V : 1));
    uint32_t *stack = malloc(sizeof(uint32_t) * (V ? V : 1));
    /* preds as per-node growable arrays */
    uint32_t **pr = calloc(V ? V : 1, sizeof(uint32_t *));
    uint32_t *pn = calloc(V ? V : 1, sizeof(uint32_t)), *pc = calloc(V ? V : 1, sizeof(uint32_t));
    size_t hcap = (size_t)g->m + 2;
    zo_item *heap = malloc(sizeof(zo_item) * hcap);
    int rc = ZO_OK;
    if (!dist || !has || !sigma || !delta || !stack || !pr || !pn || !pc || !heap) rc = ZO_ERR_OOM;
    uint64_t ns = sources ? n_sources : V;
    for (uint64_t k = 0; rc == ZO_OK && k < ns; k++) {
        uint32_t s = sources ? sources[k] : (uint32_t)k;
        if (!sources && g->node_deleted[s]) continue;
        if (s >= V || g->node_deleted[s]) {
            rc = ZO_ERR_NODE_NOT_FOUND;
            break;
        }
        for (uint32_t i = 0; i < V; i++) {
            has[i] = 0;
            sigma[i] = 0;
            delta[i] = 0.0;
            pn[i] = 0;
        }
        size_t hn = 0, sn = 0;
        dist[s] = 0.0;
        has[s] = 1;
        sigma[s] = 1;
        zo_heap_push(heap, &hn, (zo_item){0.0, s});
        while (hn > 0) {
            zo_item it = zo_heap_pop(heap, &hn);
            const double dv = it.d;
            const uint32_t v = it.node;
            const double best = has[v] ? dist[v] : dv;
            if (dv > best) continue;
            stack[sn++] = v;
            uint32_t cur = g->first_edge[v], arc, w;
            while ((w = zo_next_arc(g, &cur, &arc)) != ZO_NONE) {
                const double nd = dv + g->weight[arc];
                int shorter = !has[w] || nd < dist[w];
                int equal = has[w] && nd == dist[w];
                if (shorter) {
                    dist[w] = nd;
                    has[w] = 1;
                    sigma[w] = sigma[v];
                    pn[w] = 0;
                } else if (equal) {
                    sigma[w] += sigma[v];
                }
                if (shorter || equal) {
                    if (pn[w] == pc[w]) {
                        uint32_t nc = pc[w] ? 2 * pc[w] : 4;
                        uint32_t *np = realloc(pr[w], sizeof(uint32_t) * nc);
                        if (!np) {
                            rc = ZO_ERR_OOM;
                            break;
                        }
                        pr[w] = np;
                        pc[w] = nc;
                    }
                    pr[w][pn[w]++] = v;
                }
                if (shorter) {
                    if (hn == hcap) {
                        hcap *= 2;
                        zo_item *nh = realloc(heap, sizeof(zo_item) * hcap);
                        if (!nh) {
                            rc = ZO_ERR_OOM;
                            break;
                        }
                        heap = nh;
                    }
                    zo_heap_push(heap, &hn, (zo_item){nd, w});
                }
            }
            if (rc != ZO_OK) break;
        }
        while (rc == ZO_OK && sn > 0) {
            const uint32_t v = stack[--sn];
            const double sv = (double)sigma[v];
            const double dlt = delta[v];
            for (uint32_t i = 0; i < pn[v]; i++) {
                const uint32_t u = pr[v][i];
                delta[u] += ((double)sigma[u] / sv) * (1.0 + dlt);
            }
            if (v != s) out[v] += dlt;
        }
    }
    for (uint32_t i = 0; pr && i < V; i++) free(pr[i]);
    free(pr); free(pn); free(pc); free(heap);
    free(dist); free(has); free(sigma); free(delta); free(stack);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Eigenvector, Katz and alpha centrality (centrality.zig:1050-1201,          */
/* :1212-1267, :1278-1333)                                                    */
/* ------------------------------------------------------------------------- */

/* collectNodes (utils.zig:42-49) and buildInNeighbors (utils.zig:52-73): the
 * live ids ascending, and for each live `from` in that order, `from` appended
 * to in[to] for every successor (parallel arcs repeated, self-loops kept). */
typedef struct {
    uint32_t n;
    uint32_t *nodes;
    uint64_t *in_ptr; /* [V + 1] */
    uint32_t *in_col;
} zo_inlists;

static void zo_inlists_free(zo_inlists *L)
{
    free(L->nodes);
    free(L->in_ptr);
    free(L->in_col);
}

static int zo_inlists_build(const zo_graph *g, zo_inlists *L)
{
    const uint32_t V = g->n_cap;
    memset(L, 0, sizeof(*L));
    L->nodes = malloc(sizeof(uint32_t) * (V ? V : 1));
    L->in_ptr = calloc((size_t)V + 1, sizeof(uint64_t));
    if (!L->nodes || !L->in_ptr) {
        zo_inlists_free(L);
        return ZO_ERR_OOM;
    }
    for (uint32_t i = 0; i < V; i++)
        if (!g->node_deleted[i]) L->nodes[L->n++] = i;
    for (uint32_t k = 0; k < L->n; k++) {
        uint32_t cur = g->first_edge[L->nodes[k]], t;
        while ((t = zo_next_succ(g, &cur)) != ZO_NONE) L->in_ptr[t + 1]++;
    }
    for (uint32_t i = 0; i < V; i++) L->in_ptr[i + 1] += L->in_ptr[i];
    const uint64_t E = L->in_ptr[V];
    L->in_col = malloc(sizeof(uint32_t) * (E ? E : 1));
    uint64_t *fill = malloc(sizeof(uint64_t) * (V ? V : 1));
    if (!L->in_col || !fill) {
        free(fill);
        zo_inlists_free(L);
        return ZO_ERR_OOM;
    }
    memcpy(fill, L->in_ptr, sizeof(uint64_t) * V);
    for (uint32_t k = 0; k < L->n; k++) {
        const uint32_t v = L->nodes[k];
        uint32_t cur = g->first_edge[v], t;
        while ((t = zo_next_succ(g, &cur)) != ZO_NONE) L->in_col[fill[t]++] = v;
    }
    free(fill);
    return ZO_OK;
}

/* sum over in(v) of x, in list order (the `orelse 0.0` lookups of the
 * reference: every in-neighbour is live, so it has a value) */
static double zo_in_sum(const zo_inlists *L, const double *x, uint32_t v)
{
    double s = 0.0;
    for (uint64_t e = L->in_ptr[v]; e < L->in_ptr[v + 1]; e++) s += x[L->in_col[e]];
    return s;
}

/* Eigenvector centrality (centrality.zig:1050-1201).  out[n_cap] (0.0 for
 * non-live).  *iters_out = iterations run, *outcome_out: 0 converged,
 * 1 oscillation (average returned), 2 zero vector (uniform returned),
 * 3 iteration cap (last iterate returned), 4 at most one node.  The L2 norms
 * are summed in node order (the reference sums the first one in hash-map
 * order, so its last bits may differ). */
int zo_eigenvector(const zo_graph *g, uint64_t max_iterations, double tolerance, double *out,
                   uint64_t *iters_out, int *outcome_out)
{
    const uint32_t V = g->n_cap;
    for (uint32_t i = 0; i < V; i++) out[i] = 0.0;
    zo_inlists L;
    int rc = zo_inlists_build(g, &L);
    if (rc) return rc;
    const uint32_t n = L.n;
    *iters_out = 0;
    if (n <= 1) { /* :1061-1064 */
        for (uint32_t k = 0; k < n; k++) out[L.nodes[k]] = 1.0;
        *outcome_out = 4;
        zo_inlists_free(&L);
        return ZO_OK;
    }
    double *pp = calloc(V, sizeof(double)), *cur = calloc(V, sizeof(double)), *nw = calloc(V, sizeof(double));
    if (!pp || !cur || !nw) {
        free(pp); free(cur); free(nw);
        zo_inlists_free(&L);
        return ZO_ERR_OOM;
    }
    for (uint32_t k = 0; k < n; k++) cur[L.nodes[k]] = 1.0 + (double)k / 1000.0; /* :1078-1080 */
    int have_pp = 0;
    int outcome = 3;
    uint64_t it = 0;
    for (; it < max_iterations; it++) {
        for (uint32_t k = 0; k < n; k++) nw[L.nodes[k]] = zo_in_sum(&L, cur, L.nodes[k]);
        double l2 = 0.0;
        for (uint32_t k = 0; k < n; k++) l2 += nw[L.nodes[k]] * nw[L.nodes[k]];
        if (l2 > 0.0) {
            const double norm = sqrt(l2);
            for (uint32_t k = 0; k < n; k++) nw[L.nodes[k]] /= norm;
        }
        if (l2 == 0.0) { /* :1103-1116 */
            const double u = 1.0 / sqrt((double)n);
            for (uint32_t k = 0; k < n; k++) out[L.nodes[k]] = u;
            outcome = 2;
            break;
        }
        double d1 = 0.0;
        for (uint32_t k = 0; k < n; k++) {
            const double d = nw[L.nodes[k]] - cur[L.nodes[k]];
            d1 += d * d;
        }
        if (d1 > 0.0) d1 = sqrt(d1);
        int osc = 0;
        if (have_pp) { /* :1130-1142 */
            double d2 = 0.0;
            for (uint32_t k = 0; k < n; k++) {
                const double d = nw[L.nodes[k]] - pp[L.nodes[k]];
                d2 += d * d;
            }
            if (d2 > 0.0) d2 = sqrt(d2);
            osc = d2 < tolerance;
        }
        if (osc) { /* :1144-1167: normalised average of the last two */
            double a2 = 0.0;
            for (uint32_t k = 0; k < n; k++) {
                const uint32_t v = L.nodes[k];
                out[v] = (nw[v] + cur[v]) / 2.0;
                a2 += out[v] * out[v];
            }
            if (a2 > 0.0) {
                const double an = sqrt(a2);
                for (uint32_t k = 0; k < n; k++) out[L.nodes[k]] /= an;
            }
            outcome = 1;
            break;
        }
        if (d1 < tolerance) { /* :1169-1181 */
            for (uint32_t k = 0; k < n; k++) out[L.nodes[k]] = nw[L.nodes[k]];
            outcome = 0;
            break;
        }
        /* rotate: prev_prev <- curr <- new (:1183-1188) */
        double *t = pp;
        pp = cur;
        cur = nw;
        nw = t;
        have_pp = 1;
    }
    if (outcome == 3)
        for (uint32_t k = 0; k < n; k++) out[L.nodes[k]] = cur[L.nodes[k]];
    *iters_out = outcome == 3 ? it : it + 1;
    *outcome_out = outcome;
    free(pp); free(cur); free(nw);
    zo_inlists_free(&L);
    return ZO_OK;
}

/* Katz (centrality.zig:1212-1267) when has_beta, else alpha centrality
 * (:1278-1333) with start value `start`: x <- alpha * sum_in(x) (+ beta),
 * L1 change in node order, swap, stop when it is below tolerance. */
int zo_katz_alpha(const zo_graph *g, double alpha, double beta, int has_beta, double start,
                  uint64_t max_iterations, double tolerance, double *out, uint64_t *iters_out,
                  double *l1_out)
{
    const uint32_t V = g->n_cap;
    for (uint32_t i = 0; i < V; i++) out[i] = 0.0;
    zo_inlists L;
    int rc = zo_inlists_build(g, &L);
    if (rc) return rc;
    const uint32_t n = L.n;
    *iters_out = 0;
    *l1_out = 0.0;
    if (n == 0) {
        zo_inlists_free(&L);
        return ZO_OK;
    }
    double *cur = calloc(V, sizeof(double)), *nw = calloc(V, sizeof(double));
    if (!cur || !nw) {
        free(cur); free(nw);
        zo_inlists_free(&L);
        return ZO_ERR_OOM;
    }
    for (uint32_t k = 0; k < n; k++) cur[L.nodes[k]] = start;
    uint64_t it = 0;
    double l1 = 0.0;
    while (it < max_iterations) {
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t v = L.nodes[k];
            const double s = zo_in_sum(&L, cur, v);
            nw[v] = has_beta ? alpha * s + beta : alpha * s;
        }
        l1 = 0.0;
        for (uint32_t k = 0; k < n; k++) l1 += fabs(nw[L.nodes[k]] - cur[L.nodes[k]]);
        double *t = cur;
        cur = nw;
        nw = t;
        it++;
        if (l1 < tolerance) break;
    }
    for (uint32_t k = 0; k < n; k++) out[L.nodes[k]] = cur[L.nodes[k]];
    *iters_out = it;
    *l1_out = l1;
    free(cur); free(nw);
    zo_inlists_free(&L);
    return ZO_OK;
}
