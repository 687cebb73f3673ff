// Weighted Brandes betweenness (betweennessF64) on the gfx950.
//
// Reference: priv/zog/src/centrality.zig:529-667 with betweennessF64
// (:1355-1357): f64 arc weights, distances added in path order (a + b), a
// shortest path is one whose sum equals the node's distance exactly (`.eq`),
// path counts sigma add over every such arc (parallel arcs separately), and
// delta(u) += (sigma_u / sigma_v)(1 + delta_v) over the shortest-path DAG in
// reverse order; bc(v) += delta(v) for v != s, unnormalised.
//
// GPU form, per source (up to 1,024 sources per batch, each with its own
// rows; one wavefront per 64 entries of one source's list):
//   1. distances by frontier relaxation: every vertex whose distance dropped
//      relaxes its arcs with a 64-bit atomic min on the distance's bit
//      pattern (non-negative doubles order as unsigned integers); a vertex
//      is queued at most once per round.  It ends at the least fixed point
//      d(w) = min over arcs of fl(d(v) + w(v,w)), the distances Dijkstra
//      computes with the same additions.
//   2. the tight arcs (fl(d(v) + w) == d(w) bit for bit, v != w, w not the
//      source; with positive weights neither exception is ever tight) form the
//      shortest-path DAG: count their in-arcs, then peel it in rounds (a
//      vertex enters the round after its last tight in-arc was processed),
//      adding sigma(v) into every tight successor.
//   3. rounds in reverse: delta(v) = sigma(v) * sum over tight successors of
//      c(w), c(v) = (1 + delta(v)) / sigma(v); bc(v) += delta(v) (FP64
//      atomics across sources).
// Arc weights must be positive for the reference's semantics; a zero-weight
// cycle would never be peeled (its vertices get no sigma).
#include "common.h"
#include "queue_map.h"

#include <algorithm>

#define ZG_W_MAX_SOURCES 1024
#define ZG_W_GRID 1024
#define ZG_W_ROUNDS_PER_SYNC 16
#define ZG_W_UNSET 0xFFFFFFFFFFFFFFFFull

struct WbcDev {
    uint32_t V, S, lcap;
    const uint32_t *out_ptr, *out_col;
    const double *out_w;
    unsigned long long *dist;   // [S][V] bits of the distance, ZG_W_UNSET = unreached
    double *sig, *cc;           // [S][V]
    uint32_t *indeg;            // [S][V] tight in-arcs not yet processed
    uint32_t *stamp;            // [S][V] relax round that queued the vertex last
    uint32_t *fr;               // [2][S][V] relax frontiers
    uint32_t *fcnt;             // [2][S]
    uint32_t *reach;            // [S][V] reached vertices
    uint32_t *rcnt;             // [S]
    uint32_t *order;            // [S][V] DAG order, round by round
    uint32_t *qlen;             // [S]
    uint32_t *lstart;           // [lcap][S] round starts in order[]
    uint32_t *tot;              // [lcap] frontier sizes (relax) / entries added (peel)
    const uint32_t *src;        // [S]
    double *bc;                 // [V]
    unsigned long long *stats;  // [2] arcs, reached
};

#define WROW(s) ((size_t)(s) * a.V)

__device__ __forceinline__ double zg_bits_to_f64(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long zg_f64_to_bits(double d) { return (unsigned long long)__double_as_longlong(d); }

// Appends `item` of every lane with `want` to list[counter...] (one atomic
// per wavefront).  All lanes call it.
__device__ __forceinline__ void wbc_append(bool want, uint32_t item, uint32_t *counter, uint32_t *list, int lane)
{
    const uint64_t b = __ballot(want);
    if (!b) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(counter, (uint32_t)__popcll(b));
    base = zg_readlane32(base, 0);
    if (want) list[base + __popcll(b & ((1ull << lane) - 1))] = item;
}

__global__ void wbc_init_kernel(WbcDev a)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.S) return;
    a.fcnt[a.S + s] = 0;
    a.lstart[s] = 0;
    const uint32_t v = a.src[s];
    if (v == ZG_NONE) {
        a.fcnt[s] = 0;
        a.rcnt[s] = 0;
        a.qlen[s] = 0;
        a.lstart[a.S + s] = 0;
        return;
    }
    a.dist[WROW(s) + v] = 0ull;  // +0.0
    a.sig[WROW(s) + v] = 1.0;
    a.stamp[WROW(s) + v] = 0;
    a.fr[WROW(s)] = v;
    a.fcnt[s] = 1;
    a.reach[WROW(s)] = v;
    a.rcnt[s] = 1;
    a.order[WROW(s)] = v;
    a.qlen[s] = 1;
    a.lstart[a.S + s] = 1;
}

// Before relax round r: the size of its frontier (tot[r]) and an empty
// buffer for the next one.  One block.
__global__ void wbc_round_kernel(WbcDev a, int r)
{
    __shared__ unsigned long long red[16];
    const int cur = r & 1, nxt = cur ^ 1;
    unsigned long long t = 0;
    for (uint32_t s = threadIdx.x; s < a.S; s += blockDim.x) {
        t += a.fcnt[(size_t)cur * a.S + s];
        a.fcnt[(size_t)nxt * a.S + s] = 0;
    }
    t = zg_wave_sum_u64(t);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long T = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; w++) T += red[w];
        a.tot[r] = (uint32_t)std::min<unsigned long long>(T, 0xFFFFFFFFull);
    }
}

__global__ __launch_bounds__(256) void wbc_relax_kernel(WbcDev a, int r)
{
    __shared__ uint32_t s_lo[ZG_W_MAX_SOURCES], s_hi[ZG_W_MAX_SOURCES], s_pre[ZG_W_MAX_SOURCES + 1];
    const int cur = r & 1, nxt = cur ^ 1;
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = 0;
            hi = a.fcnt[(size_t)cur * a.S + s];
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < total; c += waves) {
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        unsigned long long *dist = a.dist + WROW(s);
        const uint32_t v = valid ? a.fr[((size_t)cur * a.S + s) * a.V + idx] : 0;
        const double dv = valid ? zg_bits_to_f64(dist[v]) : 0.0;  // latest value
        const uint32_t e0 = valid ? a.out_ptr[v] : 0, deg = valid ? a.out_ptr[v + 1] - e0 : 0;
        uint32_t maxdeg = deg;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) maxdeg = max(maxdeg, (uint32_t)__shfl_xor((int)maxdeg, off, 64));
        for (uint32_t k = 0; k < maxdeg; k++) {
            bool first = false, queue = false;
            uint32_t w = 0;
            if (k < deg) {
                w = a.out_col[e0 + k];
                const unsigned long long nb = zg_f64_to_bits(dv + a.out_w[e0 + k]);
                if (nb < dist[w]) {
                    const unsigned long long old = atomicMin(&dist[w], nb);
                    if (nb < old) {
                        first = old == ZG_W_UNSET;
                        queue = atomicExch(&a.stamp[WROW(s) + w], (uint32_t)(r + 1)) != (uint32_t)(r + 1);
                    }
                }
            }
            wbc_append(queue, w, &a.fcnt[(size_t)nxt * a.S + s], a.fr + ((size_t)nxt * a.S + s) * a.V, lane);
            wbc_append(first, w, &a.rcnt[s], a.reach + WROW(s), lane);
        }
    }
}

// Tight in-arc counts over every reached vertex, and the sweep statistics.
__global__ __launch_bounds__(256) void wbc_count_kernel(WbcDev a)
{
    __shared__ uint32_t s_lo[ZG_W_MAX_SOURCES], s_hi[ZG_W_MAX_SOURCES], s_pre[ZG_W_MAX_SOURCES + 1];
    __shared__ unsigned long long red[2][4];
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = 0;
            hi = a.rcnt[s];
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    unsigned long long arcs = 0, reached = 0;
    for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + wv; c < total; c += waves) {
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        if (idx >= s_hi[s]) continue;
        const unsigned long long *dist = a.dist + WROW(s);
        const uint32_t v = a.reach[WROW(s) + idx], src = a.src[s];
        const double dv = zg_bits_to_f64(dist[v]);
        const uint32_t e0 = a.out_ptr[v], e1 = a.out_ptr[v + 1];
        arcs += e1 - e0;
        reached += 1;
        for (uint32_t e = e0; e < e1; e++) {
            const uint32_t w = a.out_col[e];
            if (w != v && w != src && zg_f64_to_bits(dv + a.out_w[e]) == dist[w])
                atomicAdd(&a.indeg[WROW(s) + w], 1u);
        }
    }
    arcs = zg_wave_sum_u64(arcs);
    reached = zg_wave_sum_u64(reached);
    if (lane == 0) {
        red[0][wv] = arcs;
        red[1][wv] = reached;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long A = 0, R = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            A += red[0][w];
            R += red[1][w];
        }
        if (A) atomicAdd(&a.stats[0], A);
        if (R) atomicAdd(&a.stats[1], R);
    }
}

// Peel round r of the DAG: add sigma into tight successors; a successor whose
// last tight in-arc this was joins round r+1.
__global__ __launch_bounds__(256) void wbc_peel_kernel(WbcDev a, int r)
{
    __shared__ uint32_t s_lo[ZG_W_MAX_SOURCES], s_hi[ZG_W_MAX_SOURCES], s_pre[ZG_W_MAX_SOURCES + 1];
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = a.lstart[(size_t)r * a.S + s];
            hi = a.lstart[(size_t)(r + 1) * a.S + s];
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < total; c += waves) {
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        const unsigned long long *dist = a.dist + WROW(s);
        const uint32_t v = valid ? a.order[WROW(s) + idx] : 0, src = a.src[s];
        const double dv = valid ? zg_bits_to_f64(dist[v]) : 0.0;
        const double sv = valid ? a.sig[WROW(s) + v] : 0.0;
        const uint32_t e0 = valid ? a.out_ptr[v] : 0, deg = valid ? a.out_ptr[v + 1] - e0 : 0;
        uint32_t maxdeg = deg;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) maxdeg = max(maxdeg, (uint32_t)__shfl_xor((int)maxdeg, off, 64));
        for (uint32_t k = 0; k < maxdeg; k++) {
            bool ready = false;
            uint32_t w = 0;
            if (k < deg) {
                w = a.out_col[e0 + k];
                if (w != v && w != src && zg_f64_to_bits(dv + a.out_w[e0 + k]) == dist[w]) {
                    atomicAdd(&a.sig[WROW(s) + w], sv);
                    ready = atomicSub(&a.indeg[WROW(s) + w], 1u) == 1u;
                }
            }
            wbc_append(ready, w, &a.qlen[s], a.order + WROW(s), lane);
        }
    }
}

// Close peel round L (just appended): its end offset and size.  One block.
__global__ void wbc_close_kernel(WbcDev a, int L)
{
    __shared__ unsigned long long red[16];
    unsigned long long t = 0;
    for (uint32_t s = threadIdx.x; s < a.S; s += blockDim.x) {
        const uint32_t q = a.qlen[s];
        a.lstart[(size_t)(L + 1) * a.S + s] = q;
        t += q - a.lstart[(size_t)L * a.S + s];
    }
    t = zg_wave_sum_u64(t);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long T = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; w++) T += red[w];
        a.tot[L] = (uint32_t)std::min<unsigned long long>(T, 0xFFFFFFFFull);
    }
}

__global__ __launch_bounds__(256) void wbc_backward_kernel(WbcDev a, int r)
{
    __shared__ uint32_t s_lo[ZG_W_MAX_SOURCES], s_hi[ZG_W_MAX_SOURCES], s_pre[ZG_W_MAX_SOURCES + 1];
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = a.lstart[(size_t)r * a.S + s];
            hi = a.lstart[(size_t)(r + 1) * a.S + s];
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < total; c += waves) {
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        if (idx >= s_hi[s]) continue;
        const unsigned long long *dist = a.dist + WROW(s);
        const uint32_t v = a.order[WROW(s) + idx], src = a.src[s];
        const double dv = zg_bits_to_f64(dist[v]);
        double S = 0.0;
        for (uint32_t e = a.out_ptr[v], e1 = a.out_ptr[v + 1]; e < e1; e++) {
            const uint32_t w = a.out_col[e];
            if (w != v && w != src && zg_f64_to_bits(dv + a.out_w[e]) == dist[w]) S += a.cc[WROW(s) + w];
        }
        const double sv = a.sig[WROW(s) + v];
        const double delta = sv * S;
        a.cc[WROW(s) + v] = (1.0 + delta) / sv;
        if (v != src) atomicAdd(&a.bc[v], delta);
    }
}

__global__ void wbc_out_kernel(uint32_t V, const double *bc, const uint8_t *live, double *out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = live[v] ? bc[v] : 0.0;
}

int zg_wbc_ws_release(zg_graph *g)
{
    zg_wbc_ws &w = g->wbc;
    zg_dfree(w.dist); zg_dfree(w.sig); zg_dfree(w.cc); zg_dfree(w.indeg); zg_dfree(w.stamp); zg_dfree(w.fr);
    zg_dfree(w.fcnt); zg_dfree(w.reach); zg_dfree(w.rcnt); zg_dfree(w.order); zg_dfree(w.qlen);
    zg_dfree(w.lstart); zg_dfree(w.tot); zg_dfree(w.src); zg_dfree(w.bc); zg_dfree(w.stats);
    w.S = 0;
    w.lcap = 0;
    return ZG_OK;
}

static int wbc_ws_ensure(zg_graph *g, uint32_t S)
{
    zg_wbc_ws &w = g->wbc;
    if (w.S == S) return ZG_OK;
    zg_wbc_ws_release(g);
    const size_t V = g->V ? g->V : 1, SV = (size_t)S * V;
    w.lcap = 256;
    int rc = zg_dmalloc(&w.dist, SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.sig, SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.cc, SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.indeg, SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.stamp, SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.fr, 2 * SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.fcnt, 2 * (size_t)S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.reach, SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.rcnt, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.order, SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.qlen, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.lstart, (size_t)w.lcap * S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.tot, w.lcap);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.src, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.bc, V);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.stats, 2);
    if (rc != ZG_OK) {
        zg_wbc_ws_release(g);
        return rc;
    }
    w.S = S;
    return ZG_OK;
}

// Grows the per-round tables to hold round `need` (used prefix kept).
static int wbc_grow(zg_graph *g, uint32_t need, uint32_t used, hipStream_t s)
{
    zg_wbc_ws &w = g->wbc;
    if (need < w.lcap) return ZG_OK;
    uint32_t nl = w.lcap;
    while (nl <= need) nl *= 2;
    uint32_t *ls = nullptr, *tt = nullptr;
    ZG_TRY_RC(zg_dmalloc(&ls, (size_t)nl * w.S));
    int rc = zg_dmalloc(&tt, nl);
    if (rc != ZG_OK) {
        zg_dfree(ls);
        return rc;
    }
    used = std::min(std::max<uint32_t>(used, 2), w.lcap);  // rows 0 and 1 are set before the first round
    ZG_TRY(hipMemcpyAsync(ls, w.lstart, sizeof(uint32_t) * (size_t)used * w.S, hipMemcpyDeviceToDevice, s));
    ZG_TRY(hipMemcpyAsync(tt, w.tot, sizeof(uint32_t) * used, hipMemcpyDeviceToDevice, s));
    ZG_TRY(hipStreamSynchronize(s));
    zg_dfree(w.lstart);
    zg_dfree(w.tot);
    w.lstart = ls;
    w.tot = tt;
    w.lcap = nl;
    return ZG_OK;
}

static WbcDev wbc_dev(zg_graph *g)
{
    zg_wbc_ws &w = g->wbc;
    WbcDev a;
    a.V = g->V;
    a.S = w.S;
    a.lcap = w.lcap;
    a.out_ptr = g->out_ptr;
    a.out_col = g->out_col;
    a.out_w = g->out_w;
    a.dist = w.dist; a.sig = w.sig; a.cc = w.cc; a.indeg = w.indeg; a.stamp = w.stamp; a.fr = w.fr;
    a.fcnt = w.fcnt; a.reach = w.reach; a.rcnt = w.rcnt; a.order = w.order; a.qlen = w.qlen;
    a.lstart = w.lstart; a.tot = w.tot; a.src = w.src; a.bc = w.bc; a.stats = w.stats;
    return a;
}

static int wbc_sweep(zg_graph *g, const uint32_t *d_sources, uint64_t n, double *d_out, hipStream_t s, zg_bc_info *inf)
{
    const uint32_t V = g->V;
    if (n == 0 || V == 0) {
        if (V) ZG_TRY(hipMemsetAsync(d_out, 0, sizeof(double) * V, s));
        ZG_TRY(hipStreamSynchronize(s));
        inf->n_sources = n;
        return ZG_OK;
    }
    // sources per batch: as many as the call has (rounded up to 64), at most
    // ZG_W_MAX_SOURCES, and as many as half the free device memory holds at
    // 48 bytes of state per source and vertex
    const uint32_t want = (uint32_t)std::min<uint64_t>(ZG_W_MAX_SOURCES, (n + 63) / 64 * 64);
    uint32_t S = g->wbc.S;
    if (S < want) {
        size_t fr = 0, tot = 0;
        ZG_TRY(hipMemGetInfo(&fr, &tot));
        fr += (size_t)g->wbc.S * V * 48;
        const uint64_t fit = (uint64_t)(0.5 * (double)fr) / (48ull * (V ? V : 1));
        S = (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(want, fit) / 64 * 64);
    }
    ZG_TRY_RC(wbc_ws_ensure(g, S));
    zg_wbc_ws &w = g->wbc;
    const size_t SV = (size_t)S * V;
    ZG_TRY(hipMemsetAsync(w.bc, 0, sizeof(double) * (V ? V : 1), s));
    ZG_TRY(hipMemsetAsync(w.stats, 0, sizeof(unsigned long long) * 2, s));
    std::vector<uint32_t> tot(ZG_W_ROUNDS_PER_SYNC);
    for (uint64_t b0 = 0; b0 < n; b0 += S) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(S, n - b0);
        ZG_TRY(hipMemsetAsync(w.src, 0xFF, sizeof(uint32_t) * S, s));
        ZG_TRY(hipMemcpyAsync(w.src, d_sources + b0, sizeof(uint32_t) * nb, hipMemcpyDeviceToDevice, s));
        ZG_TRY(hipMemsetAsync(w.dist, 0xFF, sizeof(unsigned long long) * SV, s));
        ZG_TRY(hipMemsetAsync(w.sig, 0, sizeof(double) * SV, s));
        ZG_TRY(hipMemsetAsync(w.cc, 0, sizeof(double) * SV, s));
        ZG_TRY(hipMemsetAsync(w.indeg, 0, sizeof(uint32_t) * SV, s));
        ZG_TRY(hipMemsetAsync(w.stamp, 0xFF, sizeof(uint32_t) * SV, s));
        WbcDev a = wbc_dev(g);
        hipLaunchKernelGGL(wbc_init_kernel, dim3((S + 255) / 256), dim3(256), 0, s, a);
        ZG_TRY(hipGetLastError());
        // 1. distances: relax rounds, queued ZG_W_ROUNDS_PER_SYNC at a time
        uint32_t r = 0;
        for (bool done = false; !done;) {
            ZG_TRY_RC(wbc_grow(g, r + ZG_W_ROUNDS_PER_SYNC + 2, r + 1, s));
            a = wbc_dev(g);
            for (uint32_t k = 0; k < ZG_W_ROUNDS_PER_SYNC; k++) {
                hipLaunchKernelGGL(wbc_round_kernel, dim3(1), dim3(1024), 0, s, a, (int)(r + k));
                hipLaunchKernelGGL(wbc_relax_kernel, dim3(ZG_W_GRID), dim3(256), 0, s, a, (int)(r + k));
            }
            ZG_TRY(hipGetLastError());
            ZG_TRY(hipMemcpyAsync(tot.data(), w.tot + r, sizeof(uint32_t) * ZG_W_ROUNDS_PER_SYNC,
                                  hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            for (uint32_t k = 0; k < ZG_W_ROUNDS_PER_SYNC && !done; k++) done = tot[k] == 0;
            r += ZG_W_ROUNDS_PER_SYNC;
        }
        // 2. the tight DAG: in-arc counts, then peel rounds
        hipLaunchKernelGGL(wbc_count_kernel, dim3(ZG_W_GRID), dim3(256), 0, s, a);
        ZG_TRY(hipGetLastError());
        int last = -1;
        for (uint32_t p = 0; last < 0;) {
            ZG_TRY_RC(wbc_grow(g, p + ZG_W_ROUNDS_PER_SYNC + 2, p + 2, s));
            a = wbc_dev(g);
            for (uint32_t k = 0; k < ZG_W_ROUNDS_PER_SYNC; k++) {
                hipLaunchKernelGGL(wbc_peel_kernel, dim3(ZG_W_GRID), dim3(256), 0, s, a, (int)(p + k));
                hipLaunchKernelGGL(wbc_close_kernel, dim3(1), dim3(1024), 0, s, a, (int)(p + k + 1));
            }
            ZG_TRY(hipGetLastError());
            ZG_TRY(hipMemcpyAsync(tot.data(), w.tot + p + 1, sizeof(uint32_t) * ZG_W_ROUNDS_PER_SYNC,
                                  hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            for (uint32_t k = 0; k < ZG_W_ROUNDS_PER_SYNC; k++)
                if (tot[k] == 0) {
                    last = (int)(p + k);  // round p+k+1 is empty
                    break;
                }
            p += ZG_W_ROUNDS_PER_SYNC;
        }
        // 3. dependencies, rounds in reverse (round 0 is the source)
        a = wbc_dev(g);
        for (int q = last; q >= 1; q--)
            hipLaunchKernelGGL(wbc_backward_kernel, dim3(ZG_W_GRID), dim3(256), 0, s, a, q);
        ZG_TRY(hipGetLastError());
        inf->max_levels = std::max<uint32_t>(inf->max_levels, (uint32_t)last + 1);
        inf->push_levels += r;
        inf->batches++;
    }
    unsigned long long st[2] = {0, 0};
    ZG_TRY(hipMemcpyAsync(st, w.stats, sizeof(st), hipMemcpyDeviceToHost, s));
    if (V) {
        hipLaunchKernelGGL(wbc_out_kernel, dim3((V + 255) / 256), dim3(256), 0, s, V, w.bc, g->live, d_out);
        ZG_TRY(hipGetLastError());
    }
    ZG_TRY(hipStreamSynchronize(s));
    inf->arcs_traversed += st[0];
    inf->nodes_reached += st[1];
    inf->n_sources = n;
    return ZG_OK;
}

extern "C" int zg_bc_f64(zg_graph *g, const uint32_t *sources, uint64_t n_sources, double *out_bc, zg_bc_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out_bc && g->V) return ZG_ERR_INVALID_ARG;
    if (g->E && !g->out_w) return ZG_ERR_INVALID_ARG;  // the resource has no weights
    if (!g->weights_ok) return ZG_ERR_INVALID_ARG;
    std::vector<uint32_t> list;
    if (sources) {
        for (uint64_t i = 0; i < n_sources; i++)
            if (sources[i] >= g->V || !g->h_live[sources[i]]) return ZG_ERR_NODE_NOT_FOUND;
        list.assign(sources, sources + n_sources);
    } else {
        for (uint32_t v = 0; v < g->V; v++)
            if (g->h_live[v]) list.push_back(v);
    }
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = g->stream;
    zg_bc_info inf;
    memset(&inf, 0, sizeof(inf));
    uint32_t *d_src = nullptr;
    double *d_out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = zg_dmalloc(&d_src, list.size());
    if (rc == ZG_OK) rc = zg_dmalloc(&d_out, g->V);
    if (rc == ZG_OK && !list.empty())
        rc = zg_hip_code(hipMemcpyAsync(d_src, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice, s));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventCreate(&e0));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventCreate(&e1));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventRecord(e0, s));
    if (rc == ZG_OK) rc = wbc_sweep(g, d_src, list.size(), d_out, s, &inf);
    if (rc == ZG_OK) rc = zg_hip_code(hipEventRecord(e1, s));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventSynchronize(e1));
    if (rc == ZG_OK) {
        float ms = 0.f;
        rc = zg_hip_code(hipEventElapsedTime(&ms, e0, e1));
        inf.kernel_ms = ms;
    }
    if (rc == ZG_OK && g->V) rc = zg_hip_code(hipMemcpy(out_bc, d_out, sizeof(double) * g->V, hipMemcpyDeviceToHost));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    zg_dfree(d_src);
    zg_dfree(d_out);
    if (rc == ZG_OK && info) *info = inf;
    return rc;
}
