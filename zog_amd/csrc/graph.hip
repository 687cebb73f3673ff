// Resource build and lifetime: the device-resident successor / predecessor
// CSR that replaces the linked-list ArrayGraph for both hot paths.
//
// Reference: buildGraph (lib/zog/resource_graph.ex:120-129) appends all nodes
// then calls addEdge (priv/zog/src/models/array_graph.zig:88-108) in array
// order; addEdge prepends, so successors() yields arcs in reverse insertion
// order (array_graph.zig:94-104, 166-180).  PageRank's in-lists are filled by
// scanning live sources in id order (centrality.zig:881-899), so each is
// ascending by source with duplicates kept.
#include "common.h"

#include <cmath>

#include <algorithm>
#include <new>

#include "build_util.h"
#include "csr_host.h"
#include "pr_bins.h"

extern "C" const char *zg_strerror(int code)
{
    switch (code) {
    case ZG_OK: return "ok";
    case ZG_ERR_INVALID_ARG: return "invalid argument";
    case ZG_ERR_NODE_NOT_FOUND: return "node not found";
    case ZG_ERR_OOM: return "out of host memory";
    case ZG_ERR_GPU_OOM: return "out of device memory";
    case ZG_ERR_HIP: return "HIP runtime error";
    case ZG_ERR_NO_DEVICE: return "no gfx950 device";
    case ZG_ERR_TOO_LARGE: return "graph too large (more than 2^32-1 arcs or nodes)";
    case ZG_ERR_INVALID_HANDLE: return "invalid graph handle";
    case ZG_ERR_IO: return "file could not be read";
    case ZG_ERR_SCRATCH: return "input larger than the fixed scratch of the allocation-free entry point";
    default: return "unknown error";
    }
}

size_t zg_prof_begin(zg_graph *g, hipStream_t s)
{
    if (!g->profiling) return SIZE_MAX;
    while (g->ev_pool.size() < g->ev_used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return SIZE_MAX;
        g->ev_pool.push_back(e);
    }
    size_t i = g->ev_used;
    g->ev_used += 2;
    (void)hipEventRecord(g->ev_pool[i], s);
    return i;
}

void zg_prof_end(zg_graph *g, hipStream_t s, int slot, size_t start)
{
    if (start == SIZE_MAX) return;
    (void)hipEventRecord(g->ev_pool[start + 1], s);
    g->ev_marks.emplace_back(slot, start);
}

int zg_prof_collect(zg_graph *g, hipStream_t s)
{
    if (g->ev_marks.empty()) return ZG_OK;
    ZG_TRY(hipStreamSynchronize(s));
    for (auto &m : g->ev_marks) {
        float ms = 0.f;
        ZG_TRY(hipEventElapsedTime(&ms, g->ev_pool[m.second], g->ev_pool[m.second + 1]));
        g->prof.ms[m.first] += ms;
        g->prof.launches[m.first] += 1;
    }
    g->ev_marks.clear();
    g->ev_used = 0;
    return ZG_OK;
}

void zg_prof_release(zg_graph *g)
{
    for (hipEvent_t e : g->ev_pool) (void)hipEventDestroy(e);
    g->ev_pool.clear();
    g->ev_marks.clear();
    g->ev_used = 0;
}

extern "C" int zg_set_profiling(zg_graph *g, int on)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    std::lock_guard<std::mutex> lk(g->mu);
    g->profiling = on != 0;
    memset(&g->prof, 0, sizeof(g->prof));
    return ZG_OK;
}

extern "C" int zg_get_profile(zg_graph *g, zg_profile *out)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    *out = g->prof;
    memset(&g->prof, 0, sizeof(g->prof));
    return ZG_OK;
}

// Graphs with at least this many arcs build their CSR on the GPU.
static uint64_t g_build_gpu_min_arcs = 1u << 18;

extern "C" int zg_set_build_threshold(uint64_t min_arcs)
{
    g_build_gpu_min_arcs = min_arcs;
    return ZG_OK;
}

int zg_io_out(zg_graph *g, double **out)
{
    if (!g->io_out) ZG_TRY_RC(zg_dmalloc(&g->io_out, g->V));
    *out = g->io_out;
    return ZG_OK;
}

int zg_io_src(zg_graph *g, uint64_t n, uint32_t **src)
{
    if (n > g->io_src_cap) {
        zg_dfree(g->io_src);
        g->io_src_cap = 0;
        ZG_TRY_RC(zg_dmalloc(&g->io_src, n));
        g->io_src_cap = n;
    }
    *src = g->io_src;
    return ZG_OK;
}

static void zg_free_device(zg_graph *g)
{
    if (g->lane) {
        zg_free_device(g->lane);
        delete g->lane;
        g->lane = nullptr;
    }
    zg_dfree(g->lane_perm);
    zg_prof_release(g);
    zg_dfree(g->io_out);
    zg_dfree(g->io_src);
    zg_dfree(g->fold_succ1);
    zg_dfree(g->fold_count);
    g->io_src_cap = 0;
    zg_bc_ws_release(g);
    zg_bcq_ws_release(g);
    zg_wbc_ws_release(g);
    zg_pr_ws_release(g);
    zg_dfree(g->out_ptr); zg_dfree(g->out_col); zg_dfree(g->in_ptr); zg_dfree(g->in_col); zg_dfree(g->out_w);
    zg_bcq_order_release(g);
    zg_dfree(g->und_ptr); zg_dfree(g->und_col);
    zg_dfree(g->live);
    zg_dfree(g->ich_first); zg_dfree(g->ich_v); zg_dfree(g->ich_e0); zg_dfree(g->ich_e1);
    zg_dfree(g->och_first); zg_dfree(g->och_v); zg_dfree(g->och_e0); zg_dfree(g->och_e1);
    zg_dfree(g->ohub_idx);
    zg_dfree(g->pr_small); zg_dfree(g->pr_med); zg_dfree(g->pr_desc); zg_dfree(g->pr_large); zg_dfree(g->pr_large_seg0);
    zg_dfree(g->pr_seg_row); zg_dfree(g->pr_seg_e0); zg_dfree(g->pr_seg_e1);
    zg_dfree(g->pr_large_deg); zg_dfree(g->pr_id); zg_dfree(g->pr_col);
    zg_dfree(g->pr_long_plan); zg_dfree(g->pr_long_off);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    if (g->cap_stream) (void)hipStreamDestroy(g->cap_stream);
    g->stream = nullptr;
}

template <typename T>
static int zg_upload(T **d, const std::vector<T> &h)
{
    ZG_TRY_RC(zg_dmalloc(d, h.size()));
    if (!h.empty()) ZG_TRY(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return ZG_OK;
}

// Host form of the CSR build (zg_build_csr_gpu in graph_build.hip is the
// device form, used from ZG_GPU_BUILD_MIN_ARCS arcs on; both produce the
// same arrays): csr_host.cpp builds them, this uploads them with the weights
// and the live mask.
static int zg_build_csr_host(zg_graph *g, const uint32_t *from, const uint32_t *to, const double *weight, uint64_t m)
{
    ZgHostCsr c;
    ZG_TRY_RC(zg_host_csr(g->V, g->h_live.data(), from, to, weight, m, c));
    g->E = c.out_ptr[g->V];
    g->wmean = g->E ? c.wsum / (double)g->E : 0.0;
    g->weights_ok = c.weights_ok;
    if (weight) ZG_TRY_RC(zg_upload(&g->out_w, c.out_w));
    ZG_TRY_RC(zg_upload(&g->out_ptr, c.out_ptr));
    ZG_TRY_RC(zg_upload(&g->out_col, c.out_col));
    ZG_TRY_RC(zg_upload(&g->in_ptr, c.in_ptr));
    ZG_TRY_RC(zg_upload(&g->in_col, c.in_col));
    ZG_TRY_RC(zg_upload(&g->live, g->h_live));
    ZG_TRY_RC(zg_upload(&g->pr_id, c.pr_id));
    ZG_TRY_RC(zg_upload(&g->pr_col, c.pr_col));
    return ZG_OK;
}

static int zg_build(zg_graph *g, const uint8_t *live_mask, const uint32_t *from, const uint32_t *to,
                    const double *weight, uint64_t m)
{
    const uint32_t V = g->V;
    if (g->h_comp_off.empty()) g->h_comp_off = {0u, V};
    g->h_live.assign(V, 1);
    if (live_mask)
        for (uint32_t i = 0; i < V; i++) g->h_live[i] = live_mask[i] ? 1 : 0;
    uint32_t n_live = 0;
    for (uint32_t i = 0; i < V; i++) n_live += g->h_live[i];
    g->n_live = n_live;
    const uint64_t gpu_min = g_build_gpu_min_arcs;
    if (m >= gpu_min) ZG_TRY_RC(zg_build_csr_gpu(g, from, to, weight, m));
    else ZG_TRY_RC(zg_build_csr_host(g, from, to, weight, m));
    g->built_on_gpu = m >= gpu_min;
    // chunk lists, hub slots and PageRank row bins, on the device
    return zg_derive_tables(g);
}

static int zg_create_common(uint32_t n_cap, const uint8_t *live_mask, const uint32_t *from, const uint32_t *to,
                            const double *weight, uint64_t m, std::vector<uint32_t> comp_off, zg_graph **out)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return ZG_ERR_NO_DEVICE;
    }
    zg_graph *g = new (std::nothrow) zg_graph();
    if (!g) return ZG_ERR_OOM;
    g->V = n_cap;
    int rc = ZG_OK;
    hipError_t e = hipGetDevice(&g->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = zg_hip_code(e);
    if (rc == ZG_OK) {
        try {
            g->h_comp_off = std::move(comp_off);
            rc = zg_build(g, live_mask, from, to, weight, m);
        } catch (const std::bad_alloc &) {
            rc = ZG_ERR_OOM;
        }
    }
    if (rc != ZG_OK) {
        zg_free_device(g);
        delete g;
        return rc;
    }
    *out = g;
    return ZG_OK;
}

extern "C" int zg_graph_create(uint32_t n_cap, const uint8_t *live_mask, const uint32_t *from,
                               const uint32_t *to, uint64_t m, zg_graph **out)
{
    if (!out) return ZG_ERR_INVALID_ARG;
    *out = nullptr;
    if (m && (!from || !to)) return ZG_ERR_INVALID_ARG;
    if (m >= 0xFFFFFFFFull || n_cap == 0xFFFFFFFFu) return ZG_ERR_TOO_LARGE;
    return zg_create_common(n_cap, live_mask, from, to, nullptr, m, {}, out);
}

extern "C" int zg_graph_create_weighted(uint32_t n_cap, const uint8_t *live_mask, const uint32_t *from,
                                        const uint32_t *to, const double *weight, uint64_t m, zg_graph **out)
{
    if (!out) return ZG_ERR_INVALID_ARG;
    *out = nullptr;
    if (m && (!from || !to || !weight)) return ZG_ERR_INVALID_ARG;
    if (m >= 0xFFFFFFFFull || n_cap == 0xFFFFFFFFu) return ZG_ERR_TOO_LARGE;
    return zg_create_common(n_cap, live_mask, from, to, weight, m, {}, out);
}

extern "C" int zg_graph_create_batch(uint32_t n_graphs, const uint32_t *node_counts, const uint64_t *arc_counts,
                                     const uint32_t *from, const uint32_t *to, zg_graph **out)
{
    if (!out) return ZG_ERR_INVALID_ARG;
    *out = nullptr;
    if (n_graphs && (!node_counts || !arc_counts)) return ZG_ERR_INVALID_ARG;
    uint64_t V = 0, M = 0;
    for (uint32_t i = 0; i < n_graphs; i++) {
        V += node_counts[i];
        M += arc_counts[i];
    }
    if (M && (!from || !to)) return ZG_ERR_INVALID_ARG;
    if (M >= 0xFFFFFFFFull || V >= 0xFFFFFFFFull) return ZG_ERR_TOO_LARGE;
    std::vector<uint32_t> comp_off, gf, gt;
    try {
        comp_off.resize((size_t)n_graphs + 1);
        gf.resize(M);
        gt.resize(M);
        uint64_t vo = 0, mo = 0;
        for (uint32_t i = 0; i < n_graphs; i++) {
            comp_off[i] = (uint32_t)vo;
            for (uint64_t k = 0; k < arc_counts[i]; k++) {
                const uint32_t f = from[mo + k], t = to[mo + k];
                if (f >= node_counts[i] || t >= node_counts[i]) return ZG_ERR_NODE_NOT_FOUND;
                gf[mo + k] = (uint32_t)(vo + f);
                gt[mo + k] = (uint32_t)(vo + t);
            }
            vo += node_counts[i];
            mo += arc_counts[i];
        }
        comp_off[n_graphs] = (uint32_t)vo;
    } catch (const std::bad_alloc &) {
        return ZG_ERR_OOM;
    }
    return zg_create_common((uint32_t)V, nullptr, gf.data(), gt.data(), nullptr, M, std::move(comp_off), out);
}

extern "C" int zg_graph_components(const zg_graph *g, uint32_t *n_components)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (n_components) *n_components = (uint32_t)g->h_comp_off.size() - 1;
    return ZG_OK;
}

// ---------------------------------------------------------------------------
// The 64-lane engine's view: the same graph with every component's vertices
// renumbered by decreasing degree (in + out; ties by id).  The sweep gathers
// path-count and dependency rows, frontier and level lane masks of the
// vertices its arcs reach, most often those of the highest degrees; with them
// at the low ids those rows and masks share lines and stay in L2 (R-MAT 20,
// 4,096 sources: 177 -> 147 ms per step).  Components keep their id ranges,
// and every successor row keeps its arc order.

__global__ void lv_keys_kernel(uint32_t V, const uint32_t *out_ptr, const uint32_t *in_ptr, uint32_t *key,
                               uint32_t *val)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        key[v] = ~((out_ptr[v + 1] - out_ptr[v]) + (in_ptr[v + 1] - in_ptr[v]));  // ascending: decreasing degree
        val[v] = v;
    }
}

// component index of the vertex at each position (co: ncomp + 1 offsets)
__global__ void lv_comp_kernel(uint32_t V, const uint32_t *order, const uint32_t *co, uint32_t ncomp, uint32_t *key)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        const uint32_t v = order[i];
        uint32_t lo = 0, hi = ncomp;  // co[lo] <= v < co[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (co[mid] <= v) lo = mid;
            else hi = mid;
        }
        key[i] = lo;
    }
}

__global__ void lv_perm_kernel(uint32_t V, const uint32_t *order, uint32_t *perm)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) perm[order[i]] = i;
}

// The view's arcs in the order that makes the build give every row its arc
// order back: the build takes arcs newest (last) first, so arc i is the
// successor CSR's arc E-1-i.
__global__ void lv_arcs_kernel(uint32_t V, uint64_t E, const uint32_t *out_ptr, const uint32_t *out_col,
                               const uint32_t *perm, uint32_t *from, uint32_t *to)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < V; u += gridDim.x * blockDim.x) {
        const uint32_t pu = perm[u];
        for (uint32_t e = out_ptr[u]; e < out_ptr[u + 1]; e++) {
            const uint64_t i = E - 1 - e;
            from[i] = pu;
            to[i] = perm[out_col[e]];
        }
    }
}

int zg_lane_view(zg_graph *g, hipStream_t s)
{
    if (g->lane || !g->lane_order || g->V == 0 || g->E == 0) return ZG_OK;
    const uint32_t V = g->V;
    const uint64_t E = g->E;
    const dim3 blk(256), grd((uint32_t)std::min<uint64_t>((V + 255) / 256, 8192));
    int rc = ZG_OK;
    {
        ZgScratch sc;
        uint32_t *key, *key2, *val, *order, *from, *to, *co = nullptr;
        ZG_TRY_RC(sc.get(&key, V));
        ZG_TRY_RC(sc.get(&key2, V));
        ZG_TRY_RC(sc.get(&val, V));
        ZG_TRY_RC(sc.get(&order, V));
        ZG_TRY_RC(zg_dmalloc(&g->lane_perm, V));
        hipLaunchKernelGGL(lv_keys_kernel, grd, blk, 0, s, V, g->out_ptr, g->in_ptr, key, val);
        ZG_TRY(hipGetLastError());
        ZG_TRY_RC(zg_sort_pairs(key, key2, val, order, V, 32, s, sc));
        const uint32_t ncomp = (uint32_t)g->h_comp_off.size() - 1;
        if (ncomp > 1) {
            // stable by component: degree order within each component's range
            ZG_TRY_RC(sc.get(&co, (size_t)ncomp + 1));
            ZG_TRY(hipMemcpyAsync(co, g->h_comp_off.data(), sizeof(uint32_t) * (ncomp + 1), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(lv_comp_kernel, grd, blk, 0, s, V, order, co, ncomp, key);
            ZG_TRY(hipGetLastError());
            ZG_TRY_RC(zg_sort_pairs(key, key2, order, val, V, 32, s, sc));
            std::swap(order, val);
        }
        hipLaunchKernelGGL(lv_perm_kernel, grd, blk, 0, s, V, order, g->lane_perm);
        ZG_TRY_RC(sc.get(&from, E));
        ZG_TRY_RC(sc.get(&to, E));
        hipLaunchKernelGGL(lv_arcs_kernel, grd, blk, 0, s, V, E, g->out_ptr, g->out_col, g->lane_perm, from, to);
        ZG_TRY(hipGetLastError());
        g->h_lane_perm.resize(V);
        ZG_TRY(hipMemcpyAsync(g->h_lane_perm.data(), g->lane_perm, sizeof(uint32_t) * V, hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));

        zg_graph *c = new (std::nothrow) zg_graph();
        if (!c) return ZG_ERR_OOM;
        c->V = V;
        c->device = g->device;
        c->h_comp_off = g->h_comp_off;
        c->h_live.assign(V, 0);
        for (uint32_t v = 0; v < V; v++) c->h_live[g->h_lane_perm[v]] = g->h_live[v];
        c->n_live = g->n_live;
        rc = zg_hip_code(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        if (rc == ZG_OK) rc = zg_build_csr_gpu(c, from, to, nullptr, E, true);
        c->built_on_gpu = true;
        if (rc == ZG_OK) rc = zg_derive_tables(c);
        if (rc == ZG_OK && c->E != E) rc = ZG_ERR_INVALID_ARG;  // every arc of the view is live
        if (rc != ZG_OK) {
            (void)hipStreamSynchronize(c->stream);
            zg_free_device(c);
            delete c;
        } else {
            g->lane = c;
        }
    }
    if (rc == ZG_ERR_GPU_OOM) {
        // no room for the view: the engine runs on the resource itself
        zg_dfree(g->lane_perm);
        g->h_lane_perm.clear();
        g->lane_order = false;
        return ZG_OK;
    }
    return rc;
}

extern "C" void zg_graph_destroy(zg_graph *g)
{
    if (!g) return;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        (void)hipSetDevice(g->device);
        if (g->stream) (void)hipStreamSynchronize(g->stream);
        zg_free_device(g);
    }
    delete g;
}

extern "C" int zg_graph_info(const zg_graph *g, uint32_t *n_cap, uint32_t *n_live, uint64_t *n_arcs)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (n_cap) *n_cap = g->V;
    if (n_live) *n_live = g->n_live;
    if (n_arcs) *n_arcs = g->E;
    return ZG_OK;
}

extern "C" int zg_graph_csr(zg_graph *g, uint32_t *out_ptr, uint32_t *out_col, uint32_t *in_ptr, uint32_t *in_col,
                            uint32_t *pr_id)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    const size_t V1 = (size_t)g->V + 1, E = g->E;
    if (out_ptr) ZG_TRY(hipMemcpy(out_ptr, g->out_ptr, sizeof(uint32_t) * V1, hipMemcpyDeviceToHost));
    if (out_col && E) ZG_TRY(hipMemcpy(out_col, g->out_col, sizeof(uint32_t) * E, hipMemcpyDeviceToHost));
    if (in_ptr) ZG_TRY(hipMemcpy(in_ptr, g->in_ptr, sizeof(uint32_t) * V1, hipMemcpyDeviceToHost));
    if (in_col && E) ZG_TRY(hipMemcpy(in_col, g->in_col, sizeof(uint32_t) * E, hipMemcpyDeviceToHost));
    if (pr_id && g->V) ZG_TRY(hipMemcpy(pr_id, g->pr_id, sizeof(uint32_t) * g->V, hipMemcpyDeviceToHost));
    return g->built_on_gpu ? 1 : 0;
}

extern "C" int zg_graph_table(zg_graph *g, int which, uint32_t *out, uint64_t cap, uint64_t *n)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!n) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    const uint64_t V = g->V, nsm = (uint64_t)g->n_small + g->n_med;
    struct Part {
        const void *src;
        uint64_t words;
    };
    std::vector<Part> parts;
    std::vector<uint32_t> bins;
    switch (which) {
    case ZG_TABLE_ICH: parts = {{g->ich_v, g->n_ich}, {g->ich_e0, g->n_ich}, {g->ich_e1, g->n_ich}}; break;
    case ZG_TABLE_ICH_FIRST: parts = {{g->ich_first, V + 1}}; break;
    case ZG_TABLE_OCH: parts = {{g->och_v, g->n_och}, {g->och_e0, g->n_och}, {g->och_e1, g->n_och}}; break;
    case ZG_TABLE_OCH_FIRST: parts = {{g->och_first, V + 1}}; break;
    case ZG_TABLE_OHUB: parts = {{g->ohub_idx, V}}; break;
    case ZG_TABLE_PR_ROWS: parts = {{g->pr_small, g->n_small}, {g->pr_med, g->n_med}}; break;
    case ZG_TABLE_PR_DESC: parts = {{g->pr_desc, 4 * nsm}}; break;
    case ZG_TABLE_PR_LARGE:
        parts = {{g->pr_large, g->n_large}, {g->pr_large_deg, g->n_large}, {g->pr_large_seg0, (uint64_t)g->n_large + 1}};
        break;
    case ZG_TABLE_PR_SEG: parts = {{g->pr_seg_row, g->n_seg}, {g->pr_seg_e0, g->n_seg}, {g->pr_seg_e1, g->n_seg}}; break;
    case ZG_TABLE_PR_COL: parts = {{g->pr_col, g->E}}; break;
    case ZG_TABLE_PR_BINS: bins = {g->n_small, g->n_m4, g->n_m16, g->n_m64, g->n_large}; break;
    default: return ZG_ERR_INVALID_ARG;
    }
    uint64_t total = bins.size();
    for (const Part &p : parts) total += p.words;
    *n = total;
    if (!out) return ZG_OK;
    if (cap < total) return ZG_ERR_INVALID_ARG;
    ZG_TRY(hipSetDevice(g->device));
    for (size_t i = 0; i < bins.size(); i++) out[i] = bins[i];
    uint64_t at = 0;
    for (const Part &p : parts) {
        if (p.words) ZG_TRY(hipMemcpy(out + at, p.src, sizeof(uint32_t) * p.words, hipMemcpyDeviceToHost));
        at += p.words;
    }
    return ZG_OK;
}

extern "C" int zg_node_degrees(zg_graph *g, uint32_t *out)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out && g->V) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    // arcs of tombstoned endpoints are not in the successor CSR, as removeNode
    // deletes a node's outgoing and incoming edges (array_graph.zig:128-153)
    ZG_TRY(hipSetDevice(g->device));
    std::vector<uint32_t> ptr((size_t)g->V + 1);
    ZG_TRY(hipMemcpy(ptr.data(), g->out_ptr, sizeof(uint32_t) * ptr.size(), hipMemcpyDeviceToHost));
    for (uint32_t v = 0; v < g->V; v++) out[v] = ptr[(size_t)v + 1] - ptr[v];
    return ZG_OK;
}

extern "C" int zg_set_tuning(zg_graph *g, int param, double value)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    std::lock_guard<std::mutex> lk(g->mu);
    switch (param) {
    case ZG_TUNE_PUSH_ALPHA:
        if (!(value >= 0.0)) return ZG_ERR_INVALID_ARG;
        g->push_alpha = value;
        return ZG_OK;
    case ZG_TUNE_ENGINE:
        if (value != ZG_ENGINE_AUTO && value != ZG_ENGINE_LANES && value != ZG_ENGINE_QUEUE) return ZG_ERR_INVALID_ARG;
        g->engine = (int)value;
        return ZG_OK;
    case ZG_TUNE_PR_CACHE:
        if (!(value >= 0.0 && value <= 4294967295.0)) return ZG_ERR_INVALID_ARG;
        g->pr_xcache = (uint32_t)value;
        return ZG_OK;
    case ZG_TUNE_PR_GRID:
        if (!(value >= 0.0 && value <= 65536.0)) return ZG_ERR_INVALID_ARG;
        g->pr_grid_cap = (uint32_t)value;
        (void)zg_pr_ws_release(g);  // partial buffers are sized by the grid
        return ZG_OK;
    case ZG_TUNE_GRAPH:
        if (value != 0.0 && value != 1.0) return ZG_ERR_INVALID_ARG;
        g->use_graph = value != 0.0;
        return ZG_OK;
    case ZG_TUNE_WEIGHTED_BATCH:
        if (!(value >= 0.0 && value <= 2048.0) || value != (double)(64 * (uint32_t)(value / 64)))
            return ZG_ERR_INVALID_ARG;
        g->wbc_batch = (uint32_t)value;
        return ZG_OK;
    case ZG_TUNE_WEIGHTED_DELTA:
        if (!(value >= 0.0)) return ZG_ERR_INVALID_ARG;  // +inf allowed: no bound
        g->wbc_delta = value;
        return ZG_OK;
    case ZG_TUNE_WEIGHTED_GRID:
        if (!(value >= 0.0 && value <= 65536.0) || value != (double)(uint32_t)value) return ZG_ERR_INVALID_ARG;
        g->wbc_grid = (uint32_t)value;
        return ZG_OK;
    case ZG_TUNE_COUNT_TREE_ARCS:
        if (value != 0.0 && value != 1.0) return ZG_ERR_INVALID_ARG;
        g->count_tree = value != 0.0;
        return ZG_OK;
    case ZG_TUNE_QUEUE_ORDER:
        if (value != 0.0 && value != 1.0) return ZG_ERR_INVALID_ARG;
        if ((value != 0.0) != g->queue_order) {
            ZG_TRY(hipSetDevice(g->device));
            ZG_TRY(hipStreamSynchronize(g->stream));
            zg_bcq_order_release(g);  // rebuilt, in the other numbering, by the next sweep
        }
        g->queue_order = value != 0.0;
        return ZG_OK;
    case ZG_TUNE_PULL_GRID:
        if (!(value >= 0.0 && value <= 65536.0) || value != (double)(uint32_t)value) return ZG_ERR_INVALID_ARG;
        g->pull_grid = (uint32_t)value;
        return ZG_OK;
    case ZG_TUNE_LANE_ORDER:
        if (value != 0.0 && value != 1.0) return ZG_ERR_INVALID_ARG;
        g->lane_order = value != 0.0;
        return ZG_OK;
    case ZG_TUNE_PR_BLOCK:
        if (value != 0.0 && value != 256.0 && value != 512.0 && value != 1024.0) return ZG_ERR_INVALID_ARG;
        g->pr_block = (uint32_t)value;
        return ZG_OK;
    default: return ZG_ERR_INVALID_ARG;
    }
}

extern "C" int zg_set_source_order(zg_graph *g, int mode)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (mode < 0 || mode > 2) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    int prev = g->source_order;
    g->source_order = mode;
    return prev;
}

extern "C" int zg_set_batch_groups(zg_graph *g, int groups)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (groups < 1 || groups > ZG_MAX_GROUPS) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    int prev = g->groups;
    if (groups != g->groups) {
        (void)hipSetDevice(g->device);
        zg_bc_ws_release(g);
        zg_bcq_ws_release(g);
        g->groups = groups;
    }
    return prev;
}
