// Node features for the learned pivot scorer, computed on the resident graph.
//
// Per node: log1p(in + out degree), log1p(in degree), log1p(out degree),
// log1p(sum of the successors' out-degrees) (2-hop), log(n * PageRank), and
// log1p(core number), each standardised over the live nodes (zero mean, unit
// variance).  Column 6 is 1.0 and column 7 is 0.0, so a row is 8 floats (one
// K-step of the MLP's first MFMA layer).
//
// Core numbers follow coreNumbers in priv/zog/src/connectivity.zig:5-101: the
// graph is taken as undirected and simple (self-loops dropped, each neighbour
// once), and a node's core number is the k at which peeling removes it.  The
// GPU peels level-synchronously: at level k every live node of remaining
// degree <= k is removed in one round (core = k) and its neighbours' degrees
// drop; rounds repeat until none is left at k, then k grows.  This yields the
// same core numbers as the reference's bucket order (the k-core is unique).
#include "common.h"

#include <algorithm>
#include <cmath>

// Undirected simple adjacency (CSR) for core numbers, built once per handle.
static int zg_undirected(zg_graph *g, hipStream_t s)
{
    if (g->und_ptr || g->V == 0) return ZG_OK;
    const uint32_t V = g->V;
    std::vector<uint32_t> out_ptr(V + 1), out_col(g->E);
    ZG_TRY(hipMemcpy(out_ptr.data(), g->out_ptr, sizeof(uint32_t) * (V + 1), hipMemcpyDeviceToHost));
    if (g->E) ZG_TRY(hipMemcpy(out_col.data(), g->out_col, sizeof(uint32_t) * g->E, hipMemcpyDeviceToHost));
    std::vector<uint64_t> keys;
    keys.reserve(2 * (size_t)g->E);
    for (uint32_t u = 0; u < V; u++)
        for (uint32_t k = out_ptr[u]; k < out_ptr[u + 1]; k++) {
            const uint32_t v = out_col[k];
            if (u == v) continue;
            keys.push_back(((uint64_t)u << 32) | v);
            keys.push_back(((uint64_t)v << 32) | u);
        }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<uint32_t> ptr(V + 1, 0), col(keys.size());
    for (size_t i = 0; i < keys.size(); i++) {
        ptr[(keys[i] >> 32) + 1]++;
        col[i] = (uint32_t)keys[i];
    }
    for (uint32_t v = 0; v < V; v++) ptr[v + 1] += ptr[v];
    ZG_TRY_RC(zg_dmalloc(&g->und_ptr, V + 1));
    ZG_TRY_RC(zg_dmalloc(&g->und_col, col.size()));
    ZG_TRY(hipMemcpy(g->und_ptr, ptr.data(), sizeof(uint32_t) * (V + 1), hipMemcpyHostToDevice));
    if (!col.empty()) ZG_TRY(hipMemcpy(g->und_col, col.data(), sizeof(uint32_t) * col.size(), hipMemcpyHostToDevice));
    return ZG_OK;
}

__global__ void core_init_kernel(uint32_t V, const uint32_t *ptr, const uint8_t *live, uint32_t *deg, uint32_t *core)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    deg[v] = live[v] ? ptr[v + 1] - ptr[v] : 0;
    core[v] = live[v] ? ZG_NONE : 0;
}

// One peeling round at level k: mark (phase 0) then update neighbours (phase 1).
__global__ void core_mark_kernel(uint32_t V, const uint32_t *deg, uint32_t *core, uint8_t *removed_now, uint32_t k,
                                 uint32_t *count)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    bool r = false;
    if (v < V) {
        r = core[v] == ZG_NONE && deg[v] <= k;
        removed_now[v] = r ? 1 : 0;
        if (r) core[v] = k;
    }
    const uint64_t b = __ballot(r);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (uint32_t)__popcll(b));
}

__global__ void core_update_kernel(uint32_t V, const uint32_t *ptr, const uint32_t *col, uint32_t *deg,
                                   const uint32_t *core, const uint8_t *removed_now)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || !removed_now[v]) return;
    for (uint32_t e = ptr[v], e1 = ptr[v + 1]; e < e1; e++) {
        const uint32_t u = col[e];
        if (core[u] == ZG_NONE) atomicSub(&deg[u], 1u);
    }
}

static int zg_core_numbers_dev(zg_graph *g, uint32_t *d_core, hipStream_t s)
{
    const uint32_t V = g->V;
    if (V == 0) return ZG_OK;
    ZG_TRY_RC(zg_undirected(g, s));
    uint32_t *deg = nullptr, *cnt = nullptr;
    uint8_t *rem = nullptr;
    int rc = zg_dmalloc(&deg, V);
    if (rc == ZG_OK) rc = zg_dmalloc(&cnt, 1);
    if (rc == ZG_OK) rc = zg_dmalloc(&rem, V);
    const dim3 grid((V + 255) / 256), blk(256);
    if (rc == ZG_OK) {
        hipLaunchKernelGGL(core_init_kernel, grid, blk, 0, s, V, g->und_ptr, g->live, deg, d_core);
        rc = zg_hip_code(hipGetLastError());
    }
    uint32_t remaining = g->n_live, k = 0;
    while (rc == ZG_OK && remaining > 0) {
        uint32_t c = 0;
        rc = zg_hip_code(hipMemsetAsync(cnt, 0, sizeof(uint32_t), s));
        if (rc != ZG_OK) break;
        hipLaunchKernelGGL(core_mark_kernel, grid, blk, 0, s, V, deg, d_core, rem, k, cnt);
        hipLaunchKernelGGL(core_update_kernel, grid, blk, 0, s, V, g->und_ptr, g->und_col, deg, d_core, rem);
        rc = zg_hip_code(hipGetLastError());
        if (rc == ZG_OK) rc = zg_hip_code(hipMemcpyAsync(&c, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (rc == ZG_OK) rc = zg_hip_code(hipStreamSynchronize(s));
        if (rc != ZG_OK) break;
        if (c == 0) k++;  // nothing left at this level: next core
        else remaining -= c;
    }
    zg_dfree(deg);
    zg_dfree(cnt);
    zg_dfree(rem);
    return rc;
}

extern "C" int zg_core_numbers(zg_graph *g, uint32_t *out)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out && g->V) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    uint32_t *d = nullptr;
    ZG_TRY_RC(zg_dmalloc(&d, g->V));
    int rc = zg_core_numbers_dev(g, d, g->stream);
    if (rc == ZG_OK && g->V) rc = zg_hip_code(hipMemcpy(out, d, sizeof(uint32_t) * g->V, hipMemcpyDeviceToHost));
    zg_dfree(d);
    return rc;
}

// Raw features (before standardisation) of one node per thread.
__global__ void feat_raw_kernel(uint32_t V, const uint32_t *out_ptr, const uint32_t *out_col, const uint32_t *in_ptr,
                                const uint8_t *live, const double *pr, const uint32_t *core, uint32_t n_live,
                                float *f)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    float *row = f + (size_t)v * ZG_NFEAT;
    if (!live[v]) {
        for (int k = 0; k < ZG_NFEAT; k++) row[k] = 0.f;
        return;
    }
    const uint32_t od = out_ptr[v + 1] - out_ptr[v], id = in_ptr[v + 1] - in_ptr[v];
    double two = 0.0;
    for (uint32_t e = out_ptr[v]; e < out_ptr[v + 1]; e++) {
        const uint32_t w = out_col[e];
        two += (double)(out_ptr[w + 1] - out_ptr[w]);
    }
    row[0] = (float)log1p((double)(od + id));
    row[1] = (float)log1p((double)id);
    row[2] = (float)log1p((double)od);
    row[3] = (float)log1p(two);
    row[4] = (float)log(fmax(pr[v] * (double)n_live, 1e-30));
    row[5] = (float)log1p((double)core[v]);
    row[6] = 1.f;
    row[7] = 0.f;
}

// Sums and sums of squares of the first 6 columns over live nodes.
__global__ void feat_moments_kernel(uint32_t V, const uint8_t *live, const float *f, double *mom)
{
    __shared__ double red[12][4];
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double x[6] = {0, 0, 0, 0, 0, 0};
    const bool ok = v < V && live[v];
    for (int k = 0; k < 6; k++) x[k] = ok ? (double)f[(size_t)v * ZG_NFEAT + k] : 0.0;
    for (int k = 0; k < 12; k++) {
        double t = k < 6 ? x[k] : x[k - 6] * x[k - 6];
        t = zg_wave_sum(t);
        if (lane == 0) red[k][wv] = t;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[threadIdx.x][w];
        atomicAdd(&mom[threadIdx.x], t);
    }
}

__global__ void feat_standardize_kernel(uint32_t V, const uint8_t *live, const double *mom, uint32_t n, float *f)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || !live[v]) return;
    for (int k = 0; k < 6; k++) {
        const double mean = mom[k] / n;
        const double var = fmax(mom[6 + k] / n - mean * mean, 0.0);
        const double sd = var > 1e-12 ? sqrt(var) : 1.0;
        float *x = f + (size_t)v * ZG_NFEAT + k;
        *x = (float)(((double)*x - mean) / sd);
    }
}

int zg_pagerank_dev(zg_graph *g, double damping, uint64_t max_iterations, double tolerance, int precision,
                    double *d_out, hipStream_t s, zg_pr_info *info);

extern "C" int zg_node_features(zg_graph *g, float *d_features, void *stream)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!d_features && g->V) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    const uint32_t V = g->V;
    if (V == 0) return ZG_OK;
    double *pr = nullptr, *mom = nullptr;
    uint32_t *core = nullptr;
    int rc = zg_dmalloc(&pr, V);
    if (rc == ZG_OK) rc = zg_dmalloc(&core, V);
    if (rc == ZG_OK) rc = zg_dmalloc(&mom, 12);
    if (rc == ZG_OK) rc = zg_pagerank_dev(g, 0.85, 100, 1e-6, ZG_PR_FP32 | ZG_PR_SKIP_RESIDUAL, pr, s, nullptr);
    if (rc == ZG_OK) rc = zg_core_numbers_dev(g, core, s);
    const dim3 grid((V + 255) / 256), blk(256);
    if (rc == ZG_OK) {
        hipLaunchKernelGGL(feat_raw_kernel, grid, blk, 0, s, V, g->out_ptr, g->out_col, g->in_ptr, g->live, pr, core,
                           g->n_live, d_features);
        rc = zg_hip_code(hipMemsetAsync(mom, 0, sizeof(double) * 12, s));
    }
    if (rc == ZG_OK) {
        hipLaunchKernelGGL(feat_moments_kernel, grid, blk, 0, s, V, g->live, d_features, mom);
        hipLaunchKernelGGL(feat_standardize_kernel, grid, blk, 0, s, V, g->live, mom, g->n_live ? g->n_live : 1,
                           d_features);
        rc = zg_hip_code(hipGetLastError());
    }
    if (rc == ZG_OK) rc = zg_hip_code(hipStreamSynchronize(s));
    zg_dfree(pr);
    zg_dfree(core);
    zg_dfree(mom);
    return rc;
}
