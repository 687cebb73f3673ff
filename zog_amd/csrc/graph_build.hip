// Resource build on the GPU: the successor and predecessor CSR of buildGraph
// (lib/zog/resource_graph.ex:120-129) from the uploaded arc arrays, for
// graphs large enough that the host loops would dominate the load.
//
//   successor CSR  arcs with a tombstoned endpoint dropped (the iterator
//                  skips them, array_graph.zig:166-180); within a row the
//                  newest arc first (addEdge prepends, :88-108): a stable
//                  radix sort by source of the arcs taken in reverse order.
//   predecessor    for every source in id order, its successors in row
//   CSR            order (centrality.zig:881-899, utils.zig:52-73): a
//                  stable radix sort by head of the successor CSR's arcs.
//   PageRank ids   vertices by decreasing out-degree, ties by id: a stable
//                  sort by the complemented degree.
// The work tables derived from these (graph_derive.hip) are built on the
// device too; the host reads back only the arc count and error flags.
#include "build_util.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

namespace {

__global__ void gb_check_kernel(uint64_t m, uint32_t V, const uint32_t *from, const uint32_t *to, uint32_t *bad)
{
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x)
        if (from[e] >= V || to[e] >= V) atomicOr(bad, 1u);
}

// Arc i of the sort input is arc m-1-i (newest first); key = source, or V for
// an arc with a tombstoned endpoint (sorted after every row, then dropped).
__global__ void gb_keys_kernel(uint64_t m, uint32_t V, const uint32_t *from, const uint32_t *to, const uint8_t *live,
                               uint32_t *key, uint32_t *val, uint32_t *deg)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = (uint32_t)(m - 1 - i), f = from[e], t = to[e];
        const bool ok = live[f] && live[t];
        key[i] = ok ? f : V;
        val[i] = e;
        if (ok) atomicAdd(&deg[f], 1u);
    }
}

__global__ void gb_gather_kernel(uint32_t E, const uint32_t *arc, const uint32_t *to, const double *w, uint32_t *out_col,
                                 double *out_w, uint32_t *indeg)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < E; k += gridDim.x * blockDim.x) {
        const uint32_t e = arc[k], t = to[e];
        out_col[k] = t;
        if (w) out_w[k] = w[e];
        atomicAdd(&indeg[t], 1u);
    }
}

__global__ void gb_degree_keys_kernel(uint32_t V, const uint32_t *out_ptr, uint32_t *key, uint32_t *val)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        key[v] = ~(out_ptr[v + 1] - out_ptr[v]);  // ascending order of ~deg = decreasing degree
        val[v] = v;
    }
}

__global__ void gb_scatter_id_kernel(uint32_t V, const uint32_t *order, uint32_t *pr_id)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) pr_id[order[i]] = i;
}

__global__ void gb_relabel_kernel(uint32_t E, const uint32_t *in_col, const uint32_t *pr_id, uint32_t *pr_col)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < E; k += gridDim.x * blockDim.x)
        pr_col[k] = pr_id[in_col[k]];
}

// Sum of the weights and whether any is NaN, negative or infinite.
__global__ void gb_weight_kernel(uint32_t E, const double *w, double *sum, uint32_t *bad)
{
    __shared__ double red[4];
    double s = 0.0;
    bool b = false;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < E; k += gridDim.x * blockDim.x) {
        const double x = w[k];
        s += x;
        b |= !(x >= 0.0) || x == HUGE_VAL;
    }
    s = zg_wave_sum(s);
    if (__ballot(b) && (threadIdx.x & 63) == 0) atomicOr(bad, 1u);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(sum, red[0] + red[1] + red[2] + red[3]);
}

unsigned bits_for(uint64_t x)  // bits that hold every value <= x
{
    unsigned b = 1;
    while (b < 32 && (x >> b)) b++;
    return b;
}

}  // namespace

int zg_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s, ZgScratch &sc)
{
    if (n == 0) return ZG_OK;
    size_t bytes = 0;
    ZG_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s));
    void *tmp = nullptr;
    ZG_TRY_RC(sc.get((char **)&tmp, bytes ? bytes : 1));
    ZG_TRY(rocprim::exclusive_scan(tmp, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s));
    return ZG_OK;
}

size_t zg_scan_bytes(size_t n)
{
    size_t bytes = 0;
    if (n == 0) return 0;
    if (rocprim::exclusive_scan(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, 0u, n,
                                rocprim::plus<uint32_t>(), (hipStream_t)0) != hipSuccess)
        return 0;
    return bytes;
}

int zg_exclusive_scan_tmp(const uint32_t *in, uint32_t *out, size_t n, void *tmp, size_t cap, hipStream_t s)
{
    if (n == 0) return ZG_OK;
    size_t bytes = 0;
    ZG_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s));
    if (bytes > cap) return ZG_ERR_INVALID_ARG;
    ZG_TRY(rocprim::exclusive_scan(tmp, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s));
    return ZG_OK;
}

int zg_sort_pairs(const uint32_t *kin, uint32_t *kout, const uint32_t *vin, uint32_t *vout, uint64_t n, unsigned bits,
                  hipStream_t s, ZgScratch &sc)
{
    if (n == 0) return ZG_OK;
    size_t bytes = 0;
    ZG_TRY(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0u, bits, s));
    void *tmp = nullptr;
    ZG_TRY_RC(sc.get((char **)&tmp, bytes ? bytes : 1));
    ZG_TRY(rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, (size_t)n, 0u, bits, s));
    return ZG_OK;
}

// Fills the device arrays out_ptr, out_col, (out_w), in_ptr, in_col, pr_id,
// pr_col, live and g->E, g->wmean, g->weights_ok.
int zg_build_csr_gpu(zg_graph *g, const uint32_t *from, const uint32_t *to, const double *weight, uint64_t m,
                     bool device_arcs)
{
    const uint32_t V = g->V;
    hipStream_t s = g->stream;
    ZgScratch sc;
    const dim3 blk(256), grd(4096);
    uint32_t *d_from, *d_to, *d_bad, *key, *key2, *val, *val2, *deg;
    double *d_w = nullptr;
    ZG_TRY_RC(sc.get(&d_from, m));
    ZG_TRY_RC(sc.get(&d_to, m));
    ZG_TRY_RC(sc.get(&d_bad, 2));
    const hipMemcpyKind kind = device_arcs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (m) {
        ZG_TRY(hipMemcpyAsync(d_from, from, sizeof(uint32_t) * m, kind, s));
        ZG_TRY(hipMemcpyAsync(d_to, to, sizeof(uint32_t) * m, kind, s));
    }
    if (weight) {
        ZG_TRY_RC(sc.get(&d_w, m));
        if (m) ZG_TRY(hipMemcpyAsync(d_w, weight, sizeof(double) * m, kind, s));
    }
    ZG_TRY_RC(zg_dmalloc(&g->live, V));
    if (V) ZG_TRY(hipMemcpyAsync(g->live, g->h_live.data(), V, hipMemcpyHostToDevice, s));
    ZG_TRY(hipMemsetAsync(d_bad, 0, 2 * sizeof(uint32_t), s));
    hipLaunchKernelGGL(gb_check_kernel, grd, blk, 0, s, m, V, d_from, d_to, d_bad);
    uint32_t bad = 0;
    ZG_TRY(hipMemcpyAsync(&bad, d_bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    if (bad) return ZG_ERR_NODE_NOT_FOUND;  // addEdge: an id >= n_cap

    // successor CSR; the sort buffers also serve the PageRank-id sort of the
    // V vertices, so they hold max(m, V) entries
    const size_t nsort = std::max<size_t>((size_t)m, (size_t)V);
    ZG_TRY_RC(sc.get(&key, nsort));
    ZG_TRY_RC(sc.get(&key2, nsort));
    ZG_TRY_RC(sc.get(&val, nsort));
    ZG_TRY_RC(sc.get(&val2, nsort));
    ZG_TRY_RC(sc.get(&deg, (size_t)V + 1));
    ZG_TRY(hipMemsetAsync(deg, 0, sizeof(uint32_t) * ((size_t)V + 1), s));
    hipLaunchKernelGGL(gb_keys_kernel, grd, blk, 0, s, m, V, d_from, d_to, g->live, key, val, deg);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(zg_sort_pairs(key, key2, val, val2, m, bits_for(V), s, sc));
    ZG_TRY_RC(zg_dmalloc(&g->out_ptr, (size_t)V + 1));
    ZG_TRY_RC(zg_exclusive_scan(deg, g->out_ptr, (size_t)V + 1, s, sc));
    uint32_t E = 0;
    ZG_TRY(hipMemcpyAsync(&E, g->out_ptr + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    g->E = E;
    ZG_TRY_RC(zg_dmalloc(&g->out_col, E));
    if (weight) ZG_TRY_RC(zg_dmalloc(&g->out_w, E));
    ZG_TRY(hipMemsetAsync(deg, 0, sizeof(uint32_t) * ((size_t)V + 1), s));
    if (E) hipLaunchKernelGGL(gb_gather_kernel, grd, blk, 0, s, E, val2, d_to, d_w, g->out_col, g->out_w, deg);
    ZG_TRY(hipGetLastError());

    // predecessor CSR: heads of the successor arcs, stable, sources as values
    ZG_TRY_RC(zg_dmalloc(&g->in_ptr, (size_t)V + 1));
    ZG_TRY_RC(zg_exclusive_scan(deg, g->in_ptr, (size_t)V + 1, s, sc));
    ZG_TRY_RC(zg_dmalloc(&g->in_col, E));
    ZG_TRY_RC(zg_sort_pairs(g->out_col, key, key2, g->in_col, E, bits_for(V ? V - 1 : 0), s, sc));

    // PageRank ids
    ZG_TRY_RC(zg_dmalloc(&g->pr_id, V));
    ZG_TRY_RC(zg_dmalloc(&g->pr_col, E));
    if (V) {
        hipLaunchKernelGGL(gb_degree_keys_kernel, grd, blk, 0, s, V, g->out_ptr, key, val);
        ZG_TRY_RC(zg_sort_pairs(key, key2, val, val2, V, 32, s, sc));
        hipLaunchKernelGGL(gb_scatter_id_kernel, grd, blk, 0, s, V, val2, g->pr_id);
    }
    if (E) hipLaunchKernelGGL(gb_relabel_kernel, grd, blk, 0, s, E, g->in_col, g->pr_id, g->pr_col);
    ZG_TRY(hipGetLastError());

    // weights: mean, and whether every live arc's weight is usable
    if (weight) {
        double *d_sum;
        ZG_TRY_RC(sc.get(&d_sum, 1));
        ZG_TRY(hipMemsetAsync(d_sum, 0, sizeof(double), s));
        ZG_TRY(hipMemsetAsync(d_bad, 0, sizeof(uint32_t), s));
        if (E) hipLaunchKernelGGL(gb_weight_kernel, grd, blk, 0, s, E, g->out_w, d_sum, d_bad);
        double sum = 0.0;
        ZG_TRY(hipMemcpyAsync(&sum, d_sum, sizeof(double), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipMemcpyAsync(&bad, d_bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        g->wmean = E ? sum / (double)E : 0.0;
        g->weights_ok = bad == 0;
    }
    ZG_TRY(hipStreamSynchronize(s));
    return ZG_OK;
}
