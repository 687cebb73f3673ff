// Work tables of a resource, derived on the device from its resident CSR, so
// that one load (buildGraph, lib/zog/resource_graph.ex:120-129) runs on the
// GPU end to end and serves every later call.  The host reads back a few
// counts and, per graph of the resource, the tables' values at the graph's
// first vertex (the lane engine sizes and bounds its per-group work by
// graph).
//
//   predecessor chunks  every live vertex with predecessors: its predecessor
//                       row cut into ZG_CHUNK-arc pieces (forward-sweep work
//                       items); ich_first[v] = first chunk of v
//   successor chunks    every vertex: its successor row cut likewise, one
//                       empty chunk for a vertex without successors (the
//                       backward sweep still visits a leaf); och_first[v]
//   hub slots           a vertex with several successor chunks owns one
//                       partial-sum slot per chunk: ohub_idx[v] = first slot
//   PageRank row bins   live rows by predecessor count (pr_bins.h), each bin
//                       in PageRank-id order -- or in vertex order when the
//                       ids follow the vertices for most of them (a grid) --
//                       as a stable sort of the rows by bin; the predecessor
//                       columns copied into that row order (each row keeps
//                       its arc order, so its sum is unchanged); rows longer
//                       than ZG_PR_MED cut into ZG_PR_SEG-arc segments.
// Every table is a scan or a stable sort of per-vertex counts, so it equals
// the sequential definition entry for entry.
#include "build_util.h"
#include "pr_bins.h"

#include <algorithm>
#include <vector>

namespace {

__device__ __forceinline__ uint32_t dv_cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Per-vertex chunk counts (entry V = 0, so scans over V + 1 entries end in
// the totals), the largest successor-chunk count and the live vertices
// without successors.
__global__ void dv_count_kernel(uint32_t V, const uint32_t *in_ptr, const uint32_t *out_ptr, const uint8_t *live,
                                uint32_t *ich_n, uint32_t *och_n, uint32_t *hub_n, uint32_t *slot_n, uint32_t *maxk,
                                unsigned long long *ndang)
{
    unsigned long long dang = 0;
    uint32_t kmax = 1;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x) {
        if (v == V) {
            ich_n[v] = och_n[v] = hub_n[v] = slot_n[v] = 0;
            continue;
        }
        const uint32_t ideg = in_ptr[v + 1] - in_ptr[v], odeg = out_ptr[v + 1] - out_ptr[v];
        const bool lv = live[v] != 0;
        ich_n[v] = lv && ideg ? dv_cdiv(ideg, ZG_CHUNK) : 0u;
        const uint32_t k = odeg ? dv_cdiv(odeg, ZG_CHUNK) : 1u;
        och_n[v] = k;
        hub_n[v] = k > 1 ? 1u : 0u;
        slot_n[v] = k > 1 ? k : 0u;
        kmax = max(kmax, k);
        dang += lv && odeg == 0 ? 1 : 0;
    }
    dang = zg_wave_sum_u64(dang);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, off, 64));
    if ((threadIdx.x & 63) == 0) {
        if (dang) atomicAdd(ndang, dang);
        if (kmax > 1) atomicMax(maxk, kmax);
    }
}

__global__ void dv_chunk_kernel(uint32_t V, const uint32_t *in_ptr, const uint32_t *out_ptr, const uint8_t *live,
                                const uint32_t *ich_first, const uint32_t *och_first, const uint32_t *slot_first,
                                uint32_t *ich_v, uint32_t *ich_e0, uint32_t *ich_e1, uint32_t *och_v, uint32_t *och_e0,
                                uint32_t *och_e1, uint32_t *ohub_idx)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        uint32_t b = in_ptr[v], e = in_ptr[v + 1];
        if (live[v] && b < e) {
            uint32_t p = ich_first[v];
            for (uint32_t c = b; c < e; c += ZG_CHUNK, p++) {
                ich_v[p] = v;
                ich_e0[p] = c;
                ich_e1[p] = min(c + ZG_CHUNK, e);
            }
        }
        b = out_ptr[v];
        e = out_ptr[v + 1];
        uint32_t p = och_first[v];
        if (b == e) {
            och_v[p] = v;
            och_e0[p] = b;
            och_e1[p] = e;
        }
        for (uint32_t c = b; c < e; c += ZG_CHUNK, p++) {
            och_v[p] = v;
            och_e0[p] = c;
            och_e1[p] = min(c + ZG_CHUNK, e);
        }
        ohub_idx[v] = och_first[v + 1] - och_first[v] > 1 ? slot_first[v] : ZG_NONE;
    }
}

// Values of the prefix tables at the graphs' first vertices (and at V).
__global__ void dv_bound_kernel(uint32_t nb, const uint32_t *co, const uint32_t *in_ptr, const uint32_t *ich_first,
                                const uint32_t *och_first, const uint32_t *hub_first, uint32_t *out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
        const uint32_t v = co[i];
        out[i] = in_ptr[v];
        out[nb + i] = ich_first[v];
        out[2 * nb + i] = och_first[v];
        out[3 * nb + i] = hub_first[v];
    }
}

// act[c] = 1 + the last vertex of component c (offsets co[0..nb)) with an
// arc in or out; act must start at the components' first vertices.
__global__ void dv_act_kernel(uint32_t V, uint32_t nb, const uint32_t *co, const uint32_t *in_ptr,
                              const uint32_t *out_ptr, uint32_t *act)
{
    // a wavefront's lanes hold consecutive vertices, so they span one or two
    // graphs: one atomic per graph and wavefront, from its highest lane
    const int lane = threadIdx.x & 63;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); v0 < V; v0 += stride) {
        const uint32_t v = v0 + lane;
        const bool has = v < V && (in_ptr[v + 1] != in_ptr[v] || out_ptr[v + 1] != out_ptr[v]);
        uint32_t c = 0;
        if (has) {
            uint32_t lo = 0, hi = nb - 1;  // co[lo] <= v < co[hi]
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (co[mid] <= v) lo = mid;
                else hi = mid;
            }
            c = lo;
        }
        for (uint64_t todo = __ballot(has); todo;) {
            const int j = 63 - __builtin_clzll(todo);
            const uint32_t cj = zg_readlane32(c, j);
            if (lane == 0) atomicMax(&act[cj], v0 + (uint32_t)j + 1);
            todo &= ~__ballot(has && c == cj);
        }
    }
}

// PageRank order (vertices by PageRank id), how many ids follow their
// predecessor vertex's, and the arcs out of the first K ids.
__global__ void dv_order_kernel(uint32_t V, const uint32_t *pr_id, uint32_t *pr_order, unsigned long long *follow)
{
    unsigned long long f = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const uint32_t id = pr_id[v];
        pr_order[id] = v;
        f += v > 0 && id == pr_id[v - 1] + 1 ? 1 : 0;
    }
    f = zg_wave_sum_u64(f);
    if ((threadIdx.x & 63) == 0 && f) atomicAdd(follow, f);
}

__global__ void dv_hot_kernel(uint32_t K, const uint32_t *pr_order, const uint32_t *out_ptr, unsigned long long *hot)
{
    unsigned long long h = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < K; i += gridDim.x * blockDim.x) {
        const uint32_t v = pr_order[i];
        h += out_ptr[v + 1] - out_ptr[v];
    }
    h = zg_wave_sum_u64(h);
    if ((threadIdx.x & 63) == 0 && h) atomicAdd(hot, h);
}

// Bin of the row at position i of the row order: 0 small, 1 m4, 2 m16, 3
// m64, 4 large, 5 not live.
__global__ void dv_bin_kernel(uint32_t V, int by_pr, const uint32_t *pr_order, const uint32_t *in_ptr,
                              const uint8_t *live, uint32_t *key, uint32_t *val, uint32_t *count)
{
    __shared__ uint32_t h[6];
    if (threadIdx.x < 6) h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        const uint32_t v = by_pr ? pr_order[i] : i;
        const uint32_t d = in_ptr[v + 1] - in_ptr[v];
        const uint32_t b = !live[v]          ? 5u
                           : d <= ZG_PR_SMALL ? 0u
                           : d <= ZG_PR_M4    ? 1u
                           : d <= ZG_PR_M16   ? 2u
                           : d <= ZG_PR_MED   ? 3u
                                              : 4u;
        key[i] = b;
        val[i] = v;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 6 && h[threadIdx.x]) atomicAdd(&count[threadIdx.x], h[threadIdx.x]);
}

// Row lengths in bin order (entry n = 0).
__global__ void dv_len_kernel(uint32_t n, const uint32_t *rows, const uint32_t *in_ptr, uint32_t *len)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j <= n; j += gridDim.x * blockDim.x)
        len[j] = j < n ? in_ptr[rows[j] + 1] - in_ptr[rows[j]] : 0u;
}

// {PageRank id, first binned arc, in-degree, out-degree} of the short and
// medium rows.
__global__ void dv_desc_kernel(uint32_t n, const uint32_t *rows, const uint32_t *pos, const uint32_t *pr_id,
                               const uint32_t *in_ptr, const uint32_t *out_ptr, uint4 *desc)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t v = rows[j];
        desc[j] = make_uint4(pr_id[v], pos[j], in_ptr[v + 1] - in_ptr[v], out_ptr[v + 1] - out_ptr[v]);
    }
}

// Long rows: PageRank id, out-degree and segment count (entry n = 0).
__global__ void dv_large_kernel(uint32_t n, const uint32_t *rows, const uint32_t *pr_id, const uint32_t *in_ptr,
                                const uint32_t *out_ptr, uint32_t *large, uint32_t *large_deg, uint32_t *nseg)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j <= n; j += gridDim.x * blockDim.x) {
        if (j == n) {
            nseg[j] = 0;
            continue;
        }
        const uint32_t v = rows[j];
        large[j] = pr_id[v];
        large_deg[j] = out_ptr[v + 1] - out_ptr[v];
        nseg[j] = dv_cdiv(in_ptr[v + 1] - in_ptr[v], ZG_PR_SEG);
    }
}

__global__ void dv_seg_kernel(uint32_t n, const uint32_t *seg0, const uint32_t *pos, const uint32_t *len,
                              uint32_t *seg_row, uint32_t *seg_e0, uint32_t *seg_e1)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t a0 = pos[j], a1 = pos[j] + len[j];
        uint32_t k = seg0[j];
        for (uint32_t c = a0; c < a1; c += ZG_PR_SEG, k++) {
            seg_row[k] = j;
            seg_e0[k] = c;
            seg_e1[k] = min(c + ZG_PR_SEG, a1);
        }
    }
}

// One wavefront per row: copies row j's columns to its binned position.
__global__ void dv_rebin_kernel(uint32_t n, const uint32_t *rows, const uint32_t *in_ptr, const uint32_t *pos,
                                const uint32_t *col, uint32_t *out)
{
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n; j += (gridDim.x * blockDim.x) >> 6) {
        const uint32_t v = rows[j], a = in_ptr[v], k = in_ptr[v + 1] - a, b = pos[j];
        for (uint32_t i = lane; i < k; i += 64) out[b + i] = col[a + i];
    }
}

inline dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192))); }

}  // namespace

int zg_derive_tables(zg_graph *g)
{
    const uint32_t V = g->V;
    const uint64_t E = g->E;
    hipStream_t s = g->stream;
    const dim3 blk(256);
    ZgScratch sc;
    const size_t V1 = (size_t)V + 1;

    // chunk tables
    uint32_t *ich_n, *och_n, *hub_n, *slot_n, *hub_first, *slot_first, *maxk;
    unsigned long long *acc;  // [0] dangling rows, [1] following ids, [2] hot arcs
    ZG_TRY_RC(sc.get(&ich_n, V1));
    ZG_TRY_RC(sc.get(&och_n, V1));
    ZG_TRY_RC(sc.get(&hub_n, V1));
    ZG_TRY_RC(sc.get(&slot_n, V1));
    ZG_TRY_RC(sc.get(&hub_first, V1));
    ZG_TRY_RC(sc.get(&slot_first, V1));
    ZG_TRY_RC(sc.get(&maxk, 1));
    ZG_TRY_RC(sc.get(&acc, 3));
    ZG_TRY(hipMemsetAsync(maxk, 0, sizeof(uint32_t), s));
    ZG_TRY(hipMemsetAsync(acc, 0, 3 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(dv_count_kernel, grid_for(V1), blk, 0, s, V, g->in_ptr, g->out_ptr, g->live, ich_n, och_n, hub_n,
                       slot_n, maxk, acc);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(zg_dmalloc(&g->ich_first, V1));
    ZG_TRY_RC(zg_dmalloc(&g->och_first, V1));
    ZG_TRY_RC(zg_exclusive_scan(ich_n, g->ich_first, V1, s, sc));
    ZG_TRY_RC(zg_exclusive_scan(och_n, g->och_first, V1, s, sc));
    ZG_TRY_RC(zg_exclusive_scan(hub_n, hub_first, V1, s, sc));
    ZG_TRY_RC(zg_exclusive_scan(slot_n, slot_first, V1, s, sc));
    uint32_t tot[4] = {0, 0, 0, 0};  // predecessor chunks, successor chunks, hub slots, largest chunk count
    ZG_TRY(hipMemcpyAsync(&tot[0], g->ich_first + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipMemcpyAsync(&tot[1], g->och_first + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipMemcpyAsync(&tot[2], slot_first + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipMemcpyAsync(&tot[3], maxk, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    g->n_ich = tot[0];
    g->n_och = tot[1];
    g->n_ohub = tot[2];
    g->max_och_per_v = std::max<uint32_t>(1, tot[3]);
    ZG_TRY_RC(zg_dmalloc(&g->ich_v, g->n_ich));
    ZG_TRY_RC(zg_dmalloc(&g->ich_e0, g->n_ich));
    ZG_TRY_RC(zg_dmalloc(&g->ich_e1, g->n_ich));
    ZG_TRY_RC(zg_dmalloc(&g->och_v, g->n_och));
    ZG_TRY_RC(zg_dmalloc(&g->och_e0, g->n_och));
    ZG_TRY_RC(zg_dmalloc(&g->och_e1, g->n_och));
    ZG_TRY_RC(zg_dmalloc(&g->ohub_idx, V));
    if (V)
        hipLaunchKernelGGL(dv_chunk_kernel, grid_for(V), blk, 0, s, V, g->in_ptr, g->out_ptr, g->live, g->ich_first,
                           g->och_first, slot_first, g->ich_v, g->ich_e0, g->ich_e1, g->och_v, g->och_e0, g->och_e1,
                           g->ohub_idx);
    ZG_TRY(hipGetLastError());

    // the tables at the graphs' first vertices
    {
        const std::vector<uint32_t> &co = g->h_comp_off;
        const uint32_t nb = (uint32_t)co.size();
        uint32_t *d_co, *d_b;
        ZG_TRY_RC(sc.get(&d_co, nb));
        ZG_TRY_RC(sc.get(&d_b, 4 * (size_t)nb));
        ZG_TRY(hipMemcpyAsync(d_co, co.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(dv_bound_kernel, grid_for(nb), blk, 0, s, nb, d_co, g->in_ptr, g->ich_first, g->och_first,
                           hub_first, d_b);
        ZG_TRY(hipGetLastError());
        std::vector<uint32_t> b(4 * (size_t)nb);
        ZG_TRY(hipMemcpyAsync(b.data(), d_b, sizeof(uint32_t) * b.size(), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        g->h_cb_in.assign(b.begin(), b.begin() + nb);
        g->h_cb_ich.assign(b.begin() + nb, b.begin() + 2 * (size_t)nb);
        g->h_cb_och.assign(b.begin() + 2 * (size_t)nb, b.begin() + 3 * (size_t)nb);
        g->h_cb_hub.assign(b.begin() + 3 * (size_t)nb, b.end());
        // each graph's end past its last vertex with an arc (vertices after it
        // have none: no search reaches them unless they are its sources)
        uint32_t *d_act;
        ZG_TRY_RC(sc.get(&d_act, nb));
        ZG_TRY(hipMemcpyAsync(d_act, co.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, s));
        if (V && nb > 1) hipLaunchKernelGGL(dv_act_kernel, grid_for(V), blk, 0, s, V, nb, d_co, g->in_ptr, g->out_ptr, d_act);
        ZG_TRY(hipGetLastError());
        g->h_comp_act.resize(nb);
        ZG_TRY(hipMemcpyAsync(g->h_comp_act.data(), d_act, sizeof(uint32_t) * nb, hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
    }

    // PageRank row bins
    uint32_t *pr_order, *key, *key2, *val, *rows, *cnt;
    ZG_TRY_RC(sc.get(&pr_order, V));
    ZG_TRY_RC(sc.get(&key, V));
    ZG_TRY_RC(sc.get(&key2, V));
    ZG_TRY_RC(sc.get(&val, V));
    ZG_TRY_RC(sc.get(&rows, V));
    ZG_TRY_RC(sc.get(&cnt, 6));
    ZG_TRY(hipMemsetAsync(cnt, 0, 6 * sizeof(uint32_t), s));
    const uint32_t K = std::min<uint32_t>(V, ZG_PR_XCACHE_BYTES / sizeof(float));
    if (V) {
        hipLaunchKernelGGL(dv_order_kernel, grid_for(V), blk, 0, s, V, g->pr_id, pr_order, acc + 1);
        hipLaunchKernelGGL(dv_hot_kernel, grid_for(K), blk, 0, s, K, pr_order, g->out_ptr, acc + 2);
    }
    ZG_TRY(hipGetLastError());
    unsigned long long a3[3] = {0, 0, 0};
    ZG_TRY(hipMemcpyAsync(a3, acc, sizeof(a3), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    g->n_dangling = (uint32_t)a3[0];
    // stage the hottest PageRank values in LDS only if they serve >= 10% of
    // the gathers (PageRank ids order vertices by decreasing out-degree)
    g->pr_xcache_useful = E > 0 && a3[2] * 10 >= E;
    // Each bin lists its rows in PageRank-id order, so the rows a wavefront
    // finishes read and write neighbouring rank words -- unless the PageRank
    // ids already follow the vertex ids for most vertices (a grid: nearly
    // all degrees equal), where vertex order keeps a wavefront's predecessors
    // in a few lines and the rank words are neighbours either way (R-MAT 20:
    // 2.19 TB/s in PageRank-id order against 1.98 in vertex order; the 2048 x
    // 2048 grid: 4.5 TB/s in vertex order against 3.7).
    const int by_pr = a3[1] * 2 < (unsigned long long)V;
    if (V)
        hipLaunchKernelGGL(dv_bin_kernel, grid_for(V), blk, 0, s, V, by_pr, pr_order, g->in_ptr, g->live, key, val, cnt);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(zg_sort_pairs(key, key2, val, rows, V, 3, s, sc));
    uint32_t bc6[6];
    ZG_TRY(hipMemcpyAsync(bc6, cnt, sizeof(bc6), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    g->n_small = bc6[0];
    g->n_m4 = bc6[1];
    g->n_m16 = bc6[2];
    g->n_m64 = bc6[3];
    g->n_med = bc6[1] + bc6[2] + bc6[3];
    g->n_large = bc6[4];
    const uint32_t nsm = g->n_small + g->n_med, nrow = nsm + g->n_large;
    // binned position of each row's first arc
    uint32_t *len, *pos;
    ZG_TRY_RC(sc.get(&len, (size_t)nrow + 1));
    ZG_TRY_RC(sc.get(&pos, (size_t)nrow + 1));
    hipLaunchKernelGGL(dv_len_kernel, grid_for((uint64_t)nrow + 1), blk, 0, s, nrow, rows, g->in_ptr, len);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(zg_exclusive_scan(len, pos, (size_t)nrow + 1, s, sc));
    ZG_TRY_RC(zg_dmalloc(&g->pr_small, g->n_small));
    ZG_TRY_RC(zg_dmalloc(&g->pr_med, g->n_med));
    ZG_TRY_RC(zg_dmalloc(&g->pr_desc, nsm));
    ZG_TRY_RC(zg_dmalloc(&g->pr_large, g->n_large));
    ZG_TRY_RC(zg_dmalloc(&g->pr_large_deg, g->n_large));
    ZG_TRY_RC(zg_dmalloc(&g->pr_large_seg0, (size_t)g->n_large + 1));
    if (g->n_small)
        ZG_TRY(hipMemcpyAsync(g->pr_small, rows, sizeof(uint32_t) * g->n_small, hipMemcpyDeviceToDevice, s));
    if (g->n_med)
        ZG_TRY(hipMemcpyAsync(g->pr_med, rows + g->n_small, sizeof(uint32_t) * g->n_med, hipMemcpyDeviceToDevice, s));
    if (nsm)
        hipLaunchKernelGGL(dv_desc_kernel, grid_for(nsm), blk, 0, s, nsm, rows, pos, g->pr_id, g->in_ptr, g->out_ptr,
                           g->pr_desc);
    uint32_t *nseg;
    ZG_TRY_RC(sc.get(&nseg, (size_t)g->n_large + 1));
    hipLaunchKernelGGL(dv_large_kernel, grid_for((uint64_t)g->n_large + 1), blk, 0, s, g->n_large, rows + nsm, g->pr_id,
                       g->in_ptr, g->out_ptr, g->pr_large, g->pr_large_deg, nseg);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(zg_exclusive_scan(nseg, g->pr_large_seg0, (size_t)g->n_large + 1, s, sc));
    uint32_t ends[2] = {0, 0};  // arcs in binned order, segments
    ZG_TRY(hipMemcpyAsync(&ends[0], pos + nrow, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipMemcpyAsync(&ends[1], g->pr_large_seg0 + g->n_large, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    if ((uint64_t)ends[0] != E) return ZG_ERR_INVALID_ARG;  // every arc has a live row
    g->n_seg = ends[1];
    ZG_TRY_RC(zg_dmalloc(&g->pr_seg_row, g->n_seg));
    ZG_TRY_RC(zg_dmalloc(&g->pr_seg_e0, g->n_seg));
    ZG_TRY_RC(zg_dmalloc(&g->pr_seg_e1, g->n_seg));
    if (g->n_large)
        hipLaunchKernelGGL(dv_seg_kernel, grid_for(g->n_large), blk, 0, s, g->n_large, g->pr_large_seg0, pos + nsm,
                           len + nsm, g->pr_seg_row, g->pr_seg_e0, g->pr_seg_e1);
    ZG_TRY(hipGetLastError());
    {
        // the long rows' extents for the iteration's long-row plan (n_large
        // entries, a few thousand at most on a power-law graph)
        std::vector<uint32_t> h_e0(g->n_large), h_len(g->n_large);
        if (g->n_large) {
            ZG_TRY(hipMemcpyAsync(h_e0.data(), pos + nsm, sizeof(uint32_t) * g->n_large, hipMemcpyDeviceToHost, s));
            ZG_TRY(hipMemcpyAsync(h_len.data(), len + nsm, sizeof(uint32_t) * g->n_large, hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
        }
        ZG_TRY_RC(zg_pr_long_plan(g, h_e0.data(), h_len.data(), s));
    }
    // PageRank columns in row-bin order (short rows, then the medium bins,
    // then the long rows): the rows a wavefront sums take neighbouring column
    // words instead of words scattered over the predecessor CSR
    if (E) {
        uint32_t *binned = nullptr;
        ZG_TRY_RC(zg_dmalloc(&binned, E));
        hipLaunchKernelGGL(dv_rebin_kernel, dim3(std::min<uint32_t>((nrow + 3) / 4, 16384) + 1), blk, 0, s, nrow, rows,
                           g->in_ptr, pos, g->pr_col, binned);
        int rc = zg_hip_code(hipGetLastError());
        if (rc == ZG_OK) rc = zg_hip_code(hipStreamSynchronize(s));
        if (rc != ZG_OK) {
            zg_dfree(binned);
            return rc;
        }
        zg_dfree(g->pr_col);
        g->pr_col = binned;
    }
    ZG_TRY(hipStreamSynchronize(s));
    return ZG_OK;
}
