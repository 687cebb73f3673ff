// PageRank power iteration as a pull SpMV over the resident predecessor CSR.
//
// Reference: priv/zog/src/centrality.zig:871-965.  Per iteration
//   sink   = sum over live v with out_deg(v) = 0 of r[v] / n          (:926-931)
//   r'[v]  = (1 - d)/n + d * (sink + sum_{u in in(v), deg u > 0} r[u]/deg(u))
//   l1     = sum |r'[v] - r[v]|                                        (:947-950)
//   swap, then stop when l1 < tolerance (strict)                       (:952-957)
//
// gfx950 layout: ranks r and the gathered values x[u] = r[u]/deg(u) are FP32
// (or FP64 in the FP64 mode), double-buffered by iteration parity, indexed by
// PageRank id (vertices by decreasing out-degree: the values gathered most
// often share cache lines; rows keep the reference's arc order, so the sums
// are unchanged).  Each row
// sum is a compensated (Kahan) sum; partial sums of one row computed by
// different lanes are merged with an error-free TwoSum.  The l1 change and the
// dangling mass are reduced in FP64: every block writes its partials, and a
// one-block finish kernel adds them in a fixed order (deterministic result),
// finishes rows longer than one segment and prepares the next iteration's
// constant term base = (1-d)/n + d*sink.  The iteration kernel runs a fixed
// grid that loops over its work units (no global atomics, no fences);
// launches after convergence return at once, so several iterations are
// queued per host sync.
#include "common.h"
#include "pr_bins.h"

#include <algorithm>
#include <mutex>
#include <queue>
#include <vector>

struct PrState {
    unsigned int done;
    int outcome;      // eigenvector: how the iteration ended (ZG_EIG_*)
    unsigned long long iter;
    unsigned long long max_iter;
    double base;      // PageRank: (1-d)/n + d*sink for the next iteration; Katz: beta
    double last_l1;   // eigenvector: the L2 change
    double damping;   // PageRank: d; Katz / alpha: alpha
    double inv_n;
    double tol;
    double sumsq;     // eigenvector: sum of squares of the raw iterate
};

// What one iteration computes from the gathered sum of a row:
//   ZG_PW_PR      PageRank: r' = base + d * sum, gathered x' = r'/deg
//   ZG_PW_AFFINE  Katz / alpha: r' = alpha * sum + beta, gathered x' = r'
//   ZG_PW_RAW     eigenvector: y = sum (normalised by eig_norm_kernel)
#define ZG_PW_PR 0
#define ZG_PW_AFFINE 1
#define ZG_PW_RAW 2

template <typename T>
struct PrArgs {
    const uint32_t *in_ptr, *in_col, *out_ptr;   // in_col: predecessors as PageRank ids
    const uint32_t *small, *med, *large, *large_deg;
    const uint4 *desc;          // [small | m4 | m16 | m64]: {v, first arc, in-degree, out-degree}
    const uint4 *long_plan;     // long rows by block slot: {row, first arc, end arc, 0}
    const uint32_t *long_off;   // [n_long + 1]
    uint32_t n_small, n_m4, n_m16, n_m64, n_large, n_long;
    uint32_t nb_small, nb_m4, nb_m16, nb_m64;  // work units of each bin
    uint32_t xcache;            // values staged in LDS (0 = none)
    unsigned long long it0;     // iteration index of this run's first launch
    T *r0, *r1, *x0, *x1;       // parity 0 / 1 buffers
    T *rb0, *rb1, *rb2;         // ZG_PW_RAW: the rotating (older, current, new) iterates
    double *part;               // [2 parities][gridDim.x][2] block partials
    PrState *st;
};

template <typename T>
__device__ __forceinline__ void kahan_add(T &s, T &c, T x)
{
    T y = x - c;
    T t = s + y;
    c = (t - s) - y;
    s = t;
}

// Merge two Kahan pairs (value = s - c) with an exact TwoSum of the heads.
template <typename T>
__device__ __forceinline__ void kahan_merge(T &s, T &c, T s2, T c2)
{
    T S = s + s2;
    T bb = S - s;
    T err = (s - (S - bb)) + (s2 - bb);
    s = S;
    c = (c + c2) - err;
}

template <typename T>
__device__ __forceinline__ void wave_kahan_merge(T &s, T &c)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T s2 = __shfl_xor(s, off, 64);
        T c2 = __shfl_xor(c, off, 64);
        kahan_merge(s, c, s2, c2);
    }
}

// Gathered values: ids below K (the most-gathered vertices, PageRank ids
// being ordered by decreasing out-degree) come from the block's LDS copy.
template <typename T>
struct PrGather {
    const T *__restrict__ x;
    const T *lds;
    uint32_t K;
    __device__ __forceinline__ T operator[](uint32_t u) const { return u < K ? lds[u] : x[u]; }
};

#ifndef ZG_PR_INFLIGHT
#define ZG_PR_INFLIGHT 4  // gathers in flight per lane on strided rows
#endif

// Kahan sum of x[col[e]] for e = e0, e0+step, ... < e1, ZG_PR_INFLIGHT
// gathers in flight.
template <typename T, typename X>
__device__ __forceinline__ void pr_strided_kahan(const uint32_t *__restrict__ col, const X &x,
                                                 uint32_t e0, uint32_t e1, uint32_t step, T &s, T &c)
{
    constexpr int N = ZG_PR_INFLIGHT;
    uint32_t e = e0;
    for (; e + (N - 1) * step < e1; e += N * step) {
        uint32_t ci[N];
        T v[N];
#pragma unroll
        for (int k = 0; k < N; k++) ci[k] = col[e + k * step];
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = x[ci[k]];
#pragma unroll
        for (int k = 0; k < N; k++) kahan_add(s, c, v[k]);
    }
    if (e < e1) {
        // the last (at most N - 1) terms: loads at clamped indexes, all in
        // flight together, and each sum step kept only for a real term, so
        // the terms add in the same order as one at a time
        uint32_t ci[N - 1];
        T v[N - 1];
#pragma unroll
        for (int k = 0; k < N - 1; k++) ci[k] = col[min(e + k * step, e1 - 1)];
#pragma unroll
        for (int k = 0; k < N - 1; k++) v[k] = x[ci[k]];
#pragma unroll
        for (int k = 0; k < N - 1; k++) {
            T s2 = s, c2 = c;
            kahan_add(s2, c2, v[k]);
            const bool real = e + k * step < e1;
            s = real ? s2 : s;
            c = real ? c2 : c;
        }
    }
}


// Same, with the old rank and the out-degree already loaded.  l1 gathers
// the L1 change (PageRank, Katz / alpha) or the sum of squares (raw sums).
template <typename T, int MODE>
__device__ __forceinline__ void pr_row_finish(T *rnew, T *xnew, uint32_t v, T r_old, uint32_t outdeg, T s, T c,
                                              double base, double damping, double &l1, double &dang)
{
    const double sum = (double)s - (double)c;
    if constexpr (MODE == ZG_PW_PR) {
        const T rt = (T)(base + damping * sum);
        rnew[v] = rt;
        l1 += fabs((double)rt - (double)r_old);
        if (outdeg == 0) dang += (double)rt;
        else xnew[v] = rt / (T)outdeg;
    } else if constexpr (MODE == ZG_PW_AFFINE) {
        const T rt = (T)__dadd_rn(__dmul_rn(damping, sum), base);  // alpha * sum + beta, two roundings
        rnew[v] = rt;
        l1 += fabs((double)rt - (double)r_old);
    } else {
        const T rt = (T)sum;
        rnew[v] = rt;
        l1 += (double)rt * (double)rt;
    }
}

// The buffers of iteration `it`: gathered values, old and new iterates.
template <typename T, int MODE>
struct PrBufs {
    const T *xg, *rold;
    T *rnew, *xnew;
    __device__ __forceinline__ PrBufs(const PrArgs<T> &a, unsigned long long it)
    {
        if constexpr (MODE == ZG_PW_RAW) {
            const unsigned k = (unsigned)(it % 3);
            xg = k == 0 ? a.rb0 : (k == 1 ? a.rb1 : a.rb2);
            rold = xg;
            rnew = k == 0 ? a.rb1 : (k == 1 ? a.rb2 : a.rb0);
            xnew = rnew;
        } else {
            const bool odd = it & 1;
            xg = odd ? a.x1 : a.x0;
            rold = odd ? a.r1 : a.r0;
            rnew = odd ? a.r0 : a.r1;
            xnew = odd ? a.x0 : a.x1;
        }
    }
};

// Rows of one bin summed by groups of L lanes: unit u covers rows
// [u*(256/L), (u+1)*(256/L)) of the bin; lane groups merge their Kahan pairs.
template <typename T, int MODE, int L, typename X>
__device__ __forceinline__ void pr_rows_lanes(const PrArgs<T> &a, const uint4 *rows, uint32_t nrows,
                                              uint32_t unit, const X &x, const T *rold, T *rnew, T *xnew,
                                              double base, double damping, double &l1, double &dang, int tid)
{
    const uint32_t i = unit * (ZG_PR_BLOCK / L) + tid / L;
    const uint32_t sub = tid % L;
    const bool ok = i < nrows;
    uint4 dsc = make_uint4(0, 0, 0, 0);
    T s = 0, c = 0, r_old = 0;
    if (ok) {
        dsc = rows[i];
        if (MODE != ZG_PW_RAW && sub == 0) r_old = rold[dsc.x];
        pr_strided_kahan(a.in_col, x, dsc.y + sub, dsc.y + dsc.z, (uint32_t)L, s, c);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
        T s2 = __shfl_xor(s, off, 64);
        T c2 = __shfl_xor(c, off, 64);
        kahan_merge(s, c, s2, c2);
    }
    if (ok && sub == 0) pr_row_finish<T, MODE>(rnew, xnew, dsc.x, r_old, dsc.w, s, c, base, damping, l1, dang);
}

// One work unit of the row bins (b < nb_small + nb_m4 + nb_m16 + nb_m64)
// by 256 threads (tid in [0, 256)).
template <typename T, int MODE, typename X>
__device__ __forceinline__ void pr_unit(const PrArgs<T> &a, uint32_t b, int tid, const X &x, const T *rold, T *rnew,
                                        T *xnew, double base, double damping, double &l1, double &dang)
{
    if (b < a.nb_small) {
        // one lane per row (<= ZG_PR_SMALL predecessors)
        const uint32_t i = b * ZG_PR_BLOCK + tid;
        if (i < a.n_small) {
            const uint4 dsc = a.desc[i];
            const uint32_t e0 = dsc.y, deg = dsc.z;
            // all predecessor ids, then all gathers, in flight together
            uint32_t col[ZG_PR_SMALL];
            T val[ZG_PR_SMALL];
#pragma unroll
            for (int k = 0; k < ZG_PR_SMALL; k++) col[k] = (uint32_t)k < deg ? a.in_col[e0 + k] : 0u;
            const T r_old = MODE != ZG_PW_RAW ? rold[dsc.x] : (T)0;
#pragma unroll
            for (int k = 0; k < ZG_PR_SMALL; k++) val[k] = (uint32_t)k < deg ? x[col[k]] : (T)0;
            T s = 0, c = 0;
#pragma unroll
            for (int k = 0; k < ZG_PR_SMALL; k++)
                if ((uint32_t)k < deg) kahan_add(s, c, val[k]);
            pr_row_finish<T, MODE>(rnew, xnew, dsc.x, r_old, dsc.w, s, c, base, damping, l1, dang);
        }
    } else if (b < a.nb_small + a.nb_m4) {
        pr_rows_lanes<T, MODE, 4>(a, a.desc + a.n_small, a.n_m4, b - a.nb_small, x, rold, rnew, xnew, base, damping,
                                  l1, dang, tid);
    } else if (b < a.nb_small + a.nb_m4 + a.nb_m16) {
        pr_rows_lanes<T, MODE, 16>(a, a.desc + a.n_small + a.n_m4, a.n_m16, b - a.nb_small - a.nb_m4, x, rold, rnew,
                                   xnew, base, damping, l1, dang, tid);
    } else {
        pr_rows_lanes<T, MODE, 64>(a, a.desc + a.n_small + a.n_m4 + a.n_m16, a.n_m64,
                                   b - a.nb_small - a.nb_m4 - a.nb_m16, x, rold, rnew, xnew, base, damping, l1, dang,
                                   tid);
    }
}

// One iteration.
//
// PageRank and Katz / alpha (the fused modes) need no finishing launch: a
// launch for iteration index `it` first adds the previous launch's block
// partials of (L1 change, dangling mass) -- every block the same numbers in
// the same order -- and, when the change is below the tolerance or `it` is
// the iteration cap, block 0 records the end in the state and every block
// returns; otherwise the dangling mass gives this iteration's constant term.
// Partials alternate between two buffers by `it`.  The eigenvector
// iteration (ZG_PW_RAW) reads its index from the state and is finished by
// pr_finish_kernel and the eig_* kernels.
//
// Work: blocks [0, n_long) each sum the long rows (more than ZG_PR_MED
// predecessors) of their plan slot, one row at a time with all NT threads
// (dealt by arc count, so these blocks finish with the others); the other
// blocks are NT/256 independent 256-thread workers (quarters) over the row
// bins' work units.  All of a block's threads share its LDS copy of the
// a.xcache most-gathered values (the lowest PageRank ids), loaded once per
// launch.
// dst[0, K) (LDS, 16-byte aligned) = src[0, K), with ZG_LDS_FILL loads in
// flight per thread before the first LDS write: 16-byte loads when src is
// 16-byte aligned (the rank vectors of a V-sized layout need not be), element
// loads otherwise.  Indexes past the end are clamped to the last entry, for
// the loads and the writes alike (such a write stores the value the last
// entry's own write stores), so no load is issued under a branch.
#ifndef ZG_LDS_FILL
#define ZG_LDS_FILL 8
#endif
template <typename T, int NT>
__device__ __forceinline__ void zg_lds_fill(T *dst, const T *src, uint32_t K)
{
    constexpr int U = ZG_LDS_FILL;
    if (K == 0) return;
    constexpr uint32_t per = 16 / sizeof(T);
    const uint32_t n16 = (reinterpret_cast<uintptr_t>(src) & 15) == 0 ? K / per : 0;
    if (n16) {
        const uint4 *s16 = reinterpret_cast<const uint4 *>(src);
        uint4 *d16 = reinterpret_cast<uint4 *>(dst);
        for (uint32_t i0 = threadIdx.x; i0 < n16; i0 += U * NT) {
            uint4 t[U];
#pragma unroll
            for (int u = 0; u < U; u++) t[u] = s16[min(i0 + (uint32_t)u * NT, n16 - 1)];
#pragma unroll
            for (int u = 0; u < U; u++) d16[min(i0 + (uint32_t)u * NT, n16 - 1)] = t[u];
        }
    }
    for (uint32_t i0 = n16 * per + threadIdx.x; i0 < K; i0 += U * NT) {
        T t[U];
#pragma unroll
        for (int u = 0; u < U; u++) t[u] = src[min(i0 + (uint32_t)u * NT, K - 1)];
#pragma unroll
        for (int u = 0; u < U; u++) dst[min(i0 + (uint32_t)u * NT, K - 1)] = t[u];
    }
}

template <typename T, int MODE, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void pr_iter_kernel(PrArgs<T> a, unsigned long long it_arg)
{
    static_assert(NT % ZG_PR_BLOCK == 0, "a block is a whole number of 256-thread workers");
    constexpr bool FUSED = MODE != ZG_PW_RAW;
    constexpr uint32_t NQ = NT / ZG_PR_BLOCK;
    extern __shared__ uint4 pr_lds[];  // [a.xcache] values of type T (launch-sized)
    __shared__ double red[2][NT / 64];
    __shared__ T kred[2][NT / 64];
    __shared__ double tot[2];
    PrState *st = a.st;
    if (st->done) return;
    const unsigned long long it = FUSED ? it_arg : st->iter;
    // the previous launch's block partials (this thread's first pair: the
    // part buffer always holds both parities, so the load is unconditional),
    // then the LDS copy: both sets of loads in flight together
    const bool add_up = FUSED && it > a.it0;
    const double *pp = a.part + 2 * (size_t)gridDim.x * ((it - 1) & 1);
    const double2 p0 = reinterpret_cast<const double2 *>(pp)[min((uint32_t)threadIdx.x, gridDim.x - 1)];
    const PrBufs<T, MODE> B(a, it);
    const T *xg = B.xg;
    T *xs = reinterpret_cast<T *>(pr_lds);
    const uint32_t K = a.xcache;
    zg_lds_fill<T, NT>(xs, xg, K);
    double L = 0.0, D = 0.0;
    if (add_up) {
        if (threadIdx.x < gridDim.x) {
            L = p0.x;
            D = p0.y;
        }
        for (uint32_t q = threadIdx.x + NT; q < gridDim.x; q += NT) {
            L += pp[2 * q];
            D += pp[2 * q + 1];
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double damping = st->damping;
    double base = st->base;
    if constexpr (FUSED) {
        if (add_up) {
            L = zg_wave_sum(L);
            D = zg_wave_sum(D);
            if (lane == 0) {
                red[0][wv] = L;
                red[1][wv] = D;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double Ls = 0.0, Ds = 0.0;
                for (int w = 0; w < NT / 64; w++) {
                    Ls += red[0][w];
                    Ds += red[1][w];
                }
                tot[0] = Ls;
                tot[1] = Ds;
            }
            __syncthreads();
            L = tot[0];
            D = tot[1];
            // PageRank: (1-d)/n + d * sink; Katz / alpha: beta throughout
            const double nb = MODE == ZG_PW_PR ? (1.0 - damping) * st->inv_n + damping * (D * st->inv_n) : base;
            if (L < st->tol || it >= st->max_iter) {
                if (blockIdx.x == 0 && threadIdx.x == 0) {
                    st->iter = it;
                    st->last_l1 = L;
                    st->base = nb;
                    st->done = 1;
                }
                return;
            }
            base = nb;
        }
    }
    __syncthreads();  // the LDS copy (loads issued before the partials were added)
    const PrGather<T> x{xg, xs, K};
    const T *rold = B.rold;
    T *rnew = B.rnew;
    T *xnew = B.xnew;
    double l1 = 0.0, dang = 0.0;

    // long rows: slots b, b + gridDim.x, ... (one slot per block unless the
    // grid is smaller than the plan)
    for (uint32_t slot = blockIdx.x; slot < a.n_long; slot += gridDim.x) {
        for (uint32_t k = a.long_off[slot]; k < a.long_off[slot + 1]; k++) {
            const uint4 p = a.long_plan[k];
            T s = 0, c = 0;
            pr_strided_kahan(a.in_col, x, p.y + threadIdx.x, p.z, (uint32_t)NT, s, c);
            wave_kahan_merge(s, c);
            if (lane == 0) {
                kred[0][wv] = s;
                kred[1][wv] = c;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                T S = kred[0][0], C = kred[1][0];
                for (int w = 1; w < NT / 64; w++) kahan_merge(S, C, kred[0][w], kred[1][w]);
                const uint32_t v = a.large[p.x];
                pr_row_finish<T, MODE>(rnew, xnew, v, MODE != ZG_PW_RAW ? rold[v] : (T)0, a.large_deg[p.x], S, C,
                                       base, damping, l1, dang);
            }
            __syncthreads();
        }
    }
    // row bins: quarters of the blocks past the long-row slots (of every
    // block when the grid has no block to spare)
    const uint32_t first = a.n_long < gridDim.x ? a.n_long : 0u;
    if (blockIdx.x >= first) {
        const int tid = threadIdx.x & (ZG_PR_BLOCK - 1);
        const uint32_t q = threadIdx.x / ZG_PR_BLOCK;
        const uint32_t units = a.nb_small + a.nb_m4 + a.nb_m16 + a.nb_m64;
        for (uint32_t b = (blockIdx.x - first) * NQ + q; b < units; b += (gridDim.x - first) * NQ)
            pr_unit<T, MODE>(a, b, tid, x, rold, rnew, xnew, base, damping, l1, dang);
    }

    // block partials of (l1, dangling mass); added in a fixed order later
    l1 = zg_wave_sum(l1);
    dang = zg_wave_sum(dang);
    if (lane == 0) {
        red[0][wv] = l1;
        red[1][wv] = dang;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double L = 0.0, D = 0.0;
        for (int w = 0; w < NT / 64; w++) {
            L += red[0][w];
            D += red[1][w];
        }
        double *pp = a.part + 2 * (size_t)gridDim.x * (FUSED ? (it & 1) : 0);
        pp[2 * blockIdx.x] = L;
        pp[2 * blockIdx.x + 1] = D;
    }
}

#define ZG_PR_FINISH_THREADS 1024

// Eigenvector iteration: one block adds the iteration kernel's block
// partials (the sum of squares of the raw iterate) in a fixed order for
// eig_norm_kernel and eig_decide_kernel.
__global__ __launch_bounds__(ZG_PR_FINISH_THREADS) void pr_finish_kernel(PrArgs<double> a, uint32_t nparts)
{
    __shared__ double red[ZG_PR_FINISH_THREADS / 64];
    PrState *st = a.st;
    if (st->done) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double l1 = 0.0;
    for (uint32_t q = tid; q < nparts; q += ZG_PR_FINISH_THREADS) l1 += a.part[2 * q];
    l1 = zg_wave_sum(l1);
    if (lane == 0) red[wv] = l1;
    __syncthreads();
    if (tid == 0) {
        double L = 0.0;
        for (int w = 0; w < ZG_PR_FINISH_THREADS / 64; w++) L += red[w];
        st->sumsq = L;
    }
}

// r <- src (uniform 1/n when src == nullptr) on live rows, 0 elsewhere;
// x = r/deg on live rows with successors, 0 elsewhere; both parities.  The
// buffers are in PageRank ids; src is in PageRank ids too when src_pr, else
// in vertex ids.
template <typename T, typename S>
__global__ void pr_init_kernel(uint32_t V, const uint8_t *live, const uint32_t *out_ptr, const uint32_t *pr_id,
                               const S *src, bool src_pr, double uniform, T *r0, T *r1, T *x0, T *x1)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t p = pr_id[v];
    T r = 0;
    if (live[v]) r = src ? (T)src[src_pr ? p : v] : (T)uniform;
    uint32_t deg = out_ptr[v + 1] - out_ptr[v];
    r0[p] = r;
    r1[p] = 0;
    x0[p] = deg ? r / (T)deg : (T)0;
    x1[p] = 0;
}

template <typename T>
__global__ void pr_out_kernel(uint32_t V, const uint32_t *pr_id, const T *r, double *out)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = (double)r[pr_id[v]];
}

__global__ void pr_single_kernel(uint32_t V, const uint8_t *live, double *out)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = live[v] ? 1.0 : 0.0;
}

int zg_pr_ws_release(zg_graph *g)
{
    zg_pr_ws &w = g->pr;
    zg_dfree(w.r);
    zg_dfree(w.xh);
    zg_dfree(w.part);
    zg_dfree(w.state);
    if (w.h_state) (void)hipHostFree(w.h_state);
    w.h_state = nullptr;
    return ZG_OK;
}

// Blocks of the iteration kernel: one per unit up to a fixed cap (the
// kernel loops over units with a grid stride).
#define ZG_PR_MAX_GRID 2048

// The mixed mode leaves FP32 once the L1 change drops below this (the FP32
// rounding floor of a probability vector is ~1e-7 in L1).
#define ZG_PR_MIXED_SWITCH 1e-6
static uint32_t pr_grid(const zg_graph *g)
{
    uint32_t units = (g->n_small + ZG_PR_BLOCK - 1) / ZG_PR_BLOCK + (g->n_m4 + 63) / 64 + (g->n_m16 + 15) / 16 +
                     (g->n_m64 + 3) / 4 + g->pr_nlong;
    const uint32_t cap = g->pr_grid_cap ? g->pr_grid_cap : ZG_PR_MAX_GRID;
    return units < cap ? (units ? units : 1) : cap;
}

static int pr_ws_ensure(zg_graph *g)
{
    zg_pr_ws &w = g->pr;
    if (w.state) return ZG_OK;
    size_t V = g->V;
    // FP64 buffers (also used for the FP64 residual check) and FP32 buffers
    // share one allocation: [r64 x2][x64 x2][r32 x2][x32 x2]
    ZG_TRY_RC(zg_dmalloc((char **)&w.r, V * (4 * 8 + 4 * 4)));
    w.xh = nullptr;
    ZG_TRY_RC(zg_dmalloc(&w.part, 4 * (size_t)pr_grid(g)));  // two parities of (l1, dangling) per block
    ZG_TRY_RC(zg_dmalloc((char **)&w.state, 2 * sizeof(PrState)));
    ZG_TRY(hipMemset(w.state, 0, 2 * sizeof(PrState)));
    ZG_TRY(hipHostMalloc(&w.h_state, 2 * sizeof(PrState), hipHostMallocDefault));
    return ZG_OK;
}

// Launch shape of the iteration kernel.  Where the most-gathered values
// carry a useful share of the gathers, one 1024-thread block per CU keeps
// as many of them in LDS as one block may hold (~40K FP32 values on gfx950:
// 70% of R-MAT 20's gathers, against 35% for the 4K values of a 16 KB copy
// per 256-thread block); otherwise 256-thread blocks with no copy.
struct PrShape {
    uint32_t grid = 1, threads = ZG_PR_BLOCK, K = 0;
    size_t lds = 0;  // dynamic LDS bytes
};

static void pr_device_limits(zg_graph *g)
{
    if (g->dev_cus) return;
    int cus = 0, lds = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device) != hipSuccess ||
        lds < 65536)
        lds = 65536;
    g->dev_cus = (uint32_t)cus;
    g->dev_lds = (uint32_t)lds;
}

template <typename T>
static PrShape pr_shape(zg_graph *g)
{
    pr_device_limits(g);
    PrShape sh;
    const uint32_t units = pr_grid(g);
    // large blocks where the copy is useful or rows are long (a long row is
    // summed by a whole block)
    const uint32_t threads =
        g->pr_block ? g->pr_block : (g->pr_xcache_useful || g->n_large ? 1024u : (uint32_t)ZG_PR_BLOCK);
    sh.threads = threads;
    const uint32_t want = g->pr_xcache_useful ? std::min<uint32_t>(g->pr_xcache, g->V) : 0u;
    if (threads == ZG_PR_BLOCK) {
        sh.K = std::min<uint32_t>(want, ZG_PR_XCACHE_BYTES / sizeof(T));
        sh.grid = units;
    } else {
        // by default as many values as let the blocks of a full CU (2048
        // threads: 8 wavefronts per SIMD) share its LDS, each keeping 1 KB for
        // its own reductions; a larger cap set by ZG_TUNE_PR_CACHE gives
        // fewer, larger copies
        const uint32_t full = std::max<uint32_t>(1u, 2048u / threads);
        const uint32_t fit = (g->dev_lds / full - 1024u) / (uint32_t)sizeof(T);
        const uint32_t most = (g->dev_lds - 1024u) / (uint32_t)sizeof(T);
        sh.K = std::min<uint32_t>(want, g->pr_xcache == ZG_PR_XCACHE_DEFAULT ? fit : most);
        const uint32_t lds = sh.K * (uint32_t)sizeof(T);
        const uint32_t by_lds = lds ? std::max<uint32_t>(1u, g->dev_lds / lds) : 8u;
        const uint32_t by_threads = std::max<uint32_t>(1u, 2048u / threads);
        const uint32_t per_cu = std::min(by_lds, by_threads);
        const uint32_t quarters = (units + threads / ZG_PR_BLOCK - 1) / (threads / ZG_PR_BLOCK);
        sh.grid = std::max<uint32_t>(1u, std::min<uint32_t>(g->dev_cus * per_cu, quarters));
        if (g->pr_grid_cap) sh.grid = std::min(sh.grid, g->pr_grid_cap);
    }
    sh.lds = (size_t)sh.K * sizeof(T);
    return sh;
}

// Long rows (more than ZG_PR_MED predecessors) to iteration blocks: as many
// slots as the default grid's share of blocks matches the long rows' share
// of the arcs (at least one, at most half the grid), rows dealt largest
// first to the least loaded slot, so a slot's block finishes with the
// blocks that sum the row bins.  Each slot's rows in dealing order.
int zg_pr_long_plan(zg_graph *g, const uint32_t *h_e0, const uint32_t *h_len, hipStream_t s)
{
    zg_dfree(g->pr_long_plan);
    zg_dfree(g->pr_long_off);
    g->pr_nlong = 0;
    const uint32_t nl = g->n_large;
    if (nl == 0) return ZG_OK;
    const PrShape sh = pr_shape<float>(g);
    uint64_t long_arcs = 0;
    for (uint32_t i = 0; i < nl; i++) long_arcs += h_len[i];
    const double share = g->E ? (double)long_arcs / (double)g->E : 1.0;
    uint32_t P = (uint32_t)std::llround((double)sh.grid * share);
    P = std::max<uint32_t>(1u, std::min<uint32_t>(P, sh.grid > 1 ? sh.grid / 2 : 1u));
    P = std::min<uint32_t>(P, nl);
    std::vector<uint32_t> order(nl);
    for (uint32_t i = 0; i < nl; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return h_len[x] > h_len[y]; });
    // least loaded slot first (ties: lowest slot)
    using Slot = std::pair<uint64_t, uint32_t>;
    std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
    for (uint32_t b = 0; b < P; b++) heap.push({0ull, b});
    std::vector<std::vector<uint32_t>> lists(P);
    for (uint32_t i : order) {
        Slot t = heap.top();
        heap.pop();
        lists[t.second].push_back(i);
        t.first += h_len[i];
        heap.push(t);
    }
    std::vector<uint4> plan;
    plan.reserve(nl);
    std::vector<uint32_t> off(P + 1, 0);
    for (uint32_t b = 0; b < P; b++) {
        off[b] = (uint32_t)plan.size();
        for (uint32_t i : lists[b]) plan.push_back(make_uint4(i, h_e0[i], h_e0[i] + h_len[i], 0u));
    }
    off[P] = (uint32_t)plan.size();
    ZG_TRY_RC(zg_dmalloc(&g->pr_long_plan, nl));
    ZG_TRY_RC(zg_dmalloc(&g->pr_long_off, (size_t)P + 1));
    ZG_TRY(hipMemcpyAsync(g->pr_long_plan, plan.data(), sizeof(uint4) * nl, hipMemcpyHostToDevice, s));
    ZG_TRY(hipMemcpyAsync(g->pr_long_off, off.data(), sizeof(uint32_t) * (P + 1), hipMemcpyHostToDevice, s));
    ZG_TRY(hipStreamSynchronize(s));
    g->pr_nlong = P;
    return ZG_OK;
}

template <typename T, int MODE>
static void pr_launch_iter(const PrShape &sh, hipStream_t s, const PrArgs<T> &a, unsigned long long it = 0)
{
    if (sh.threads == 1024)
        hipLaunchKernelGGL((pr_iter_kernel<T, MODE, 1024>), dim3(sh.grid), dim3(1024), sh.lds, s, a, it);
    else if (sh.threads == 512)
        hipLaunchKernelGGL((pr_iter_kernel<T, MODE, 512>), dim3(sh.grid), dim3(512), sh.lds, s, a, it);
    else
        hipLaunchKernelGGL((pr_iter_kernel<T, MODE, ZG_PR_BLOCK>), dim3(sh.grid), dim3(ZG_PR_BLOCK), sh.lds, s, a,
                           it);
}

// The dynamic LDS above the default 64 KB has to be allowed per kernel.
template <typename T, int MODE>
static int pr_allow_lds(const PrShape &sh)
{
    if (sh.lds <= 65536) return ZG_OK;
    const void *f = sh.threads == 1024 ? reinterpret_cast<const void *>(&pr_iter_kernel<T, MODE, 1024>)
                                       : reinterpret_cast<const void *>(&pr_iter_kernel<T, MODE, 512>);
    ZG_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds));
    return ZG_OK;
}

template <typename T>
static PrArgs<T> pr_args(zg_graph *g, T *r0, T *r1, T *x0, T *x1, PrState *st, double *part = nullptr)
{
    PrArgs<T> a;
    a.in_ptr = g->in_ptr;
    a.in_col = g->pr_col;
    a.out_ptr = g->out_ptr;
    a.small = g->pr_small;
    a.med = g->pr_med;
    a.desc = g->pr_desc;
    a.large = g->pr_large;
    a.large_deg = g->pr_large_deg;
    a.long_plan = g->pr_long_plan;
    a.long_off = g->pr_long_off;
    a.n_long = g->pr_nlong;
    a.it0 = 0;
    a.n_small = g->n_small;
    a.n_m4 = g->n_m4;
    a.n_m16 = g->n_m16;
    a.n_m64 = g->n_m64;
    a.n_large = g->n_large;
    a.nb_small = (g->n_small + ZG_PR_BLOCK - 1) / ZG_PR_BLOCK;
    a.nb_m4 = (g->n_m4 + ZG_PR_BLOCK / 4 - 1) / (ZG_PR_BLOCK / 4);
    a.nb_m16 = (g->n_m16 + ZG_PR_BLOCK / 16 - 1) / (ZG_PR_BLOCK / 16);
    a.nb_m64 = (g->n_m64 + ZG_PR_BLOCK / 64 - 1) / (ZG_PR_BLOCK / 64);
    a.xcache = pr_shape<T>(g).K;
    a.r0 = r0;
    a.r1 = r1;
    a.x0 = x0;
    a.x1 = x1;
    a.rb0 = a.rb1 = a.rb2 = nullptr;
    a.part = part ? part : g->pr.part;
    a.st = st;
    return a;
}

// Runs a fused-mode power iteration (PageRank, Katz / alpha) in precision
// T from the initialised buffers and the device state: launches for the
// iteration indices a.it0, a.it0 + 1, ... up to st.max_iter, the last one
// only adding up (it records the end when nothing before did), eight per
// host sync; launches after the end return at once.  Returns with the host
// mirror of the final state filled.
template <typename T, int MODE>
static int pr_run(zg_graph *g, hipStream_t s, const PrArgs<T> &a, PrState *h_st, uint64_t max_iter, bool prof = true)
{
    const PrShape sh = pr_shape<T>(g);
    ZG_TRY_RC((pr_allow_lds<T, MODE>(sh)));
    const uint64_t K = 8;
    unsigned long long it = a.it0;
    bool synced = false;
    while (it <= max_iter) {
        const uint64_t k = std::min<uint64_t>(K, max_iter + 1 - it);
        for (uint64_t i = 0; i < k; i++, it++) {
            const size_t p = prof ? zg_prof_begin(g, s) : SIZE_MAX;
            pr_launch_iter<T, MODE>(sh, s, a, it);
            zg_prof_end(g, s, ZG_PROF_PR_ITER, p);
        }
        ZG_TRY(hipGetLastError());
        ZG_TRY(hipMemcpyAsync(h_st, a.st, sizeof(PrState), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        synced = true;
        if (h_st->done) break;
    }
    if (!synced) {
        ZG_TRY(hipMemcpyAsync(h_st, a.st, sizeof(PrState), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
    }
    return ZG_OK;
}

// Per-kernel profile of a fused run that computed `computed` iterations:
// its launches past the closing one returned at once and are dropped; the
// closing launch (it only adds the last partials up) is timed with the
// iterations, whose count is what the profile reports.
static int pr_collect_run(zg_graph *g, hipStream_t s, uint64_t computed)
{
    if (g->ev_marks.empty()) return ZG_OK;
    if (g->ev_marks.size() > computed + 1) g->ev_marks.resize(computed + 1);
    const bool closing = g->ev_marks.size() == computed + 1;
    ZG_TRY_RC(zg_prof_collect(g, s));
    if (closing && g->prof.launches[ZG_PROF_PR_ITER] > 0) g->prof.launches[ZG_PROF_PR_ITER] -= 1;
    return ZG_OK;
}

// The buffers one PageRank call works in: the handle's workspace, or the
// fixed scratch of zg_pagerank_device_noalloc (then no events are created:
// no kernel time, no per-kernel profile).
struct PrMem {
    double *r64 = nullptr, *x64 = nullptr;  // [2][V] each (FP64 phases, residual check)
    float *r32 = nullptr, *x32 = nullptr;   // [2][V] each (FP32 phases)
    double *part = nullptr;                 // [2 parities][grid][2] block partials
    PrState *d_st = nullptr, *h_st = nullptr;  // [2] device state and its host mirror
    bool timed = true;
};

static int zg_pagerank_impl(zg_graph *g, double damping, uint64_t max_iterations, double tolerance,
                            int precision, double *d_out, hipStream_t s, zg_pr_info *info, const PrMem *mem = nullptr)
{
    const bool check = !(precision & ZG_PR_SKIP_RESIDUAL);
    precision &= 0xFF;
    if (precision != ZG_PR_FP32 && precision != ZG_PR_FP64 && precision != ZG_PR_MIXED) return ZG_ERR_INVALID_ARG;
    const bool mixed = precision == ZG_PR_MIXED;
    if (g->h_comp_off.size() > 2) return ZG_ERR_INVALID_ARG;  // one PageRank per resource
    const uint32_t V = g->V;
    const uint32_t n = g->n_live;
    zg_pr_info inf;
    memset(&inf, 0, sizeof(inf));
    inf.precision = precision;
    const dim3 vb(256), vg((V + 255) / 256);
    if (n <= 1) {
        // centrality.zig:875-879: n = 0 gives no scores (every id is
        // non-live: 0.0), n = 1 gives 1.0
        if (V) {
            hipLaunchKernelGGL(pr_single_kernel, vg, vb, 0, s, V, g->live, d_out);
            ZG_TRY(hipGetLastError());
        }
        ZG_TRY(hipStreamSynchronize(s));
        if (info) *info = inf;
        return ZG_OK;
    }
    PrMem ws;
    if (!mem) {
        ZG_TRY_RC(pr_ws_ensure(g));
        ws.r64 = (double *)g->pr.r;
        ws.x64 = ws.r64 + 2 * (size_t)V;
        ws.r32 = (float *)(ws.x64 + 2 * (size_t)V);
        ws.x32 = ws.r32 + 2 * (size_t)V;
        ws.part = g->pr.part;
        ws.d_st = (PrState *)g->pr.state;
        ws.h_st = (PrState *)g->pr.h_state;
        mem = &ws;
    }
    double *r64 = mem->r64, *x64 = mem->x64, *part = mem->part;
    float *r32 = mem->r32, *x32 = mem->x32;
    PrState *d_st = mem->d_st;
    PrState *h_st = mem->h_st;
    const bool timed = mem->timed;

    const double inv_n = 1.0 / (double)n;
    PrState st0;
    memset(&st0, 0, sizeof(st0));
    st0.max_iter = max_iterations;
    st0.damping = damping;
    st0.inv_n = inv_n;
    st0.tol = tolerance;
    // sink of the uniform start vector: n_dangling * (1/n) / n
    st0.base = (1.0 - damping) * inv_n + damping * ((double)g->n_dangling * inv_n * inv_n);
    st0.done = max_iterations == 0 ? 1u : 0u;

    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) {
        ZG_TRY(hipEventCreate(&e0));
        ZG_TRY(hipEventCreate(&e1));
    }
    int rc = ZG_OK;
    do {
        // phase 1: FP64 for ZG_PR_FP64, FP32 otherwise (the mixed mode stops
        // it at the FP32 noise floor and finishes in FP64)
        const bool fp64 = precision == ZG_PR_FP64;
        PrState st = st0;
        if (mixed) st.tol = std::max(tolerance, ZG_PR_MIXED_SWITCH);
        ZG_TRY(hipMemcpyAsync(d_st, &st, sizeof(PrState), hipMemcpyHostToDevice, s));
        if (!fp64) {
            hipLaunchKernelGGL((pr_init_kernel<float, float>), vg, vb, 0, s, V, g->live, g->out_ptr, g->pr_id,
                               (const float *)nullptr, false, inv_n, r32, r32 + V, x32, x32 + V);
        } else {
            hipLaunchKernelGGL((pr_init_kernel<double, double>), vg, vb, 0, s, V, g->live, g->out_ptr, g->pr_id,
                               (const double *)nullptr, false, inv_n, r64, r64 + V, x64, x64 + V);
        }
        ZG_TRY(hipGetLastError());
        if (timed) ZG_TRY(hipEventRecord(e0, s));
        if (!fp64) {
            PrArgs<float> a = pr_args<float>(g, r32, r32 + V, x32, x32 + V, d_st, part);
            rc = pr_run<float, ZG_PW_PR>(g, s, a, h_st, max_iterations, timed);
        } else {
            PrArgs<double> a = pr_args<double>(g, r64, r64 + V, x64, x64 + V, d_st, part);
            rc = pr_run<double, ZG_PW_PR>(g, s, a, h_st, max_iterations, timed);
        }
        if (rc != ZG_OK) break;
        uint64_t iters = h_st->iter;
        if (timed) ZG_TRY_RC(pr_collect_run(g, s, iters));
        bool final64 = fp64;
        if (mixed && !(h_st->last_l1 < tolerance) && iters < max_iterations) {
            // phase 2: FP64 from the FP32 iterate; the buffers of parity
            // (iters & 1) hold the start vector, the state carries on
            final64 = true;
            const bool odd = iters & 1;
            const float *rs = odd ? r32 + V : r32;
            double *ra = odd ? r64 + V : r64, *rb = odd ? r64 : r64 + V;
            double *xa = odd ? x64 + V : x64, *xb = odd ? x64 : x64 + V;
            hipLaunchKernelGGL((pr_init_kernel<double, float>), vg, vb, 0, s, V, g->live, g->out_ptr, g->pr_id, rs,
                               true, 0.0, ra, rb, xa, xb);
            ZG_TRY(hipGetLastError());
            PrState st2 = *h_st;
            st2.tol = tolerance;
            st2.done = 0;
            ZG_TRY(hipMemcpyAsync(d_st, &st2, sizeof(PrState), hipMemcpyHostToDevice, s));
            PrArgs<double> a = pr_args<double>(g, r64, r64 + V, x64, x64 + V, d_st, part);
            a.it0 = iters;  // the FP32 phase's block partials are not read: st2.base carries its dangling mass
            rc = pr_run<double, ZG_PW_PR>(g, s, a, h_st, max_iterations, timed);
            if (rc != ZG_OK) break;
            const uint64_t more = h_st->iter - iters;
            if (timed) ZG_TRY_RC(pr_collect_run(g, s, more));
            inf.fp64_iterations = more;
            iters = h_st->iter;
        }
        if (timed) {
            ZG_TRY(hipEventRecord(e1, s));
            ZG_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            ZG_TRY(hipEventElapsedTime(&ms, e0, e1));
            inf.kernel_ms = ms;
        }
        const bool final_odd = iters & 1;
        inf.iterations = iters;
        inf.last_l1 = h_st->last_l1;
        if (!final64)
            hipLaunchKernelGGL(pr_out_kernel<float>, vg, vb, 0, s, V, g->pr_id, final_odd ? r32 + V : r32, d_out);
        else
            hipLaunchKernelGGL(pr_out_kernel<double>, vg, vb, 0, s, V, g->pr_id, final_odd ? r64 + V : r64, d_out);
        ZG_TRY(hipGetLastError());

        if (check) {
            // One FP64 iteration from the returned vector: its l1 change is
            // ||G(r) - r||_1, and ||r - r*||_1 <= that / (1 - d).
            PrState st1 = st0;
            st1.base = h_st->base;  // constant term built from r's dangling mass
            st1.max_iter = 1;
            st1.tol = -1.0;
            st1.done = 0;
            ZG_TRY(hipMemcpyAsync(d_st + 1, &st1, sizeof(PrState), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((pr_init_kernel<double, double>), vg, vb, 0, s, V, g->live, g->out_ptr, g->pr_id,
                               (const double *)d_out, false, 0.0, r64, r64 + V, x64, x64 + V);
            ZG_TRY(hipGetLastError());
            PrArgs<double> a = pr_args<double>(g, r64, r64 + V, x64, x64 + V, d_st + 1, part);
            const PrShape sh = pr_shape<double>(g);
            rc = pr_allow_lds<double, ZG_PW_PR>(sh);
            if (rc != ZG_OK) break;
            pr_launch_iter<double, ZG_PW_PR>(sh, s, a, 0);  // the application
            pr_launch_iter<double, ZG_PW_PR>(sh, s, a, 1);  // its change, added up
            ZG_TRY(hipGetLastError());
            ZG_TRY(hipMemcpyAsync(h_st + 1, d_st + 1, sizeof(PrState), hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            inf.residual_l1 = h_st[1].last_l1;
            inf.error_bound_l1 = damping < 1.0 ? inf.residual_l1 / (1.0 - damping) : INFINITY;
        }
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc != ZG_OK) return rc;
    ZG_TRY(hipStreamSynchronize(s));
    if (info) *info = inf;
    return ZG_OK;
}

// For other parts of the library that already hold the handle's lock.
int zg_pagerank_dev(zg_graph *g, double damping, uint64_t max_iterations, double tolerance, int precision,
                    double *d_out, hipStream_t s, zg_pr_info *info)
{
    return zg_pagerank_impl(g, damping, max_iterations, tolerance, precision, d_out, s, info);
}

extern "C" int zg_pagerank_device(zg_graph *g, double damping, uint64_t max_iterations, double tolerance,
                                  int precision, double *d_out, void *stream, zg_pr_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!d_out && g->V) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    return zg_pagerank_impl(g, damping, max_iterations, tolerance, precision, d_out, s, info);
}

extern "C" int zg_pagerank(zg_graph *g, double damping, uint64_t max_iterations, double tolerance,
                           int precision, double *out, zg_pr_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out && g->V) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    double *d_out = nullptr;
    ZG_TRY_RC(zg_io_out(g, &d_out));
    int rc = zg_pagerank_impl(g, damping, max_iterations, tolerance, precision, d_out, g->stream, info);
    if (rc == ZG_OK && g->V) {
        hipError_t e = hipMemcpy(out, d_out, sizeof(double) * g->V, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = zg_hip_code(e);
    }
    return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free PageRank: zg_pagerank_device's computation with every
// buffer it would take from the handle's workspace (rank and gathered-value
// vectors, block partials, long-row segment partials, the iteration state)
// carved from one fixed device array instead.  The array is a module global:
// each device the library runs on holds its own copy, from the module's load
// until the process ends.  Its name is a host-side stand-in, not a device
// address; the address is taken once per device with hipGetSymbolAddress and
// only that pointer is passed to kernels and copies.  Nothing in it has to
// start at zero: every call writes the state, the vectors (init kernel, both
// parities) and the partials before reading them.  Calls on one device take
// turns (a per-device lock); two processes on one device each hold their own
// array.
#define ZG_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define ZG_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[ZG_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu[ZG_NOALLOC_MAX_DEVICES];
static unsigned char *g_scratch_ptr[ZG_NOALLOC_MAX_DEVICES];
static PrState g_scratch_state[ZG_NOALLOC_MAX_DEVICES][2];  // host mirror of the iteration state

// Device address of `scratch` on device `dev` (current), taken on the first
// call there.  The caller holds g_scratch_mu[dev].
static int scratch_address(int dev, unsigned char **out)
{
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return e != hipSuccess ? zg_hip_code(e) : ZG_ERR_HIP;
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return ZG_OK;
}

extern "C" int zg_pagerank_device_noalloc(zg_graph *g, double damping, uint64_t max_iterations, double tolerance,
                                          int precision, double *d_out, void *stream, zg_pr_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!d_out && g->V) return ZG_ERR_INVALID_ARG;
    const int prec = precision & 0xFF;
    if (prec != ZG_PR_FP32 && prec != ZG_PR_FP64 && prec != ZG_PR_MIXED) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->device < 0 || g->device >= ZG_NOALLOC_MAX_DEVICES) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> sl(g_scratch_mu[g->device]);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    // layout: [state][block partials, two parities]
    //         [FP64 r, x: both parities][FP32 r, x: both parities]
    const size_t V = g->V;
    const bool need64 = prec != ZG_PR_FP32 || !(precision & ZG_PR_SKIP_RESIDUAL);
    const bool need32 = prec != ZG_PR_FP64;
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_st = take(2 * sizeof(PrState));
    const size_t o_part = take(4 * sizeof(double) * pr_grid(g));
    const size_t o_64 = need64 ? take(4 * sizeof(double) * V) : 0;
    const size_t o_32 = need32 ? take(4 * sizeof(float) * V) : 0;
    if (off > ZG_NOALLOC_SCRATCH_BYTES) return ZG_ERR_SCRATCH;
    unsigned char *base = nullptr;
    ZG_TRY_RC(scratch_address(g->device, &base));
    PrMem m;
    m.d_st = (PrState *)(base + o_st);
    m.h_st = g_scratch_state[g->device];
    m.part = (double *)(base + o_part);
    if (need64) {
        m.r64 = (double *)(base + o_64);
        m.x64 = m.r64 + 2 * V;
    }
    if (need32) {
        m.r32 = (float *)(base + o_32);
        m.x32 = m.r32 + 2 * V;
    }
    m.timed = false;
    return zg_pagerank_impl(g, damping, max_iterations, tolerance, precision, d_out, s, info, &m);
}

// ---------------------------------------------------------------------------
// Eigenvector, Katz and alpha centrality: the same pull over the predecessor
// lists (collectNodes / buildInNeighbors, priv/zog/src/utils.zig:42-73, are
// the live ids ascending and, per vertex, its predecessors ascending with
// parallel arcs repeated), in FP64 with the PageRank kernels' row bins.
//   eigenvector (centrality.zig:1050-1201): y = sum over in(v) of x, then
//     y /= ||y||_2 (when nonzero), then in order: a zero y ends with the
//     uniform 1/sqrt(n); from the second iteration on, ||y - x_prev_prev||_2
//     below tol ends with the normalised average (y + x)/2; ||y - x||_2
//     below tol ends with y; after max_iterations the last y is returned.
//     Start 1 + i/1000 for the i-th live id; n <= 1 gives 1.0.
//   Katz (:1212-1267): x' = alpha * sum + beta from beta; alpha centrality
//     (:1278-1333): x' = alpha * sum from `initial`; the L1 change, then the
//     swap, then stop when it is below tol.

// y /= ||y|| on the new iterate; block partials of ||y - x||^2 and
// ||y - x_prev_prev||^2.
__global__ __launch_bounds__(256) void eig_norm_kernel(PrArgs<double> a, uint32_t V)
{
    __shared__ double red[2][4];
    const PrState *st = a.st;
    if (st->done) return;
    const unsigned long long it = st->iter;
    const unsigned k = (unsigned)(it % 3);
    const double *cur = k == 0 ? a.rb0 : (k == 1 ? a.rb1 : a.rb2);
    double *nw = k == 0 ? a.rb1 : (k == 1 ? a.rb2 : a.rb0);
    const double *pp = k == 0 ? a.rb2 : (k == 1 ? a.rb0 : a.rb1);
    const double s2 = st->sumsq, norm = sqrt(s2);
    double d1 = 0.0, d2 = 0.0;
    if (s2 > 0.0)
        for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < V; p += gridDim.x * blockDim.x) {
            const double y = nw[p] / norm;
            nw[p] = y;
            const double e1 = y - cur[p], e2 = y - pp[p];
            d1 += e1 * e1;
            d2 += e2 * e2;
        }
    d1 = zg_wave_sum(d1);
    d2 = zg_wave_sum(d2);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = d1;
        red[1][threadIdx.x >> 6] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        a.part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// The eigenvector iteration's end-of-step checks, in the reference's order.
// One block.
__global__ __launch_bounds__(1024) void eig_decide_kernel(PrArgs<double> a, uint32_t nparts)
{
    __shared__ double red[2][16];
    PrState *st = a.st;
    if (st->done) return;
    double d1 = 0.0, d2 = 0.0;
    for (uint32_t q = threadIdx.x; q < nparts; q += blockDim.x) {
        d1 += a.part[2 * q];
        d2 += a.part[2 * q + 1];
    }
    d1 = zg_wave_sum(d1);
    d2 = zg_wave_sum(d2);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = d1;
        red[1][threadIdx.x >> 6] = d2;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double D1 = 0.0, D2 = 0.0;
    for (int w = 0; w < 16; w++) {
        D1 += red[0][w];
        D2 += red[1][w];
    }
    const unsigned long long it = st->iter;
    if (st->sumsq == 0.0) {  // the zero vector
        st->outcome = ZG_EIG_ZERO;
        st->done = 1;
        return;
    }
    const double l2 = D1 > 0.0 ? sqrt(D1) : D1;
    const double l2b = D2 > 0.0 ? sqrt(D2) : D2;
    st->last_l1 = l2;
    if (it >= 1 && l2b < st->tol) {
        st->outcome = ZG_EIG_OSCILLATING;
        st->done = 1;
    } else if (l2 < st->tol) {
        st->outcome = ZG_EIG_CONVERGED;
        st->done = 1;
    } else {
        st->iter = it + 1;
        if (it + 1 >= st->max_iter) {
            st->outcome = ZG_EIG_CAP;
            st->done = 1;
        }
    }
}

// out[p] = (u[p] + w[p]) / 2 and block partials of its sum of squares.
__global__ __launch_bounds__(256) void eig_avg_kernel(uint32_t V, const double *u, const double *w, double *out,
                                                      double *part)
{
    __shared__ double red[4];
    double s2 = 0.0;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < V; p += gridDim.x * blockDim.x) {
        const double y = (u[p] + w[p]) / 2.0;
        out[p] = y;
        s2 += y * y;
    }
    s2 = zg_wave_sum(s2);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s2;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// out /= sqrt(sum of the partials) when positive (each block adds the
// partials in the same order).
__global__ __launch_bounds__(256) void eig_scale_kernel(uint32_t V, double *out, const double *part, uint32_t nparts)
{
    __shared__ double tot;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t q = 0; q < nparts; q++) t += part[q];
        tot = t;
    }
    __syncthreads();
    if (!(tot > 0.0)) return;
    const double norm = sqrt(tot);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < V; p += gridDim.x * blockDim.x) out[p] /= norm;
}

__global__ void pw_fill_kernel(uint32_t V, const uint8_t *live, double value, double *out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = live[v] ? value : 0.0;
}

// b0[pr_id[v]] = src[v] (or `fill`) on live rows, 0 elsewhere; b1 = 0 (and
// b2 = 0 when given).
__global__ void pw_init_kernel(uint32_t V, const uint8_t *live, const uint32_t *pr_id, const double *src, double fill,
                               double *b0, double *b1, double *b2)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t p = pr_id[v];
    b0[p] = live[v] ? (src ? src[v] : fill) : 0.0;
    b1[p] = 0.0;
    if (b2) b2[p] = 0.0;
}

// kind: ZG_PW_AFFINE (alpha, beta, start) or ZG_PW_RAW (eigenvector).
static int zg_power_impl(zg_graph *g, int kind, double alpha, double beta, double start, uint64_t max_iterations,
                         double tolerance, double *d_out, hipStream_t s, zg_pr_info *info)
{
    if (g->h_comp_off.size() > 2) return ZG_ERR_INVALID_ARG;  // one iteration per resource
    const uint32_t V = g->V, n = g->n_live;
    zg_pr_info inf;
    memset(&inf, 0, sizeof(inf));
    inf.precision = ZG_PR_FP64;
    const dim3 vb(256), vg((V + 255) / 256);
    if (n == 0 || (kind == ZG_PW_RAW && n == 1)) {
        // Katz / alpha of no nodes: no scores; eigenvector of <= 1 node: 1.0
        if (V) {
            if (n == 0) ZG_TRY(hipMemsetAsync(d_out, 0, sizeof(double) * V, s));
            else hipLaunchKernelGGL(pr_single_kernel, vg, vb, 0, s, V, g->live, d_out);
            ZG_TRY(hipGetLastError());
        }
        ZG_TRY(hipStreamSynchronize(s));
        inf.outcome = kind == ZG_PW_RAW ? ZG_EIG_TRIVIAL : 0;
        if (info) *info = inf;
        return ZG_OK;
    }
    ZG_TRY_RC(pr_ws_ensure(g));
    double *r64 = (double *)g->pr.r;
    double *b0 = r64, *b1 = r64 + V, *b2 = r64 + 2 * (size_t)V;
    PrState *d_st = (PrState *)g->pr.state;
    PrState *h_st = (PrState *)g->pr.h_state;
    PrState st0;
    memset(&st0, 0, sizeof(st0));
    st0.max_iter = max_iterations;
    st0.damping = alpha;
    st0.base = beta;
    st0.inv_n = 1.0 / (double)n;
    st0.tol = tolerance;
    st0.done = max_iterations == 0 ? 1u : 0u;
    st0.outcome = ZG_EIG_CAP;
    double *d_start = nullptr;
    if (kind == ZG_PW_RAW) {
        // the start vector 1 + i/1000 over the live ids in ascending order
        std::vector<double> h((size_t)V, 0.0);
        uint64_t i = 0;
        for (uint32_t v = 0; v < V; v++)
            if (g->h_live[v]) h[v] = 1.0 + (double)(i++) / 1000.0;
        ZG_TRY_RC(zg_dmalloc(&d_start, V));
        hipError_t e = hipMemcpyAsync(d_start, h.data(), sizeof(double) * V, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            zg_dfree(d_start);
            return zg_hip_code(e);
        }
    }
    hipLaunchKernelGGL(pw_init_kernel, vg, vb, 0, s, V, g->live, g->pr_id, (const double *)d_start, start, b0, b1,
                       kind == ZG_PW_RAW ? b2 : nullptr);
    ZG_TRY(hipMemcpyAsync(d_st, &st0, sizeof(PrState), hipMemcpyHostToDevice, s));
    ZG_TRY(hipGetLastError());
    PrArgs<double> a = pr_args<double>(g, b0, b1, b0, b1, d_st);  // Katz / alpha gather x = r
    a.rb0 = b0;
    a.rb1 = b1;
    a.rb2 = b2;
    const uint32_t grid = pr_grid(g);
    const PrShape sh = pr_shape<double>(g);
    int rc = kind == ZG_PW_RAW ? pr_allow_lds<double, ZG_PW_RAW>(sh) : pr_allow_lds<double, ZG_PW_AFFINE>(sh);
    if (rc != ZG_OK) {
        zg_dfree(d_start);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = zg_hip_code(hipEventCreate(&e0));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventCreate(&e1));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventRecord(e0, s));
    if (kind == ZG_PW_RAW) {
        uint64_t launched = 0;
        const uint64_t K = 8;
        while (rc == ZG_OK && launched < max_iterations) {
            const uint64_t k = std::min<uint64_t>(K, max_iterations - launched);
            for (uint64_t i = 0; i < k; i++) {
                const size_t p = zg_prof_begin(g, s);
                pr_launch_iter<double, ZG_PW_RAW>(sh, s, a);
                hipLaunchKernelGGL(pr_finish_kernel, dim3(1), dim3(ZG_PR_FINISH_THREADS), 0, s, a, sh.grid);
                hipLaunchKernelGGL(eig_norm_kernel, dim3(grid), dim3(256), 0, s, a, V);
                hipLaunchKernelGGL(eig_decide_kernel, dim3(1), dim3(1024), 0, s, a, grid);
                zg_prof_end(g, s, ZG_PROF_PR_ITER, p);
            }
            rc = zg_hip_code(hipGetLastError());
            launched += k;
            if (rc == ZG_OK) rc = zg_hip_code(hipMemcpyAsync(h_st, d_st, sizeof(PrState), hipMemcpyDeviceToHost, s));
            if (rc == ZG_OK) rc = zg_hip_code(hipStreamSynchronize(s));
            if (rc == ZG_OK && h_st->done) break;
        }
        if (rc == ZG_OK && max_iterations == 0) {
            rc = zg_hip_code(hipMemcpyAsync(h_st, d_st, sizeof(PrState), hipMemcpyDeviceToHost, s));
            if (rc == ZG_OK) rc = zg_hip_code(hipStreamSynchronize(s));
        }
    } else if (rc == ZG_OK) {
        rc = pr_run<double, ZG_PW_AFFINE>(g, s, a, h_st, max_iterations);
    }
    if (rc == ZG_OK) {
        // iterations that ran (the eigenvector counts its last, deciding one)
        const uint64_t it = h_st->iter;
        const int oc = h_st->outcome;
        uint64_t ran = it;
        if (kind == ZG_PW_RAW && max_iterations > 0 && oc != ZG_EIG_CAP) ran = it + 1;
        if (kind == ZG_PW_RAW) {
            if (g->ev_marks.size() > ran) g->ev_marks.resize(ran);
            rc = zg_prof_collect(g, s);
        } else {
            rc = pr_collect_run(g, s, ran);
        }
        inf.iterations = ran;
        inf.last_l1 = h_st->last_l1;
        inf.outcome = kind == ZG_PW_RAW ? oc : 0;
        const double *res = nullptr;
        if (kind != ZG_PW_RAW) {
            res = (it & 1) ? b1 : b0;
        } else if (oc == ZG_EIG_ZERO) {
            hipLaunchKernelGGL(pw_fill_kernel, vg, vb, 0, s, V, g->live, 1.0 / sqrt((double)n), d_out);
        } else if (oc == ZG_EIG_OSCILLATING) {
            const unsigned k3 = (unsigned)(it % 3);
            const double *cur = k3 == 0 ? b0 : (k3 == 1 ? b1 : b2);
            const double *nw = k3 == 0 ? b1 : (k3 == 1 ? b2 : b0);
            double *tmp = k3 == 0 ? b2 : (k3 == 1 ? b0 : b1);
            hipLaunchKernelGGL(eig_avg_kernel, dim3(grid), dim3(256), 0, s, V, nw, cur, tmp, g->pr.part);
            hipLaunchKernelGGL(eig_scale_kernel, dim3(grid), dim3(256), 0, s, V, tmp, (const double *)g->pr.part,
                               grid);
            res = tmp;
        } else {
            // converged: the new iterate; cap (or no iteration): the current one
            const unsigned k3 = (unsigned)((oc == ZG_EIG_CONVERGED ? it + 1 : it) % 3);
            res = k3 == 0 ? b0 : (k3 == 1 ? b1 : b2);
        }
        if (res) hipLaunchKernelGGL(pr_out_kernel<double>, vg, vb, 0, s, V, g->pr_id, res, d_out);
        if (rc == ZG_OK) rc = zg_hip_code(hipGetLastError());
        if (rc == ZG_OK) rc = zg_hip_code(hipEventRecord(e1, s));
        if (rc == ZG_OK) rc = zg_hip_code(hipEventSynchronize(e1));
        float ms = 0.f;
        if (rc == ZG_OK) rc = zg_hip_code(hipEventElapsedTime(&ms, e0, e1));
        inf.kernel_ms = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    zg_dfree(d_start);
    if (rc == ZG_OK && info) *info = inf;
    return rc;
}

static int zg_power_host(zg_graph *g, int kind, double alpha, double beta, double start, uint64_t max_iterations,
                         double tolerance, double *out, zg_pr_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out && g->V) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    double *d_out = nullptr;
    ZG_TRY_RC(zg_io_out(g, &d_out));
    int rc = zg_power_impl(g, kind, alpha, beta, start, max_iterations, tolerance, d_out, g->stream, info);
    if (rc == ZG_OK && g->V) rc = zg_hip_code(hipMemcpy(out, d_out, sizeof(double) * g->V, hipMemcpyDeviceToHost));
    return rc;
}

extern "C" int zg_eigenvector(zg_graph *g, uint64_t max_iterations, double tolerance, double *out, zg_pr_info *info)
{
    return zg_power_host(g, ZG_PW_RAW, 0.0, 0.0, 0.0, max_iterations, tolerance, out, info);
}

extern "C" int zg_katz(zg_graph *g, double alpha, double beta, uint64_t max_iterations, double tolerance, double *out,
                       zg_pr_info *info)
{
    return zg_power_host(g, ZG_PW_AFFINE, alpha, beta, beta, max_iterations, tolerance, out, info);
}

extern "C" int zg_alpha_centrality(zg_graph *g, double alpha, double initial, uint64_t max_iterations,
                                   double tolerance, double *out, zg_pr_info *info)
{
    return zg_power_host(g, ZG_PW_AFFINE, alpha, 0.0, initial, max_iterations, tolerance, out, info);
}

// ---------------------------------------------------------------------------
// Gather floor: the time of one pass sum_e x[pr_col[e]] over the resident
// predecessor columns (PageRank ids), one thread per arc, no row structure.
// It bounds what any pull iteration over these columns can reach.

template <typename T>
__global__ __launch_bounds__(256) void pw_gather_kernel(const uint32_t *__restrict__ col, const T *__restrict__ x,
                                                        uint32_t E, T *sink)
{
    T s = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) s += x[col[e]];
    if (s == (T)-1) sink[0] = s;  // never true (x is zero): keeps the loads
}

// The same pass in the iteration kernel's launch shape (blocks of NT
// threads, its grid) with its LDS copy of the most-gathered values (ids
// below K), ZG_PR_INFLIGHT gathers in flight per thread: the floor of an
// iteration that keeps that copy.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void pw_gather_lds_kernel(const uint32_t *__restrict__ col, const T *__restrict__ x,
                                                           uint32_t E, uint32_t K, T *sink)
{
    extern __shared__ uint4 pw_lds[];
    T *xs = reinterpret_cast<T *>(pw_lds);
    zg_lds_fill<T, NT>(xs, x, K);
    __syncthreads();
    constexpr int N = ZG_PR_INFLIGHT;
    const uint32_t step = gridDim.x * NT;
    T s = 0;
    uint32_t e = blockIdx.x * NT + threadIdx.x;
    for (; e + (N - 1) * step < E; e += N * step) {
        uint32_t u[N];
#pragma unroll
        for (int k = 0; k < N; k++) u[k] = col[e + k * step];
#pragma unroll
        for (int k = 0; k < N; k++) s += u[k] < K ? xs[u[k]] : x[u[k]];
    }
    for (; e < E; e += step) {
        const uint32_t u = col[e];
        s += u < K ? xs[u] : x[u];
    }
    if (s == (T)-1) sink[0] = s;  // never true (x is zero): keeps the loads
}

template <typename T>
static int pw_gather_lds_launch(const PrShape &sh, hipStream_t s, const uint32_t *col, const T *x, uint32_t E, T *sink,
                                bool allow)
{
    const void *f = sh.threads == 1024  ? reinterpret_cast<const void *>(&pw_gather_lds_kernel<T, 1024>)
                    : sh.threads == 512 ? reinterpret_cast<const void *>(&pw_gather_lds_kernel<T, 512>)
                                        : reinterpret_cast<const void *>(&pw_gather_lds_kernel<T, 256>);
    if (allow && sh.lds > 65536) ZG_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds));
    if (sh.threads == 1024)
        hipLaunchKernelGGL((pw_gather_lds_kernel<T, 1024>), dim3(sh.grid), dim3(1024), sh.lds, s, col, x, E, sh.K, sink);
    else if (sh.threads == 512)
        hipLaunchKernelGGL((pw_gather_lds_kernel<T, 512>), dim3(sh.grid), dim3(512), sh.lds, s, col, x, E, sh.K, sink);
    else
        hipLaunchKernelGGL((pw_gather_lds_kernel<T, 256>), dim3(sh.grid), dim3(256), sh.lds, s, col, x, E, sh.K, sink);
    return ZG_OK;
}

extern "C" int zg_gather_floor(zg_graph *g, int precision, int reps, double *us_per_pass)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    const bool with_lds = (precision & ZG_FLOOR_LDS_COPY) != 0;
    precision &= ~ZG_FLOOR_LDS_COPY;
    if ((precision != 32 && precision != 64) || reps < 1 || !us_per_pass) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = g->stream;
    const uint32_t V = g->V, E = (uint32_t)g->E;
    // the iteration kernel's shape and LDS copy in this precision
    const PrShape sh = precision == 32 ? pr_shape<float>(g) : pr_shape<double>(g);
    void *x = nullptr;
    ZG_TRY(hipMalloc(&x, (size_t)(precision / 8) * (V ? V : 1) + 64));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = zg_hip_code(hipMemsetAsync(x, 0, (size_t)(precision / 8) * (V ? V : 1) + 64, s));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventCreate(&e0));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventCreate(&e1));
    const dim3 grid(8192), blk(256);
    for (int pass = 0; pass < 2 && rc == ZG_OK; pass++) {  // the first pass warms the caches
        if (pass) rc = zg_hip_code(hipEventRecord(e0, s));
        for (int i = 0; i < reps && rc == ZG_OK && E; i++) {
            if (with_lds && precision == 32)
                rc = pw_gather_lds_launch<float>(sh, s, g->pr_col, (const float *)x, E, (float *)x + (V ? V : 1),
                                                 pass == 0 && i == 0);
            else if (with_lds)
                rc = pw_gather_lds_launch<double>(sh, s, g->pr_col, (const double *)x, E, (double *)x + (V ? V : 1),
                                                  pass == 0 && i == 0);
            else if (precision == 32)
                hipLaunchKernelGGL(pw_gather_kernel<float>, grid, blk, 0, s, g->pr_col, (const float *)x, E,
                                   (float *)x + (V ? V : 1));
            else
                hipLaunchKernelGGL(pw_gather_kernel<double>, grid, blk, 0, s, g->pr_col, (const double *)x, E,
                                   (double *)x + (V ? V : 1));
        }
        if (rc == ZG_OK) rc = zg_hip_code(hipGetLastError());
        if (pass && rc == ZG_OK) rc = zg_hip_code(hipEventRecord(e1, s));
    }
    float ms = 0.f;
    if (rc == ZG_OK) rc = zg_hip_code(hipEventSynchronize(e1));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventElapsedTime(&ms, e0, e1));
    if (rc == ZG_OK) *us_per_pass = 1000.0 * ms / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(x);
    return rc;
}
