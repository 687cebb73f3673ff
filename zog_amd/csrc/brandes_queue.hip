// Unweighted Brandes, per-source queue engine, for high-diameter graphs.
//
// Reference: priv/zog/src/centrality.zig:694-798.  The reference keeps, per
// source, a BFS queue whose order is also the stack of the dependency sweep;
// this engine keeps exactly that per source, for S sources at once:
//   rec[s][v]  one 16-byte record {depth (u32), exponent (i32), value}: the
//              value is first the path-count mantissa sig (FP64): sigma =
//              sig * 2^exponent, sig in [1, 2) -- every count has its own
//              exponent, so a level whose counts run from 1 to 2^4000 (a
//              2048 x 2048 grid) is still represented; counts below 2^53 are
//              exact -- and, once the dependency sweep has passed v, cc =
//              (1 + delta) / sig, the dependency term c = (1+delta)/sigma
//              scaled by 2^exponent (sig(v) is read only by v itself then)
//   queue[s][i], lstart[L][s] (queue offset of level L), qlen[s].
// One lane handles one (source, vertex) pair; a wavefront always holds pairs
// of a single source, so discoveries are appended to that source's queue
// with one atomic per wavefront (ballot + prefix).
//
// Why a second engine: the 64-lane engine (brandes.hip) shares one row
// access among the group's sources, which pays when they reach a vertex at
// the same level (small-world graphs).  On a high-diameter graph with
// scattered sources each vertex sits at ~64 different depths, so a row
// access serves one lane; here every access serves its own pair.
//
// Forward level d (one launch): for v at depth d, first its path count,
// pulled from its predecessors at depth d-1 over the predecessor CSR (the
// sum of their counts, parallel arcs counted once each, aligned to the
// largest exponent, then normalised), then for each arc v->w with depth[w]
// unset, a CAS to d+1 (the winner appends w).  Backward level d: delta(v) =
// sig(v) * sum over successors w at depth d+1 of cc(w) 2^(e_v - e_w)
// (e_w >= e_v since sigma_w >= sigma_v, so only negligible terms can
// underflow), cc(v) = (1 + delta(v)) / sig(v); delta(v) is kept per (source,
// vertex) and added into bc[v] source by source after the batch, so the
// result repeats bit for bit.
//
// The level loop is device-driven: every kernel reads its per-source ranges
// from lstart/qlen, so the host queues ZG_Q_LEVELS_PER_SYNC levels at a time
// and syncs only to find the last level.
#include "build_util.h"
#include "queue_map.h"

#include <algorithm>
#include <climits>

#define ZG_Q_MAX_SOURCES 1024  // sources per batch (LDS range tables)
#ifndef ZG_Q_GRID
#define ZG_Q_GRID 1024
#endif
#define ZG_Q_LEVELS_PER_SYNC 32
#define ZG_Q_WIDE 32           // vertices with more predecessors are summed by the whole wavefront
#define ZG_Q_BYTES_PER_SV 28   // workspace bytes per (source, vertex)

struct BcqDev {
    uint32_t V, S, lcap;
    const uint32_t *out_ptr, *out_col, *in_ptr, *in_col;
    const uint4 *adj;               // [V][4] inline adjacency (ZG_Q_INLINE), or null
    zg_qrec *rec;                   // [S][V] {depth, exponent, sig or cc}
    uint32_t *queue;                // [S][V]
    uint32_t *lstart;               // [lcap][S]
    uint32_t *cpre;                 // [lcap][S + 1] chunk prefix of each level (total last)
    uint32_t *qlen;                 // [S]
    uint32_t *newcnt;               // [lcap] entries added at each level
    const uint32_t *src;            // [S]
    const double *wts;              // [S] source weights, or null (all 1)
    double *dlt;                    // [S][V] weighted dependencies of the batch (summed by bcq_reduce_kernel)
    double *bc;                     // [V]
    unsigned long long *stats;      // [2] arcs, reached
};

#define QROW(s) ((size_t)(s) * a.V)

// 2^k * x (ldexp saturates to 0 or infinity; exponents of path counts stay
// far inside the int range, and only those of reached vertices are used).
__device__ __forceinline__ double zg_scale(double x, int k)
{
    return ldexp(x, k);
}

// Inline adjacency of a vertex with at most ZG_Q_INLINE predecessors and
// successors: one 64-byte record {counts, predecessors in predecessor-CSR
// order, successors in successor-CSR order}, so a level step reads one line
// for both lists instead of the four CSR arrays' lines.  Word 0 is
// in-degree | out-degree << 16, or ZG_NONE for a vertex the CSR serves.
#define ZG_Q_INLINE 7

__global__ void bcq_adj_kernel(uint32_t V, const uint32_t *in_ptr, const uint32_t *in_col, const uint32_t *out_ptr,
                               const uint32_t *out_col, uint4 *adj)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const uint32_t i0 = in_ptr[v], id = in_ptr[v + 1] - i0, o0 = out_ptr[v], od = out_ptr[v + 1] - o0;
        uint32_t w[16];
#pragma unroll
        for (int k = 0; k < 16; k++) w[k] = 0u;
        if (id <= ZG_Q_INLINE && od <= ZG_Q_INLINE) {
            w[0] = id | (od << 16);
#pragma unroll
            for (int k = 0; k < ZG_Q_INLINE; k++) {
                if ((uint32_t)k < id) w[1 + k] = in_col[i0 + k];
                if ((uint32_t)k < od) w[8 + k] = out_col[o0 + k];
            }
        } else {
            w[0] = ZG_NONE;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) adj[(size_t)v * 4 + q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    }
}

// One 16-byte record as a single vector load, so the compiler does not
// split it into a depth load and a later load of the rest under the depth
// test (which would add a dependent round trip).
__device__ __forceinline__ zg_qrec zg_qload(const zg_qrec *p)
{
    const uint4 r = *reinterpret_cast<const uint4 *>(p);
    zg_qrec q;
    q.depth = r.x;
    q.exp = r.y;
    q.val = __hiloint2double((int)r.w, (int)r.z);
    return q;
}

// A vertex's inline record, loaded whole (four 16-byte loads of one line).
struct ZgQAdj {
    uint32_t w[16];
    __device__ __forceinline__ void load(const uint4 *adj, uint32_t v)
    {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 x = adj[(size_t)v * 4 + q];
            w[4 * q] = x.x;
            w[4 * q + 1] = x.y;
            w[4 * q + 2] = x.z;
            w[4 * q + 3] = x.w;
        }
    }
    __device__ __forceinline__ bool inl() const { return w[0] != ZG_NONE; }
    __device__ __forceinline__ uint32_t ideg() const { return w[0] & 0xFFFFu; }
    __device__ __forceinline__ uint32_t odeg() const { return w[0] >> 16; }
    // predecessor / successor k of batch b (k = 4 b + t, t static)
    __device__ __forceinline__ uint32_t pred(int b, int t) const { return b == 0 ? w[1 + t] : (t < 3 ? w[5 + t] : 0u); }
    __device__ __forceinline__ uint32_t succ(int b, int t) const { return b == 0 ? w[8 + t] : (t < 3 ? w[12 + t] : 0u); }
};

__global__ void bcq_init_kernel(BcqDev a)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.S) return;
    const uint32_t v = a.src[s];
    if (v == ZG_NONE) {
        a.qlen[s] = 0;
        a.lstart[s] = 0;
        a.lstart[a.S + s] = 0;
        return;
    }
    a.rec[QROW(s) + v] = zg_qrec{0u, 0u, 1.0};
    a.queue[QROW(s)] = v;
    a.qlen[s] = 1;
    a.lstart[s] = 0;        // level 0 starts at 0
    a.lstart[a.S + s] = 1;  // level 1 starts at 1
    atomicAdd(&a.stats[0], (unsigned long long)(a.out_ptr[v + 1] - a.out_ptr[v]));
    atomicAdd(&a.stats[1], 1ull);
}

// The per-source ranges and chunk prefix of level d into LDS (published by
// the close step), and the total; false when this workgroup has no chunk
// of the level (then it touches nothing and leaves).
__device__ __forceinline__ bool bcq_level_tables(const BcqDev &a, int d, uint32_t *s_lo, uint32_t *s_hi,
                                                 uint32_t *s_pre, uint32_t &total, ZgQWalk &wk)
{
    const uint32_t *pre = a.cpre + (size_t)d * (a.S + 1);
    total = pre[a.S];
    wk = zg_q_walk(total);
    // the block's first chunk is its wavefront 0's
    const uint32_t first = wk.c - (threadIdx.x >> 6);
    if (first >= wk.end) return false;
    // 256-thread blocks, at most ZG_Q_MAX_SOURCES sources: every load of the
    // tables is issued before the first LDS write (indexes past the last
    // source are clamped to it, so their writes repeat its values)
    constexpr int U = (ZG_Q_MAX_SOURCES + 255) / 256;
    const uint32_t *lo = a.lstart + (size_t)d * a.S, *hi = lo + a.S;
    uint32_t tp[U], tl[U], th[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t s = min(threadIdx.x + 256u * u, a.S - 1);
        tp[u] = pre[s];
        tl[u] = lo[s];
        th[u] = hi[s];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const uint32_t s = min(threadIdx.x + 256u * u, a.S - 1);
        s_pre[s] = tp[u];
        s_lo[s] = tl[u];
        s_hi[s] = th[u];
    }
    if (threadIdx.x == 0) s_pre[a.S] = total;
    __syncthreads();
    return true;
}

__global__ __launch_bounds__(256) void bcq_forward_kernel(BcqDev a, int d)
{
    __shared__ uint32_t s_lo[ZG_Q_MAX_SOURCES], s_hi[ZG_Q_MAX_SOURCES], s_pre[ZG_Q_MAX_SOURCES + 1];
    __shared__ unsigned long long red[2][4];
    uint32_t total;
    ZgQWalk wk;
    if (!bcq_level_tables(a, d, s_lo, s_hi, s_pre, total, wk)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t dprev = (uint32_t)(d - 1);
    unsigned long long arcs = 0, reached = 0;
    for (uint32_t c = wk.c; c < wk.end; c += wk.stride) {
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        zg_qrec *rec = a.rec + QROW(s);
        // a lane past the level's end shadows its last entry, so every load
        // below is unconditional (in flight together, one wait) and in bounds;
        // only valid lanes write
        const uint32_t v = a.queue[QROW(s) + (valid ? idx : s_hi[s] - 1)];
        ZgQAdj J{};
        bool inl = false;
        if (a.adj) {
            J.load(a.adj, v);
            inl = valid && J.inl();
        }
        // depths of the inline predecessors, kept so that a successor that is
        // also a predecessor (every arc of a symmetric graph) needs no second
        // record load: depths only go from unset to set during the sweep, so
        // a value read here is as good as one read later (the claim is a CAS)
        uint32_t unset = 0u;  // bit k: inline predecessor k was unreached
        uint32_t npred = 0u;  // inline predecessors whose depth was read
        if (d > 0) {
            // path count of v from its predecessors at depth d-1
            const uint32_t i0 = valid && !inl ? a.in_ptr[v] : 0;
            const uint32_t ideg = !valid ? 0 : inl ? J.ideg() : a.in_ptr[v + 1] - i0;
            // narrow rows: one lane, the sum rescaled whenever a larger exponent appears
            int E = 0;
            double x = 0.0;
            auto fold = [&](const zg_qrec &q) {
                if (q.depth != dprev) return;
                const int eu = (int)q.exp;
                if (x == 0.0) E = eu;
                else if (eu > E) {
                    x = zg_scale(x, E - eu);
                    E = eu;
                }
                x += zg_scale(q.val, eu - E);
            };
            // a second batch of four only where some lane has more than four
            const bool more = __ballot(inl && ideg > 4) != 0;
            if (inl) {
                // the record's predecessors, four at a time, in the same order;
                // the loads are unconditional (v's own record past ideg, whose
                // depth is d, so the fold skips it) and masked afterwards
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    if (b == 1 && !more) break;
                    uint32_t u[4];
                    zg_qrec q[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) u[t] = (uint32_t)(4 * b + t) < ideg ? J.pred(b, t) : v;
#pragma unroll
                    for (int t = 0; t < 4; t++) q[t] = zg_qload(rec + u[t]);
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        if ((uint32_t)(4 * b + t) >= ideg) continue;
                        fold(q[t]);
                        if (q[t].depth == ZG_NONE) unset |= 1u << (4 * b + t);
                    }
                }
                npred = ideg;
            } else if (ideg <= ZG_Q_WIDE) {
                // four arcs at a time: their loads overlap instead of chaining
                for (uint32_t k0 = 0; k0 < ideg; k0 += 4) {
                    uint32_t u[4];
                    zg_qrec q[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) u[t] = k0 + t < ideg ? a.in_col[i0 + k0 + t] : 0u;
#pragma unroll
                    for (int t = 0; t < 4; t++) q[t] = k0 + t < ideg ? rec[u[t]] : zg_qrec{ZG_NONE, 0u, 0.0};
#pragma unroll
                    for (int t = 0; t < 4; t++) fold(q[t]);
                }
            }
            // wide rows: the whole wavefront, largest exponent first, then the sum
            for (uint64_t wide = __ballot(ideg > ZG_Q_WIDE); wide; wide &= wide - 1) {
                const int j = __builtin_ctzll(wide);
                const uint32_t ij = zg_readlane32(i0, j), dj = zg_readlane32(ideg, j);
                int emax = INT_MIN;
                for (uint32_t k = lane; k < dj; k += 64) {
                    const zg_qrec q = rec[a.in_col[ij + k]];
                    if (q.depth == dprev) emax = max(emax, (int)q.exp);
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) emax = max(emax, __shfl_xor(emax, off, 64));
                double part = 0.0;
                for (uint32_t k = lane; k < dj; k += 64) {
                    const zg_qrec q = rec[a.in_col[ij + k]];
                    if (q.depth == dprev) part += zg_scale(q.val, (int)q.exp - emax);
                }
                part = zg_wave_sum(part);
                if (lane == j) {
                    x = part;
                    E = emax;
                }
            }
            if (valid) {
                const int k = ilogb(x);  // x >= 1: v has a predecessor at depth d-1
                rec[v] = zg_qrec{(uint32_t)d, (uint32_t)(E + k), ldexp(x, -k)};  // depth unchanged
            }
        }
        const uint32_t e0 = valid && !inl ? a.out_ptr[v] : 0;
        const uint32_t deg = !valid ? 0 : inl ? J.odeg() : a.out_ptr[v + 1] - e0;
        if (valid && d > 0) {  // level 0 (the source) is counted by bcq_init_kernel
            arcs += deg;
            reached += 1;
        }
        uint32_t maxdeg = deg;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) maxdeg = max(maxdeg, (uint32_t)__shfl_xor((int)maxdeg, off, 64));
        for (uint32_t k0 = 0; k0 < maxdeg; k0 += 4) {
            // the heads and their depths of four arcs first, then the claims
            uint32_t w[4], dw[4];
#pragma unroll
            for (int t = 0; t < 4; t++)
                w[t] = k0 + t < deg ? (inl ? J.succ((int)(k0 >> 2), t) : a.out_col[e0 + k0 + t]) : 0u;
            bool need[4], any_need = false;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                bool known = false;
                dw[t] = 0u;
#pragma unroll
                for (int k = 0; k < ZG_Q_INLINE; k++)
                    if ((uint32_t)k < npred && J.w[1 + k] == w[t]) {
                        dw[t] = (unset >> k) & 1u ? ZG_NONE : 0u;
                        known = true;
                    }
                need[t] = k0 + t < deg && !known;
                any_need |= need[t];
            }
            if (any_need) {
                // the depths not known yet, four loads in flight
                uint32_t ld[4];
#pragma unroll
                for (int t = 0; t < 4; t++) ld[t] = rec[need[t] ? w[t] : v].depth;
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (need[t]) dw[t] = ld[t];
            }
            // the four claims in flight together, then one append for the
            // wavefront's winners (arc by arc, lanes in order within an arc)
            uint32_t prev[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                prev[t] = 0u;  // any value but ZG_NONE
                if (k0 + t < deg && dw[t] == ZG_NONE)
                    prev[t] = atomicCAS(&rec[w[t]].depth, ZG_NONE, (uint32_t)(d + 1));
            }
            bool won[4];
#pragma unroll
            for (int t = 0; t < 4; t++) won[t] = prev[t] == ZG_NONE;
            uint64_t bw[4];
            uint32_t nwon = 0;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                bw[t] = __ballot(won[t]);
                nwon += (uint32_t)__popcll(bw[t]);
            }
            if (nwon) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&a.qlen[s], nwon);
                base = zg_readlane32(base, 0);
                const uint64_t below = (1ull << lane) - 1;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (won[t]) a.queue[QROW(s) + base + __popcll(bw[t] & below)] = w[t];
                    base += (uint32_t)__popcll(bw[t]);
                }
            }
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
        const unsigned long long A = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const unsigned long long R = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if (A) atomicAdd(&a.stats[0], A);
        if (R) atomicAdd(&a.stats[1], R);
    }
}

// Close level L (just appended): the offset of level L+1, the level's size,
// and the prefix over sources of the level's 64-entry chunks (the forward
// and backward launches of level L read it instead of each workgroup
// scanning the ranges again).  One block of 1024 threads; S <= 1024.
__global__ __launch_bounds__(1024) void bcq_close_kernel(BcqDev a, int L)
{
    __shared__ uint32_t wtot[16];
    const uint32_t s = threadIdx.x;
    const int lane = s & 63, wv = s >> 6;
    uint32_t added = 0, cnt = 0;
    if (s < a.S) {
        const uint32_t q = a.qlen[s];
        a.lstart[(size_t)(L + 1) * a.S + s] = q;
        added = q - a.lstart[(size_t)L * a.S + s];
        cnt = (added + 63) / 64;
    }
    const uint32_t incl = zg_wave_incl_scan(cnt, lane);
    if (lane == 63) wtot[wv] = incl;
    const uint32_t tot_added = (uint32_t)zg_wave_sum_u64(added);
    if (lane == 0 && tot_added) atomicAdd(&a.newcnt[L], tot_added);
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wv; w++) off += wtot[w];
    uint32_t *pre = a.cpre + (size_t)L * (a.S + 1);
    if (s < a.S) pre[s] = off + incl - cnt;
    if (s == 0) {
        uint32_t t = 0;
        for (int w = 0; w < 16; w++) t += wtot[w];
        pre[a.S] = t;
    }
}

__global__ __launch_bounds__(256) void bcq_backward_kernel(BcqDev a, int d)
{
    __shared__ uint32_t s_lo[ZG_Q_MAX_SOURCES], s_hi[ZG_Q_MAX_SOURCES], s_pre[ZG_Q_MAX_SOURCES + 1];
    uint32_t total;
    ZgQWalk wk;
    if (!bcq_level_tables(a, d, s_lo, s_hi, s_pre, total, wk)) return;
    const int lane = threadIdx.x & 63;
    const uint32_t dnext = (uint32_t)(d + 1);
    for (uint32_t c = wk.c; c < wk.end; c += wk.stride) {
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        zg_qrec *rec = a.rec + QROW(s);
        // a lane past the level's end shadows its last entry (unconditional
        // loads, in flight together) and writes nothing
        const uint32_t v = a.queue[QROW(s) + (valid ? idx : s_hi[s] - 1)];
        const double ws = a.wts ? a.wts[s] : 1.0;  // loaded with the rest, used at the end
        double S = 0.0;
        ZgQAdj J{};
        bool inl = false;
        if (a.adj) {
            J.load(a.adj, v);
            inl = J.inl();
        }
        zg_qrec me;
        const bool more = __ballot(inl && J.odeg() > 4) != 0;
        if (inl) {
            // the record's successors, four at a time, in the same order;
            // past the out-degree the loads read v's own record (depth d,
            // skipped), which is needed anyway and comes in the same round
            const uint32_t od = J.odeg();
#pragma unroll
            for (int b = 0; b < 2; b++) {
                if (b == 1 && !more) break;
                uint32_t w[4];
                zg_qrec q[4];
#pragma unroll
                for (int t = 0; t < 4; t++) w[t] = (uint32_t)(4 * b + t) < od ? J.succ(b, t) : v;
#pragma unroll
                for (int t = 0; t < 4; t++) q[t] = zg_qload(rec + w[t]);
                if (b == 0) me = zg_qload(rec + v);
                const int ev = (int)me.exp;
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if ((uint32_t)(4 * b + t) < od && q[t].depth == dnext) S += zg_scale(q[t].val, ev - (int)q[t].exp);
            }
        } else {
            me = rec[v];
            const int ev = (int)me.exp;
            for (uint32_t e = a.out_ptr[v], e1 = a.out_ptr[v + 1]; e < e1; e += 4) {
                // four arcs at a time: heads, then the records (depth, exponent
                // and, at depth d+1, the dependency term)
                uint32_t w[4];
                zg_qrec q[4];
#pragma unroll
                for (int t = 0; t < 4; t++) w[t] = e + t < e1 ? a.out_col[e + t] : 0u;
#pragma unroll
                for (int t = 0; t < 4; t++) q[t] = e + t < e1 ? rec[w[t]] : zg_qrec{ZG_NONE, 0u, 0.0};
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (q[t].depth == dnext) S += zg_scale(q[t].val, ev - (int)q[t].exp);
            }
        }
        if (!valid) continue;
        const double sv = me.val;
        const double delta = sv * S;
        rec[v].val = (1.0 + delta) / sv;  // from here on the dependency term
        // kept per (source, vertex) and summed in source order after the batch
        // (levels >= 1 only: v is never the source here)
        a.dlt[QROW(s) + v] = a.wts ? delta * ws : delta;
    }
}

// bc[v] += the batch's dependencies of v, source by source in slot order
// (one thread per vertex, so the sum has one order whatever the timing).
__global__ void bcq_reduce_kernel(uint32_t V, uint32_t nb, const double *dlt, double *bc)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double acc = bc[v];
    for (uint32_t s = 0; s < nb; s++) acc += dlt[(size_t)s * V + v];
    bc[v] = acc;
}

// perm: internal id of each vertex (locality order), or null
__global__ void bcq_out_kernel(uint32_t V, const double *bc, const uint8_t *live, const uint32_t *perm, double *out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = live[v] ? bc[perm ? perm[v] : v] : 0.0;
}

// ---------------------------------------------------------------------------
// Locality order (ZG_TUNE_QUEUE_ORDER, on by default).  A level touches, for
// every queued (source, vertex) pair, the 16-byte records of the vertex's
// neighbours.  With the vertices numbered in the order of a sequential
// breadth-first search (over each vertex's successors, then its
// predecessors, in arc order, from the lowest live vertex with an arc;
// vertices it does not reach follow in id order), the rings of another
// source's search run along stretches of consecutive ids, so neighbouring
// records share lines (2048 x 2048 grid, 512 sources: 15% less time than
// row-major ids; the same levels in an arbitrary order inside each level
// gain 4%).  The engine runs on a copy of the CSR under these ids in which
// every row keeps its arc order, so every sum, and the result, is the same
// bit for bit.
//
// A level [lo, hi) of the search yields the sequential order exactly: a new
// vertex belongs to the first (parent, arc) pair that reaches it -- the
// smallest key (parent - lo) << 32 | arc, by a 64-bit atomic min -- and each
// parent places its children in arc order after the children of the parents
// before it (a scan of the per-parent counts).

__global__ void bcq_start_kernel(uint32_t V, const uint32_t *out_ptr, const uint32_t *in_ptr, const uint8_t *live,
                                 uint32_t *start)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        if (live[v] && (out_ptr[v + 1] > out_ptr[v] || in_ptr[v + 1] > in_ptr[v])) atomicMin(start, v);
}

__global__ void bcq_seed_kernel(const uint32_t *start, uint32_t *perm, uint32_t *inv)
{
    const uint32_t v = *start;
    if (v != ZG_NONE) {
        perm[v] = 0;
        inv[0] = v;
    }
}

// Arc k of parent i (successors first, then predecessors), or ZG_NONE past
// its lists.
struct BcqSearchArcs {
    const uint32_t *out_ptr, *out_col, *in_ptr, *in_col;
};

// phase 0: claim (atomic min of the key); 1: count the arcs that won;
// 2: number the winners in arc order from base[i - lo]
__global__ void bcq_bfs_level_kernel(BcqSearchArcs g, uint32_t lo, uint32_t hi, int phase, uint32_t *perm, uint32_t *inv,
                                     unsigned long long *best, uint32_t *cnt, const uint32_t *base)
{
    for (uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
        const uint32_t v = inv[i];
        uint32_t k = 0, won = 0, at = phase == 2 ? hi + base[i - lo] : 0;
        for (int dir = 0; dir < 2; dir++) {
            const uint32_t *ptr = dir ? g.in_ptr : g.out_ptr, *col = dir ? g.in_col : g.out_col;
            for (uint32_t e = ptr[v]; e < ptr[v + 1]; e++, k++) {
                const uint32_t w = col[e];
                if (perm[w] != ZG_NONE) continue;
                const unsigned long long key = ((unsigned long long)(i - lo) << 32) | k;
                if (phase == 0) {
                    if (best[w] > key) atomicMin(&best[w], key);
                } else if (best[w] == key) {
                    if (phase == 1) {
                        won++;
                    } else {
                        perm[w] = at;  // only this (parent, arc) has w's key
                        inv[at++] = w;
                    }
                }
            }
        }
        if (phase == 1) cnt[i - lo] = won;
    }
    if (phase == 1 && blockIdx.x == 0 && threadIdx.x == 0) cnt[hi - lo] = 0;
}

__global__ void bcq_rest_flag_kernel(uint32_t V, const uint32_t *perm, uint32_t *flag)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += gridDim.x * blockDim.x)
        flag[v] = v < V && perm[v] == ZG_NONE ? 1u : 0u;
}

// The vertices the search did not reach, in id order after the others.
__global__ void bcq_rest_kernel(uint32_t V, const uint32_t *flag, const uint32_t *rank, uint32_t base, uint32_t *perm,
                                uint32_t *inv)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        if (flag[v]) {
            perm[v] = base + rank[v];
            inv[base + rank[v]] = v;
        }
}

__global__ void bcq_pdeg_kernel(uint32_t V, const uint32_t *inv, const uint32_t *ptr, uint32_t *deg)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= V; i += gridDim.x * blockDim.x)
        deg[i] = i < V ? ptr[inv[i] + 1] - ptr[inv[i]] : 0u;
}

// One wavefront per row: row i is row inv[i] with its heads renumbered.
__global__ void bcq_prow_kernel(uint32_t V, const uint32_t *inv, const uint32_t *perm, const uint32_t *ptr,
                                const uint32_t *col, const uint32_t *nptr, uint32_t *ncol)
{
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < V; i += (gridDim.x * blockDim.x) >> 6) {
        const uint32_t v = inv[i], a = ptr[v], k = ptr[v + 1] - a, b = nptr[i];
        for (uint32_t j = lane; j < k; j += 64) ncol[b + j] = perm[col[a + j]];
    }
}

__global__ void bcq_map_src_kernel(uint32_t S, const uint32_t *perm, uint32_t *src)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S && src[s] != ZG_NONE) src[s] = perm[src[s]];
}

void zg_bcq_order_release(zg_graph *g)
{
    zg_dfree(g->qperm);
    zg_dfree(g->q_out_ptr);
    zg_dfree(g->q_out_col);
    zg_dfree(g->q_in_ptr);
    zg_dfree(g->q_in_col);
    zg_dfree(g->qadj);
}

static int bcq_build_order_into(zg_graph *g, hipStream_t s)
{
    const uint32_t V = g->V;
    const uint64_t E = g->E;
    const size_t V1 = (size_t)V + 1;
    ZgScratch sc;
    uint32_t *inv, *start, *flag, *rank, *deg;
    unsigned long long *best;
    ZG_TRY_RC(sc.get(&inv, V));
    ZG_TRY_RC(sc.get(&start, 1));
    ZG_TRY_RC(sc.get(&flag, V1));
    ZG_TRY_RC(sc.get(&rank, V1));
    ZG_TRY_RC(sc.get(&deg, V1));
    ZG_TRY_RC(sc.get(&best, V));
    size_t scan_cap = zg_scan_bytes(V1);
    void *scan_tmp = nullptr;
    ZG_TRY_RC(sc.get((char **)&scan_tmp, scan_cap ? scan_cap : 1));
    ZG_TRY_RC(zg_dmalloc(&g->qperm, V));
    ZG_TRY(hipMemsetAsync(g->qperm, 0xFF, sizeof(uint32_t) * V, s));
    ZG_TRY(hipMemsetAsync(best, 0xFF, sizeof(unsigned long long) * V, s));
    ZG_TRY(hipMemsetAsync(start, 0xFF, sizeof(uint32_t), s));
    const dim3 blk(256), grd(std::min<uint32_t>((V + 255) / 256, 8192));
    hipLaunchKernelGGL(bcq_start_kernel, grd, blk, 0, s, V, g->out_ptr, g->in_ptr, g->live, start);
    hipLaunchKernelGGL(bcq_seed_kernel, dim3(1), dim3(1), 0, s, start, g->qperm, inv);
    ZG_TRY(hipGetLastError());
    uint32_t h_start = ZG_NONE;
    ZG_TRY(hipMemcpyAsync(&h_start, start, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    const BcqSearchArcs arcs{g->out_ptr, g->out_col, g->in_ptr, g->in_col};
    // level [lo, hi): numbered; its children get [hi, hi + count)
    uint32_t lo = 0, hi = h_start != ZG_NONE ? 1u : 0u;
    while (hi > lo) {
        const uint32_t n = hi - lo;
        const dim3 lg(std::min<uint32_t>((n + 255) / 256, 4096));
        hipLaunchKernelGGL(bcq_bfs_level_kernel, lg, blk, 0, s, arcs, lo, hi, 0, g->qperm, inv, best, flag, rank);
        hipLaunchKernelGGL(bcq_bfs_level_kernel, lg, blk, 0, s, arcs, lo, hi, 1, g->qperm, inv, best, flag, rank);
        ZG_TRY(hipGetLastError());
        ZG_TRY_RC(zg_exclusive_scan_tmp(flag, rank, (size_t)n + 1, scan_tmp, scan_cap, s));
        hipLaunchKernelGGL(bcq_bfs_level_kernel, lg, blk, 0, s, arcs, lo, hi, 2, g->qperm, inv, best, flag, rank);
        ZG_TRY(hipGetLastError());
        uint32_t added = 0;
        ZG_TRY(hipMemcpyAsync(&added, rank + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        lo = hi;
        hi += added;
    }
    hipLaunchKernelGGL(bcq_rest_flag_kernel, grd, blk, 0, s, V, g->qperm, flag);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(zg_exclusive_scan(flag, rank, V1, s, sc));
    hipLaunchKernelGGL(bcq_rest_kernel, grd, blk, 0, s, V, flag, rank, hi, g->qperm, inv);
    ZG_TRY(hipGetLastError());
    // the CSR under the new ids, rows in their arc order
    for (int dir = 0; dir < 2; dir++) {
        const uint32_t *ptr = dir ? g->in_ptr : g->out_ptr, *col = dir ? g->in_col : g->out_col;
        uint32_t **nptr = dir ? &g->q_in_ptr : &g->q_out_ptr, **ncol = dir ? &g->q_in_col : &g->q_out_col;
        ZG_TRY_RC(zg_dmalloc(nptr, V1));
        ZG_TRY_RC(zg_dmalloc(ncol, E));
        hipLaunchKernelGGL(bcq_pdeg_kernel, grd, blk, 0, s, V, inv, ptr, deg);
        ZG_TRY(hipGetLastError());
        ZG_TRY_RC(zg_exclusive_scan(deg, *nptr, V1, s, sc));
        hipLaunchKernelGGL(bcq_prow_kernel, dim3(std::min<uint32_t>((V + 3) / 4, 16384)), blk, 0, s, V, inv, g->qperm,
                           ptr, col, *nptr, *ncol);
        ZG_TRY(hipGetLastError());
    }
    ZG_TRY(hipStreamSynchronize(s));
    return ZG_OK;
}

// Builds the order and the renumbered CSR, or leaves none of it behind.
static int bcq_build_order(zg_graph *g, hipStream_t s)
{
    const int rc = bcq_build_order_into(g, s);
    if (rc != ZG_OK) {
        (void)hipStreamSynchronize(s);
        zg_bcq_order_release(g);
    }
    return rc;
}

int zg_bcq_ws_release(zg_graph *g)
{
    zg_bcq_ws &w = g->bcq;
    zg_dfree(w.rec); zg_dfree(w.queue); zg_dfree(w.dlt); zg_dfree(w.lstart); zg_dfree(w.cpre);
    zg_dfree(w.qlen); zg_dfree(w.newcnt); zg_dfree(w.src); zg_dfree(w.wts); zg_dfree(w.bc);
    zg_dfree(w.stats);
    w.S = 0;
    w.lcap = 0;
    return ZG_OK;
}

static int bcq_ws_ensure(zg_graph *g, uint32_t S)
{
    zg_bcq_ws &w = g->bcq;
    if (w.S == S) return ZG_OK;
    zg_bcq_ws_release(g);
    const size_t V = g->V;
    w.lcap = 1024;
    int rc = ZG_OK;
    if (rc == ZG_OK) rc = zg_dmalloc(&w.rec, S * V);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.queue, S * V);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.dlt, S * V);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.lstart, (size_t)w.lcap * S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.cpre, (size_t)w.lcap * (S + 1));
    if (rc == ZG_OK) rc = zg_dmalloc(&w.qlen, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.newcnt, w.lcap);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.src, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.wts, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.bc, V);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.stats, 2);
    if (rc != ZG_OK) {
        zg_bcq_ws_release(g);
        return rc;
    }
    w.S = S;
    return ZG_OK;
}

// Grows the per-level tables to hold level `need` (prefix copied; level-major
// layout keeps the used prefix contiguous).
static int bcq_grow_levels(zg_graph *g, uint32_t need, uint32_t used, hipStream_t s)
{
    zg_bcq_ws &w = g->bcq;
    if (need < w.lcap) return ZG_OK;
    uint32_t nl = w.lcap;
    while (nl <= need) nl *= 2;
    uint32_t *ls = nullptr, *nc = nullptr, *cp = nullptr;
    ZG_TRY_RC(zg_dmalloc(&ls, (size_t)nl * w.S));
    int rc = zg_dmalloc(&nc, nl);
    if (rc == ZG_OK) rc = zg_dmalloc(&cp, (size_t)nl * (w.S + 1));
    if (rc != ZG_OK) {
        zg_dfree(ls);
        zg_dfree(nc);
        return rc;
    }
    const size_t keep = std::min(used, w.lcap);
    ZG_TRY(hipMemsetAsync(nc, 0, sizeof(uint32_t) * nl, s));
    ZG_TRY(hipMemcpyAsync(ls, w.lstart, sizeof(uint32_t) * keep * w.S, hipMemcpyDeviceToDevice, s));
    ZG_TRY(hipMemcpyAsync(cp, w.cpre, sizeof(uint32_t) * keep * (w.S + 1), hipMemcpyDeviceToDevice, s));
    ZG_TRY(hipMemcpyAsync(nc, w.newcnt, sizeof(uint32_t) * keep, hipMemcpyDeviceToDevice, s));
    ZG_TRY(hipStreamSynchronize(s));
    zg_dfree(w.lstart);
    zg_dfree(w.newcnt);
    zg_dfree(w.cpre);
    w.lstart = ls;
    w.newcnt = nc;
    w.cpre = cp;
    w.lcap = nl;
    return ZG_OK;
}

static BcqDev bcq_dev(zg_graph *g, bool weighted)
{
    zg_bcq_ws &w = g->bcq;
    BcqDev a;
    a.V = g->V;
    a.S = w.S;
    a.lcap = w.lcap;
    const bool ord = g->qperm != nullptr;  // locality order built (and on)
    a.out_ptr = ord ? g->q_out_ptr : g->out_ptr;
    a.out_col = ord ? g->q_out_col : g->out_col;
    a.in_ptr = ord ? g->q_in_ptr : g->in_ptr;
    a.in_col = ord ? g->q_in_col : g->in_col;
    a.adj = g->qadj;
    a.rec = w.rec;
    a.queue = w.queue;
    a.lstart = w.lstart;
    a.cpre = w.cpre;
    a.qlen = w.qlen;
    a.newcnt = w.newcnt;
    a.src = w.src;
    a.wts = weighted ? w.wts : nullptr;
    a.bc = w.bc;
    a.dlt = w.dlt;
    a.stats = w.stats;
    return a;
}

int zg_bcq_sweep(zg_graph *g, const uint32_t *d_sources, uint64_t n, const double *d_weights, double *h_totals,
                 double *d_out, hipStream_t s, zg_bc_info *inf)
{
    const uint32_t V = g->V;
    // sources per batch: as many as the call has (rounded up to 64), at most
    // ZG_Q_MAX_SOURCES, and as many as half the free device memory holds at
    // ZG_Q_BYTES_PER_SV bytes of state per source and vertex; an existing
    // larger workspace is kept
    const uint32_t want = (uint32_t)std::min<uint64_t>(ZG_Q_MAX_SOURCES, (n + 63) / 64 * 64);
    uint32_t S = g->bcq.S;
    if (S < want) {
        size_t fr = 0, tot = 0;
        ZG_TRY(hipMemGetInfo(&fr, &tot));
        fr += (size_t)g->bcq.S * V * ZG_Q_BYTES_PER_SV;  // released before the new one is made
        const uint64_t fit = (uint64_t)(0.5 * (double)fr) / ((uint64_t)ZG_Q_BYTES_PER_SV * (V ? V : 1));
        S = (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(want, fit) / 64 * 64);
    }
    ZG_TRY_RC(bcq_ws_ensure(g, S));
    zg_bcq_ws &w = g->bcq;
    if (g->queue_order && !g->qperm && V) ZG_TRY_RC(bcq_build_order(g, s));  // once per resource
    const bool ord = g->qperm != nullptr;
    if (!g->qadj && V) {  // the inline adjacency, once per resource (and order)
        ZG_TRY_RC(zg_dmalloc(&g->qadj, (size_t)V * 4));
        hipLaunchKernelGGL(bcq_adj_kernel, dim3(std::min<uint32_t>((V + 255) / 256, 8192)), dim3(256), 0, s, V,
                           ord ? g->q_in_ptr : g->in_ptr, ord ? g->q_in_col : g->in_col,
                           ord ? g->q_out_ptr : g->out_ptr, ord ? g->q_out_col : g->out_col, g->qadj);
        ZG_TRY(hipGetLastError());
    }
    ZG_TRY(hipMemsetAsync(w.bc, 0, sizeof(double) * (V ? V : 1), s));
    std::vector<uint32_t> newcnt;
    for (uint64_t b0 = 0; b0 < n; b0 += S) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(S, n - b0);
        ZG_TRY(hipMemsetAsync(w.src, 0xFF, sizeof(uint32_t) * S, s));
        ZG_TRY(hipMemcpyAsync(w.src, d_sources + b0, sizeof(uint32_t) * nb, hipMemcpyDeviceToDevice, s));
        if (ord) {
            hipLaunchKernelGGL(bcq_map_src_kernel, dim3((S + 255) / 256), dim3(256), 0, s, S, g->qperm, w.src);
            ZG_TRY(hipGetLastError());
        }
        if (d_weights)
            ZG_TRY(hipMemcpyAsync(w.wts, d_weights + b0, sizeof(double) * nb, hipMemcpyDeviceToDevice, s));
        ZG_TRY(hipMemsetAsync(w.rec, 0xFF, sizeof(zg_qrec) * (size_t)S * V, s));  // depth unset
        ZG_TRY(hipMemsetAsync(w.dlt, 0, sizeof(double) * (size_t)nb * V, s));      // unreached: no dependency
        ZG_TRY(hipMemsetAsync(w.newcnt, 0, sizeof(uint32_t) * w.lcap, s));
        ZG_TRY(hipMemsetAsync(w.stats, 0, sizeof(unsigned long long) * 2, s));
        BcqDev a = bcq_dev(g, d_weights != nullptr);
        ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bcq_init_kernel, dim3((S + 255) / 256), dim3(256), 0, s, a);
        // level 0's chunk prefix (its offsets are already set; re-set to the same values)
        ZG_PROF_LAUNCH(ZG_PROF_BC_FINALIZE, bcq_close_kernel, dim3(1), dim3(1024), 0, s, a, 0);
        ZG_TRY(hipGetLastError());
        // forward: queue levels in chunks, sync to find the last one
        uint32_t d = 0;
        int maxd = -1;
        while (maxd < 0) {
            ZG_TRY_RC(bcq_grow_levels(g, d + ZG_Q_LEVELS_PER_SYNC + 2, d + 2, s));
            a = bcq_dev(g, d_weights != nullptr);
            for (uint32_t k = 0; k < ZG_Q_LEVELS_PER_SYNC; k++) {
                ZG_PROF_LAUNCH(ZG_PROF_BC_PUSH, bcq_forward_kernel, dim3(ZG_Q_GRID), dim3(256), 0, s, a, (int)(d + k));
                ZG_PROF_LAUNCH(ZG_PROF_BC_FINALIZE, bcq_close_kernel, dim3(1), dim3(1024), 0, s, a, (int)(d + k + 1));
            }
            ZG_TRY(hipGetLastError());
            newcnt.resize(ZG_Q_LEVELS_PER_SYNC);
            ZG_TRY(hipMemcpyAsync(newcnt.data(), w.newcnt + d + 1, sizeof(uint32_t) * ZG_Q_LEVELS_PER_SYNC,
                                  hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            for (uint32_t k = 0; k < ZG_Q_LEVELS_PER_SYNC; k++)
                if (newcnt[k] == 0) {
                    maxd = (int)(d + k);  // level d+k+1 is empty
                    break;
                }
            d += ZG_Q_LEVELS_PER_SYNC;
            inf->push_levels += ZG_Q_LEVELS_PER_SYNC;
        }
        for (int L = maxd; L >= 1; L--)
            ZG_PROF_LAUNCH(ZG_PROF_BC_BACKWARD, bcq_backward_kernel, dim3(ZG_Q_GRID), dim3(256), 0, s, a, L);
        if (V)
            ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bcq_reduce_kernel, dim3((V + 255) / 256), dim3(256), 0, s, V, nb, w.dlt,
                           w.bc);
        ZG_TRY(hipGetLastError());
        unsigned long long st[2];
        ZG_TRY(hipMemcpyAsync(st, w.stats, sizeof(st), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        inf->arcs_traversed += st[0];
        inf->nodes_reached += st[1];
        inf->max_levels = std::max<uint32_t>(inf->max_levels, (uint32_t)maxd + 1);
        inf->batches++;
        if (h_totals) {
            // dependency total of a source: sum over its levels L of (L - 1) x level size
            std::vector<uint32_t> ls((size_t)(maxd + 2) * S);
            ZG_TRY(hipMemcpyAsync(ls.data(), w.lstart, sizeof(uint32_t) * ls.size(), hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            for (uint32_t i = 0; i < nb; i++) {
                double t = 0.0;
                for (int L = 2; L <= maxd; L++)
                    t += (double)(L - 1) * (double)(ls[(size_t)(L + 1) * S + i] - ls[(size_t)L * S + i]);
                h_totals[b0 + i] = t;
            }
        }
    }
    if (V) {
        ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bcq_out_kernel, dim3((V + 255) / 256), dim3(256), 0, s, V, w.bc, g->live,
                       g->qperm, d_out);
        ZG_TRY(hipGetLastError());
    }
    inf->n_sources = n;
    return ZG_OK;
}
