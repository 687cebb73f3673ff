// Unweighted Brandes betweenness as a batched, level-synchronous sweep.
//
// Reference: priv/zog/src/centrality.zig:694-798 (ArrayGraph branch).  Per
// source s: BFS where an arc v->w with dist[w] == dist[v] + 1 adds sigma[v]
// to sigma[w] (once per parallel arc; self-loops never qualify), then in
// reverse BFS order delta[u] += (sigma[u]/sigma[w]) * (1 + delta[w]) over the
// predecessors u of w, and bc[w] += delta[w] for w != s.
//
// gfx950 design.  Sources are processed 64 at a time as a *group*: lane l of
// every wavefront owns source slot l, and per-vertex state is stored as a
// 64-wide row, so one wavefront reads one vertex's state for all 64 sources
// with a single coalesced 512-byte access.  A batch sweeps up to
// ZG_MAX_GROUPS groups in the same launches (blockIdx.y = group).
//
// Forward (level d -> d+1), pull over predecessor chunks: for a vertex v with
// unvisited lanes U, every predecessor u whose frontier mask cur[u] meets U
// contributes sigma[u][l] for the lanes l in cur[u] & U.  The lanes load 64
// predecessors and their masks at once; a ballot picks the ones that
// contribute, and each contributes one coalesced row read.  New lanes are
// published in the next-frontier mask.  A finalize kernel (fixed grid, one
// contiguous chunk range per block) folds them into the visited mask and
// appends the vertex's successor chunks (with the lane mask) to the
// level-(d+1) work list, reserving space with one atomic per block.
//
// Backward (level d, from the deepest level up): for an entry (v, lanes at
// depth d), delta_l(v) = sigma_l(v) * sum_{w in succ(v)} c_l(w), where
// c_l(w) = (1 + delta_l(w)) / sigma_l(w) is published when w's level is done.
// Within the levels the lane table covers, c_l(w) is written over sigma_l(w)
// in place (sigma_l(w) is not read again once w's level is done): a parent
// at level d reads it only for the lanes the table puts at level d+1, so a
// same-level neighbour's fresh value is never seen.  Past the table (and
// when the counts are exported through a hook), c is double-buffered by
// level parity: level d reads buffer (d+1)&1 and writes buffer d&1, both
// cleared before the backward pass.  An arc v->w out of a reached v has
// depth(w) <= depth(v) + 1, so in buffer (d+1)&1 the only lanes of w already
// written are those with depth(w) = d+1: c_l(w) != 0 exactly when w is a BFS
// child of v for source l, with no depth compare.  No atomics for ordinary vertices; the
// per-lane bc contributions are summed across the wavefront before one
// read-modify-write per (vertex, level).
//
// Path counts are FP64 (exact integers below 2^53, equal to the reference's
// u64 counts there) with a per-(source, level) power-of-two shift k_L taken
// only when a level's largest exponent passes ZG_SHIFT_T, so counts on
// high-diameter graphs never overflow; every k is 0 until then.
#include "build_util.h"

#include <algorithm>
#include <cmath>

struct GroupRange {
    uint32_t off[ZG_MAX_GROUPS];
    uint32_t cnt[ZG_MAX_GROUPS];
};

// Vertex range [vlo, vhi) and predecessor-chunk range [clo, chi) a group
// works on: the components (graphs) its sources belong to.
struct GroupSpan {
    uint32_t vlo[ZG_MAX_GROUPS], vhi[ZG_MAX_GROUPS];
    uint32_t clo[ZG_MAX_GROUPS], chi[ZG_MAX_GROUPS];
    uint32_t used[ZG_MAX_GROUPS];   // source slots in use
};

struct BcDev {
    uint32_t V;
    uint32_t Vloc;              // per-group capacity: state of vertex v of group g sits at v - vlo[g]
    uint32_t vlo[ZG_MAX_GROUPS];
    uint32_t G;                 // groups in this batch (gridDim.y)
    uint32_t n_ich;
    const uint32_t *in_ptr, *in_col, *out_ptr, *out_col;
    const uint32_t *ich_v, *ich_e0, *ich_e1;
    const uint32_t *och_first, *och_v, *och_e0, *och_e1;
    const uint32_t *ohub_idx;
    const uint8_t *live;
    double *sig, *cc, *hubacc, *bcg;   // cc: [2][Gcap][V][64]
    uint32_t *dep;
    uint32_t *maxdep;           // [Gcap][V]
    uint64_t *lvl;              // [lmcap][Gcap][V] lanes whose level-L frontier held v
    uint32_t lmcap;
    // 1: a lane finished at a level below lmcap keeps its (1 + delta) / sigma
    // in place of its path count (nothing reads the count after that), and
    // its parents, which read only child lanes there, find it in sig; 0 (the
    // path-count export, which reads sig after the backward pass): every
    // level's terms go to the cc parity buffers
    uint32_t inplace;
    uint64_t *vis, *fr;         // fr: [3][Gcap][V]
    unsigned long long *stats;  // [Gcap][ZG_FIN_BLOCKS][ZG_BC_STATS]
    unsigned long long *tree;   // [Gcap][ZG_TREE_SLOTS] per-lane tree arcs seen by the backward pass, or null
    uint32_t *touched;          // [Gcap][V]
    uint32_t *flist;
    uint64_t *fmask;
    uint32_t *hlist;
    uint64_t *hmask;
    int *emax;                  // [L][Gcap][64]
    uint32_t *foff, *hoff;      // [L][Gcap] level offsets in flist / hlist
    uint32_t *modes;            // [L][Gcap]
    const uint32_t *src;        // [Gcap][64]
    const double *wts;          // [Gcap][64] source weights, or null (all 1)
    unsigned long long *dsum;   // [Gcap][64] dependency totals, or null (not wanted)
    // folded sweep (zg_bc_folded), or null: succ1[t] = the source t is folded
    // into (ZG_NONE if none), fcount[m] = sources folded into m; per lane,
    // vertices reached beyond the source (rsum) and folded members reached (fsum)
    const uint32_t *succ1, *fcount;
    unsigned long long *rsum, *fsum;    // [Gcap][64]
    double push_alpha;
    unsigned long long Eg[ZG_MAX_GROUPS];  // arcs inside each group's span
    zg_bc_ws::Counters *cnt;
    size_t fcap, hcap;
    uint32_t Gcap, n_ohub;
};

// Direction of a group's level.
#define ZG_MODE_IDLE 0u    // the level is empty: the group's sweep is done
#define ZG_MODE_PULL 1u
#define ZG_MODE_PUSH 2u

// Levels queued per host sync: the first chunk, doubling up to the cap.
#define ZG_LEVEL_CHUNK0 12
// Lane table of the backward pass: at most this many levels and bytes.
#define ZG_LVL_LEVELS 32
#define ZG_LVL_BYTES ((size_t)2 << 30)
#define ZG_LEVEL_CHUNK_MAX 64
// Fixed grid of the push kernel (it strides over the level's entries).
#define ZG_PUSH_GRID 1024

// Arguments of the graph-replayed level kernels, kept in device memory so one
// captured graph serves every chunk of levels of every batch: the sweep's
// BcDev, the batch's spans, and the device-side level counter.
struct BcLevelArgs {
    BcDev a;
    GroupSpan sp;
    int *cur;
};

// Per-group state is indexed by the vertex's position in the group's span.
#define MASK(g, v) ((size_t)(g) * a.Vloc + ((v) - a.vlo[g]))
#define ROW(g, v) (MASK(g, v) * 64)
#define EMAX(L, g, l) ((((size_t)(L)) * a.Gcap + (g)) * 64 + (l))
#define LVLM(L, g, v) ((size_t)(L) * a.Gcap * a.Vloc + MASK(g, v))

__device__ __forceinline__ uint64_t *fr_buf(const BcDev &a, int k)
{
    return a.fr + (size_t)k * a.Gcap * a.Vloc;
}

__device__ __forceinline__ double *cc_buf(const BcDev &a, int k)
{
    return a.cc + (size_t)k * a.Gcap * a.Vloc * 64;
}

// Batch reset over a group's vertex span: the visited mask starts as the
// group's unused lanes, the frontier buffers and the first nlev levels of
// the lane table as empty.
__global__ void bc_reset_kernel(BcDev a, GroupSpan sp, uint32_t nlev)
{
    const uint32_t g = blockIdx.y;
    const uint32_t v = sp.vlo[g] + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= sp.vhi[g]) return;
    const uint32_t nl = sp.used[g];
    const uint64_t used = nl >= 64 ? ~0ull : ((1ull << nl) - 1);
    a.vis[MASK(g, v)] = ~used;
    fr_buf(a, 0)[MASK(g, v)] = 0;
    fr_buf(a, 1)[MASK(g, v)] = 0;
    fr_buf(a, 2)[MASK(g, v)] = 0;
    for (uint32_t L = 0; L < nlev; L++) a.lvl[LVLM(L, g, v)] = 0;
}

// Lane-table rows [from[g], to[g]) of group g (positions past this batch's
// span that an earlier batch's larger span wrote) cleared for the first nlev
// levels.
struct LvlExtent {
    uint32_t from[ZG_MAX_GROUPS], to[ZG_MAX_GROUPS];
};
__global__ void bc_lvl_clear_kernel(BcDev a, LvlExtent x, uint32_t nlev)
{
    const uint32_t g = blockIdx.y;
    const uint32_t p = x.from[g] + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= x.to[g]) return;
    for (uint32_t L = 0; L < nlev; L++) a.lvl[(size_t)L * a.Gcap * a.Vloc + (size_t)g * a.Vloc + p] = 0;
}

// Path-count rows of a group's span start at zero (push levels and
// multi-chunk vertices add into them); exported depths start unreached.
__global__ void bc_zero_rows_kernel(BcDev a, GroupSpan sp)
{
    const uint32_t g = blockIdx.y;
    const uint32_t v = sp.vlo[g] + blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= sp.vhi[g]) return;
    const size_t r = ROW(g, v) + (threadIdx.x & 63);
    a.sig[r] = 0.0;
    if (a.dep) a.dep[r] = ZG_NONE;
}

// Level 0: every lane's source gets depth 0, sigma 1 and joins frontier 0.
__global__ void bc_source_kernel(BcDev a)
{
    const uint32_t g = blockIdx.y, l = threadIdx.x;
    uint32_t s = a.src[g * 64 + l];
    unsigned long long arcs = 0, reached = 0;
    if (s != ZG_NONE) {
        if (s >= a.V || !a.live[s]) {
            atomicOr(&a.cnt[g].err, ZG_BC_ERR_SOURCE);
        } else {
            const uint64_t bit = 1ull << l;
            atomicOr((unsigned long long *)&a.vis[MASK(g, s)], (unsigned long long)bit);
            atomicOr((unsigned long long *)&fr_buf(a, 0)[MASK(g, s)], (unsigned long long)bit);
            a.sig[ROW(g, s) + l] = 1.0;
            if (a.dep) a.dep[ROW(g, s) + l] = 0;
            a.maxdep[MASK(g, s)] = 0;
            uint32_t c0 = a.och_first[s], k = a.och_first[s + 1] - c0;
            uint32_t pos = atomicAdd(&a.cnt[g].fcount, k);
            for (uint32_t q = 0; q < k; q++) {
                if (pos + q < a.fcap) {
                    a.flist[g * a.fcap + pos + q] = c0 + q;
                    a.fmask[g * a.fcap + pos + q] = bit;
                } else {
                    atomicOr(&a.cnt[g].err, ZG_BC_ERR_CAPACITY);
                }
            }
            arcs = a.out_ptr[s + 1] - a.out_ptr[s];
            reached = 1;
        }
    }
    arcs = zg_wave_sum_u64(arcs);
    reached = zg_wave_sum_u64(reached);
    if (l == 0) {
        a.cnt[g].farcs[0] = arcs;
        // slot 0 belongs to finalize block 0, which runs strictly later
        a.stats[(size_t)g * ZG_FIN_BLOCKS * ZG_BC_STATS] += arcs;
        a.stats[(size_t)g * ZG_FIN_BLOCKS * ZG_BC_STATS + 1] += reached;
    }
}

// Opens level d for every group: records where the level's work lists end
// (everything finalize(d-1) appended), picks the direction from the level's
// successor-arc count, and clears the counters level d accumulates into.
__device__ __forceinline__ void bc_level_body(const BcDev &a, int d)
{
    const uint32_t g = threadIdx.x;
    if (g >= a.G) return;
    zg_bc_ws::Counters &c = a.cnt[g];
    const uint32_t fc = c.fcount;
    a.foff[(size_t)(d + 1) * a.Gcap + g] = fc;
    a.hoff[(size_t)(d + 1) * a.Gcap + g] = c.hcount;
    const uint32_t start = a.foff[(size_t)d * a.Gcap + g];
    const unsigned long long fa = c.farcs[d & 1];
    uint32_t mode = ZG_MODE_IDLE;
    if (fc > start)
        mode = (a.push_alpha > 0.0 && (double)fa * a.push_alpha < (double)a.Eg[g]) ? ZG_MODE_PUSH : ZG_MODE_PULL;
    c.mode = mode;
    a.modes[(size_t)d * a.Gcap + g] = mode;
    c.tcount = 0;
    c.farcs[(d + 1) & 1] = 0;
}

__global__ void bc_level_kernel(BcDev a, int d) { bc_level_body(a, d); }

// Graph-replayed form: the level index lives on the device and advances here.
__global__ void bc_level_kernel_g(const BcLevelArgs *__restrict__ p)
{
    const int d = *p->cur + 1;
    __syncthreads();
    if (threadIdx.x == 0) *p->cur = d;
    bc_level_body(p->a, d);
}

// One predecessor chunk of the pull at level d (discovers level d+1); k is
// the lane's level-d exponent shift.  Everything that needs only the chunk
// is loaded in one round -- v's visited lanes, its in-degree (a vertex of
// several chunks adds atomically) and the chunk's first 64 predecessors (a
// lane past the end re-reads the last one; predecessor chunks are never
// empty) -- so a chunk costs four dependent round trips: chunk, vertex,
// predecessor masks, path-count rows.
__device__ __forceinline__ void bc_pull_chunk(const BcDev &a, uint32_t g, int lane, int d, uint32_t ch, int k)
{
    const uint32_t v = a.ich_v[ch], e0 = a.ich_e0[ch], e1 = a.ich_e1[ch];
    const uint32_t e_first = e0 + lane;
    const uint32_t u_first = a.in_col[e_first < e1 ? e_first : e1 - 1];
    const uint64_t unvis = ~a.vis[MASK(g, v)];
    const uint32_t ideg = a.in_ptr[v + 1] - a.in_ptr[v];
    // (no vertex id and no in-degree is ZG_NONE: the test is unvis == 0, and
    // naming the other two keeps their loads ahead of the branch, in the
    // same round as v's visited lanes)
    if (unvis == 0 && u_first != ZG_NONE && ideg != ZG_NONE) return;
    const uint64_t *cur = fr_buf(a, d % 3);
    double acc = 0.0;
    uint64_t m = 0;
    for (uint32_t base = e0; base < e1; base += 64) {
        const uint32_t e = base + lane;
        const uint32_t u = base == e0 ? u_first : a.in_col[e < e1 ? e : e1 - 1];
        uint64_t f = cur[MASK(g, u)] & unvis;
        if (e >= e1) f = 0;
        uint64_t act = __ballot(f != 0);
        // contributing predecessors, eight row reads in flight per round
        while (act) {
            uint32_t uj[8];
            uint64_t fj[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (act) {
                    const int j = __builtin_ctzll(act);
                    act &= act - 1;
                    uj[q] = zg_readlane32(u, j);
                    fj[q] = zg_readlane64(f, j);
                } else {
                    uj[q] = 0;
                    fj[q] = 0;
                }
                m |= fj[q];
            }
            double val[8];
#pragma unroll
            for (int q = 0; q < 8; q++) val[q] = ((fj[q] >> lane) & 1) ? a.sig[ROW(g, uj[q]) + lane] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; q++) acc += val[q];
        }
    }
    if (m == 0) return;
    const bool mine = (m >> lane) & 1;
    const bool multi = ideg > ZG_CHUNK;
    const size_t r = ROW(g, v) + lane;
    if (mine) {
        const double val = ldexp(acc, -k);
        if (multi) {
            atomicAdd(&a.sig[r], val);
        } else {
            a.sig[r] = val;
            const int ex = ilogb(val);
            if (ex > ZG_SHIFT_T) atomicMax(&a.emax[EMAX(d + 1, g, lane)], ex);
        }
        if (a.dep) a.dep[r] = (uint32_t)(d + 1);
    }
    if (lane == 0) {
        uint64_t *nxt = fr_buf(a, (d + 1) % 3) + MASK(g, v);
        if (multi) atomicOr((unsigned long long *)nxt, (unsigned long long)m);
        else *nxt = m;
    }
}


// Forward level d: discover level d+1 by pulling over predecessor chunks (a
// fixed grid striding over the group's chunks).
__device__ __forceinline__ void bc_forward_body(const BcDev &a, int d, const GroupSpan &sp)
{
    const uint32_t g = blockIdx.y;
    if (a.cnt[g].mode != ZG_MODE_PULL) return;
    const int lane = threadIdx.x & 63;
    const uint32_t nch = sp.chi[g] - sp.clo[g];
    const int k = a.emax[EMAX(d, g, lane)];  // written for level d+1 only during this level
    for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < nch; i += gridDim.x * 4)
        bc_pull_chunk(a, g, lane, d, sp.clo[g] + i, k);
}

__global__ __launch_bounds__(256) void bc_forward_kernel(BcDev a, int d, GroupSpan sp) { bc_forward_body(a, d, sp); }

__global__ __launch_bounds__(256) void bc_forward_kernel_g(const BcLevelArgs *__restrict__ p)
{
    bc_forward_body(p->a, *p->cur, p->sp);
}

__device__ __forceinline__ void push_flush_touched(const BcDev &a, uint32_t g, int lane, uint32_t tbuf, int tn)
{
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&a.cnt[g].tcount, (uint32_t)tn);
    base = zg_readlane32(base, 0);
    if (lane < tn) {
        if (base + lane < a.Vloc) a.touched[(size_t)g * a.Vloc + base + lane] = tbuf;
        else atomicOr(&a.cnt[g].err, ZG_BC_ERR_CAPACITY);
    }
}

// Forward level d, push over the level-d work list (sparse frontiers): each
// entry's lanes add their count into every successor not yet visited by that
// lane (FP64 atomics; exact for integer counts below 2^53).
__device__ __forceinline__ void bc_push_body(const BcDev &a, int d)
{
    const uint32_t g = blockIdx.y;
    if (a.cnt[g].mode != ZG_MODE_PUSH) return;
    const int lane = threadIdx.x & 63;
    const uint32_t off = a.foff[(size_t)d * a.Gcap + g];
    const uint32_t cnt = a.foff[(size_t)(d + 1) * a.Gcap + g] - off;
    uint64_t *next = fr_buf(a, (d + 1) % 3);
    uint32_t tbuf = 0;  // lane k holds the k-th vertex this wavefront reached first
    int tn = 0;
    for (uint32_t wid = blockIdx.x * 4 + (threadIdx.x >> 6); wid < cnt; wid += gridDim.x * 4) {
    const size_t idx = g * a.fcap + off + wid;
    const uint32_t oc = a.flist[idx];
    const uint64_t mask = a.fmask[idx];
    const uint32_t v = a.och_v[oc], e0 = a.och_e0[oc], e1 = a.och_e1[oc];
    const bool mine = (mask >> lane) & 1;
    const double val = mine ? ldexp(a.sig[ROW(g, v) + lane], -a.emax[EMAX(d, g, lane)]) : 0.0;
    for (uint32_t base = e0; base < e1; base += 64) {
        const uint32_t e = base + lane;
        uint32_t w = 0;
        uint64_t m = 0;
        if (e < e1) {
            w = a.out_col[e];
            m = mask & ~a.vis[MASK(g, w)];
        }
        uint64_t act = __ballot(m != 0);
        while (act) {
            const int j = __builtin_ctzll(act);
            act &= act - 1;
            const uint32_t wj = zg_readlane32(w, j);
            const uint64_t mj = zg_readlane64(m, j);
            if ((mj >> lane) & 1) {
                const size_t r = ROW(g, wj) + lane;
                atomicAdd(&a.sig[r], val);
                if (a.dep) a.dep[r] = (uint32_t)(d + 1);
            }
            unsigned long long old = 0;
            if (lane == 0) old = atomicOr((unsigned long long *)&next[MASK(g, wj)], (unsigned long long)mj);
            if (zg_readlane64(old, 0) == 0) {
                // first to reach wj at this level: record it for the sparse finalize
                if (lane == tn) tbuf = wj;
                if (++tn == 64) {
                    push_flush_touched(a, g, lane, tbuf, tn);
                    tn = 0;
                }
            }
        }
    }
    }
    if (tn) push_flush_touched(a, g, lane, tbuf, tn);
}

__global__ __launch_bounds__(256) void bc_push_kernel(BcDev a, int d) { bc_push_body(a, d); }

__global__ __launch_bounds__(256) void bc_push_kernel_g(const BcLevelArgs *__restrict__ p) { bc_push_body(p->a, *p->cur); }

// Per-lane counts of the vertices a finalize block adds to level d+1; each
// adds d to its lane's dependency total sum_t (dist(s,t) - 1).
__device__ __forceinline__ void bc_count_lanes(uint32_t *lcnt, uint64_t nm)
{
    while (nm) {
        const int l = __builtin_ctzll(nm);
        nm &= nm - 1;
        atomicAdd(&lcnt[l], 1u);
    }
}

__device__ __forceinline__ void bc_flush_lanes(const BcDev &a, uint32_t g, int d, const uint32_t *lcnt)
{
    const uint32_t t = threadIdx.x;
    if (t < 64 && lcnt[t] && d > 0 && a.dsum)
        atomicAdd(&a.dsum[g * 64 + t], (unsigned long long)lcnt[t] * (unsigned long long)d);
    if (t < 64 && lcnt[t] && a.rsum) atomicAdd(&a.rsum[g * 64 + t], (unsigned long long)lcnt[t]);
}

// Folded sweep: lanes of nm whose source is the vertex v is folded into
// (v reached by its own group's representative) count one more folded
// member reached.
__device__ __forceinline__ void bc_count_folded(const BcDev &a, uint32_t g, uint32_t v, uint64_t nm, uint32_t *fcnt)
{
    const uint32_t m = a.succ1[v];
    if (m == ZG_NONE) return;
    while (nm) {
        const int l = __builtin_ctzll(nm);
        nm &= nm - 1;
        if (a.src[g * 64 + l] == m) atomicAdd(&fcnt[l], 1u);
    }
}

__device__ __forceinline__ void bc_flush_folded(const BcDev &a, uint32_t g, const uint32_t *fcnt)
{
    const uint32_t t = threadIdx.x;
    if (t < 64 && fcnt[t]) atomicAdd(&a.fsum[g * 64 + t], (unsigned long long)fcnt[t]);
}

// Clears the frontier buffer of level d-1 (it becomes level d+2's) for the
// vertices of level d-1's work list.
__device__ __forceinline__ void bc_clear_stale(const BcDev &a, int d, uint32_t g, uint32_t i0, uint32_t step)
{
    if (d < 1) return;
    const uint32_t off = a.foff[(size_t)(d - 1) * a.Gcap + g];
    const uint32_t n = a.foff[(size_t)d * a.Gcap + g] - off;
    for (uint32_t i = i0; i < n; i += step) {
        const uint32_t oc = a.flist[g * a.fcap + off + i];
        fr_buf(a, (d + 2) % 3)[MASK(g, a.och_v[oc])] = 0;
    }
}

// Exclusive prefix sum over a 1024-thread block (16 wavefronts); returns the
// block total through *total.  All threads must call it.
__device__ __forceinline__ uint32_t block_excl_scan_1024(uint32_t x, uint32_t *lds, uint32_t *total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t incl = zg_wave_incl_scan(x, lane);
    if (lane == 63) lds[wv] = incl;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t t = threadIdx.x < 16 ? lds[threadIdx.x] : 0;
        uint32_t ti = zg_wave_incl_scan(t, lane);
        if (threadIdx.x < 16) lds[16 + threadIdx.x] = ti - t;  // exclusive wave offsets
        if (threadIdx.x == 15) lds[32] = ti;
    }
    __syncthreads();
    const uint32_t r = lds[16 + wv] + incl - x;
    *total = lds[32];
    __syncthreads();
    return r;
}

// Finalize level d: fold level d+1 into the visited masks, append its work
// lists, clear the frontier buffer of level d-1.  Block b owns predecessor
// chunks [b*R, (b+1)*R); a vertex is handled by the block holding its first
// chunk.
__device__ __forceinline__ void bc_finalize_body(const BcDev &a, int d, int pushed, const GroupSpan &sp)
{
    __shared__ uint32_t lds[40];
    __shared__ uint32_t base[2];
    __shared__ unsigned long long sred[4][ZG_FIN_THREADS / 64];
    const uint32_t g = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
    if (a.cnt[g].mode != ZG_MODE_PULL) return;
    const int lane = tid & 63;
    bc_clear_stale(a, d, g, b * ZG_FIN_THREADS + tid, gridDim.x * ZG_FIN_THREADS);
    const uint64_t *next = fr_buf(a, (d + 1) % 3);
    __shared__ uint32_t lcnt[64], fcnt[64];
    const bool lanes = a.dsum || a.succ1;
    if (lanes) {
        if (tid < 64) lcnt[tid] = fcnt[tid] = 0;
        __syncthreads();
    }
    const uint32_t nch = sp.chi[g] - sp.clo[g];
    const uint32_t R = (nch + gridDim.x - 1) / gridDim.x;
    const uint32_t c0 = sp.clo[g] + min(b * R, nch), c1 = sp.clo[g] + min(b * R + R, nch);

    // pass A: fold masks, count list entries, statistics
    uint32_t nf = 0, nh = 0;
    unsigned long long arcs = 0, reached = 0, farcs = 0, visits = 0;
    for (uint32_t c = c0 + tid; c < c1; c += ZG_FIN_THREADS) {
        const uint32_t v = a.ich_v[c];
        if (a.ich_e0[c] != a.in_ptr[v]) continue;
        const uint64_t nm = next[MASK(g, v)];
        if (!nm) continue;
        a.vis[MASK(g, v)] |= nm;
        a.maxdep[MASK(g, v)] = (uint32_t)(d + 1);  // levels only grow
        if ((uint32_t)(d + 1) < a.lmcap) a.lvl[LVLM(d + 1, g, v)] = nm;
        if (lanes) bc_count_lanes(lcnt, nm);
        if (a.succ1) bc_count_folded(a, g, v, nm, fcnt);
        const uint32_t kf = a.och_first[v + 1] - a.och_first[v];
        nf += kf;
        nh += kf > 1 ? 1 : 0;
        const unsigned long long pc = __popcll(nm);
        const uint32_t od = a.out_ptr[v + 1] - a.out_ptr[v];
        arcs += pc * od;
        reached += pc;
        farcs += od;
        visits++;
        if (pushed || a.in_ptr[v + 1] - a.in_ptr[v] > ZG_CHUNK) {
            // summed by atomics: the level-(d+1) exponent is known only now
            uint64_t mm = nm;
            while (mm) {
                const int l = __builtin_ctzll(mm);
                mm &= mm - 1;
                const int ex = ilogb(a.sig[ROW(g, v) + l]);
                if (ex > ZG_SHIFT_T) atomicMax(&a.emax[EMAX(d + 1, g, l)], ex);
            }
        }
    }
    uint32_t tf, th;
    (void)block_excl_scan_1024(nf, lds, &tf);
    (void)block_excl_scan_1024(nh, lds, &th);
    if (tid == 0) {
        base[0] = tf ? atomicAdd(&a.cnt[g].fcount, tf) : 0;
        base[1] = th ? atomicAdd(&a.cnt[g].hcount, th) : 0;
    }
    arcs = zg_wave_sum_u64(arcs);
    reached = zg_wave_sum_u64(reached);
    farcs = zg_wave_sum_u64(farcs);
    visits = zg_wave_sum_u64(visits);
    if (lane == 0) {
        sred[0][tid >> 6] = arcs;
        sred[1][tid >> 6] = reached;
        sred[2][tid >> 6] = farcs;
        sred[3][tid >> 6] = visits;
    }
    __syncthreads();
    if (lanes) bc_flush_lanes(a, g, d, lcnt);
    if (a.succ1) bc_flush_folded(a, g, fcnt);
    if (tid == 0) {
        unsigned long long A = 0, Rr = 0, F = 0, Vi = 0;
        for (int w = 0; w < ZG_FIN_THREADS / 64; w++) {
            A += sred[0][w];
            Rr += sred[1][w];
            F += sred[2][w];
            Vi += sred[3][w];
        }
        unsigned long long *st = a.stats + ((size_t)g * ZG_FIN_BLOCKS + b) * ZG_BC_STATS;
        st[0] += A;
        st[1] += Rr;
        st[2] += Vi;
        st[3] += F;
        if (F) atomicAdd(&a.cnt[g].farcs[(d + 1) & 1], F);
    }
    if (tf == 0 && th == 0) return;

    // pass B: write the entries, tile by tile
    uint32_t fpos = base[0], hpos = base[1];
    for (uint32_t t0 = c0; t0 < c1; t0 += ZG_FIN_THREADS) {
        const uint32_t c = t0 + tid;
        uint32_t v = 0, kf = 0, kh = 0;
        uint64_t nm = 0;
        if (c < c1) {
            v = a.ich_v[c];
            if (a.ich_e0[c] == a.in_ptr[v]) {
                nm = next[MASK(g, v)];
                if (nm) {
                    kf = a.och_first[v + 1] - a.och_first[v];
                    kh = kf > 1 ? 1 : 0;
                }
            }
        }
        uint32_t sf_tot, sh_tot;
        const uint32_t of = block_excl_scan_1024(kf, lds, &sf_tot);
        const uint32_t oh = block_excl_scan_1024(kh, lds, &sh_tot);
        if (kf) {
            const uint32_t pos = fpos + of, q0 = a.och_first[v];
            for (uint32_t q = 0; q < kf; q++) {
                if (pos + q < a.fcap) {
                    a.flist[g * a.fcap + pos + q] = q0 + q;
                    a.fmask[g * a.fcap + pos + q] = nm;
                } else {
                    atomicOr(&a.cnt[g].err, ZG_BC_ERR_CAPACITY);
                }
            }
        }
        if (kh) {
            const uint32_t pos = hpos + oh;
            if (pos < a.hcap) {
                a.hlist[g * a.hcap + pos] = v;
                a.hmask[g * a.hcap + pos] = nm;
            } else {
                atomicOr(&a.cnt[g].err, ZG_BC_ERR_CAPACITY);
            }
        }
        fpos += sf_tot;
        hpos += sh_tot;
    }
}

__global__ __launch_bounds__(ZG_FIN_THREADS) void bc_finalize_kernel(BcDev a, int d, int pushed, GroupSpan sp)
{
    bc_finalize_body(a, d, pushed, sp);
}

__global__ __launch_bounds__(ZG_FIN_THREADS) void bc_finalize_kernel_g(const BcLevelArgs *__restrict__ p)
{
    bc_finalize_body(p->a, *p->cur, 0, p->sp);
}

// Finalize level d after a push level: same as bc_finalize_kernel, over the
// vertices the push reached first (a fixed grid striding over the list).
#define ZG_SPARSE_FIN_BLOCKS 64
#ifndef ZG_FIN_CHUNKS
#define ZG_FIN_CHUNKS 4096        // predecessor chunks per dense-finalize workgroup (at most ZG_FIN_BLOCKS groups)
#endif
#ifndef ZG_SPARSE_FIN_VERTS
#define ZG_SPARSE_FIN_VERTS 8192  // span vertices per sparse-finalize workgroup
#endif
__device__ __forceinline__ void bc_finalize_sparse_body(const BcDev &a, int d)
{
    __shared__ uint32_t lds[40];
    __shared__ uint32_t base[2];
    __shared__ unsigned long long sred[4][ZG_FIN_THREADS / 64];
    const uint32_t g = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
    if (a.cnt[g].mode != ZG_MODE_PUSH) return;
    const int lane = tid & 63;
    bc_clear_stale(a, d, g, b * ZG_FIN_THREADS + tid, gridDim.x * ZG_FIN_THREADS);
    const uint64_t *next = fr_buf(a, (d + 1) % 3);
    const uint32_t *tl = a.touched + (size_t)g * a.Vloc;
    const uint32_t nt = min(a.cnt[g].tcount, a.Vloc);
    __shared__ uint32_t lcnt[64], fcnt[64];
    const bool lanes = a.dsum || a.succ1;
    if (lanes) {
        if (tid < 64) lcnt[tid] = fcnt[tid] = 0;
        __syncthreads();
    }
    const uint32_t R = (nt + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = min(b * R, nt), t1 = min(t0 + R, nt);

    uint32_t nf = 0, nh = 0;
    unsigned long long arcs = 0, reached = 0, farcs = 0, visits = 0;
    for (uint32_t t = t0 + tid; t < t1; t += ZG_FIN_THREADS) {
        const uint32_t v = tl[t];
        const uint64_t nm = next[MASK(g, v)];
        a.vis[MASK(g, v)] |= nm;
        a.maxdep[MASK(g, v)] = (uint32_t)(d + 1);
        if ((uint32_t)(d + 1) < a.lmcap) a.lvl[LVLM(d + 1, g, v)] = nm;
        if (lanes) bc_count_lanes(lcnt, nm);
        if (a.succ1) bc_count_folded(a, g, v, nm, fcnt);
        const uint32_t kf = a.och_first[v + 1] - a.och_first[v];
        nf += kf;
        nh += kf > 1 ? 1 : 0;
        const unsigned long long pc = __popcll(nm);
        const uint32_t od = a.out_ptr[v + 1] - a.out_ptr[v];
        arcs += pc * od;
        reached += pc;
        farcs += od;
        visits++;
        uint64_t mm = nm;
        while (mm) {
            const int l = __builtin_ctzll(mm);
            mm &= mm - 1;
            const int ex = ilogb(a.sig[ROW(g, v) + l]);
            if (ex > ZG_SHIFT_T) atomicMax(&a.emax[EMAX(d + 1, g, l)], ex);
        }
    }
    uint32_t tf, th;
    (void)block_excl_scan_1024(nf, lds, &tf);
    (void)block_excl_scan_1024(nh, lds, &th);
    if (tid == 0) {
        base[0] = tf ? atomicAdd(&a.cnt[g].fcount, tf) : 0;
        base[1] = th ? atomicAdd(&a.cnt[g].hcount, th) : 0;
    }
    arcs = zg_wave_sum_u64(arcs);
    reached = zg_wave_sum_u64(reached);
    farcs = zg_wave_sum_u64(farcs);
    visits = zg_wave_sum_u64(visits);
    if (lane == 0) {
        sred[0][tid >> 6] = arcs;
        sred[1][tid >> 6] = reached;
        sred[2][tid >> 6] = farcs;
        sred[3][tid >> 6] = visits;
    }
    __syncthreads();
    if (lanes) bc_flush_lanes(a, g, d, lcnt);
    if (a.succ1) bc_flush_folded(a, g, fcnt);
    if (tid == 0) {
        unsigned long long A = 0, Rr = 0, F = 0, Vi = 0;
        for (int w = 0; w < ZG_FIN_THREADS / 64; w++) {
            A += sred[0][w];
            Rr += sred[1][w];
            F += sred[2][w];
            Vi += sred[3][w];
        }
        unsigned long long *st = a.stats + ((size_t)g * ZG_FIN_BLOCKS + b) * ZG_BC_STATS;
        st[0] += A;
        st[1] += Rr;
        st[2] += Vi;
        st[3] += F;
        if (F) atomicAdd(&a.cnt[g].farcs[(d + 1) & 1], F);
    }
    if (tf == 0 && th == 0) return;
    uint32_t fpos = base[0], hpos = base[1];
    for (uint32_t tt = t0; tt < t1; tt += ZG_FIN_THREADS) {
        const uint32_t t = tt + tid;
        uint32_t v = 0, kf = 0, kh = 0;
        uint64_t nm = 0;
        if (t < t1) {
            v = tl[t];
            nm = next[MASK(g, v)];
            kf = a.och_first[v + 1] - a.och_first[v];
            kh = kf > 1 ? 1 : 0;
        }
        uint32_t sf_tot, sh_tot;
        const uint32_t of = block_excl_scan_1024(kf, lds, &sf_tot);
        const uint32_t oh = block_excl_scan_1024(kh, lds, &sh_tot);
        if (kf) {
            const uint32_t pos = fpos + of, q0 = a.och_first[v];
            for (uint32_t q = 0; q < kf; q++) {
                if (pos + q < a.fcap) {
                    a.flist[g * a.fcap + pos + q] = q0 + q;
                    a.fmask[g * a.fcap + pos + q] = nm;
                } else {
                    atomicOr(&a.cnt[g].err, ZG_BC_ERR_CAPACITY);
                }
            }
        }
        if (kh) {
            const uint32_t pos = hpos + oh;
            if (pos < a.hcap) {
                a.hlist[g * a.hcap + pos] = v;
                a.hmask[g * a.hcap + pos] = nm;
            } else {
                atomicOr(&a.cnt[g].err, ZG_BC_ERR_CAPACITY);
            }
        }
        fpos += sf_tot;
        hpos += sh_tot;
    }
}

__global__ __launch_bounds__(ZG_FIN_THREADS) void bc_finalize_sparse_kernel(BcDev a, int d)
{
    bc_finalize_sparse_body(a, d);
}

__global__ __launch_bounds__(ZG_FIN_THREADS) void bc_finalize_sparse_kernel_g(const BcLevelArgs *__restrict__ p)
{
    bc_finalize_sparse_body(p->a, *p->cur);
}

// What the dependency step of vertex v in group g needs besides its
// successors' sum, per lane: the path count, the level's exponent shift,
// the lane's source, its weight and (folded sweep) its target weight, and
// the group accumulator of v (read by lane 0 only).  Loaded in one round.
struct BcFinishIn {
    double sv, wl, bc_old;
    int k;
    uint32_t srcl;
    double tau;
};

__device__ __forceinline__ BcFinishIn bc_finish_load(const BcDev &a, uint32_t g, uint32_t v, int lane, int d,
                                                     bool active)
{
    BcFinishIn in;
    in.sv = active ? a.sig[ROW(g, v) + lane] : 1.0;  // only the active lanes' lines of the row
    in.k = a.emax[EMAX(d, g, lane)];
    in.srcl = a.src[g * 64 + lane];
    in.wl = a.wts ? a.wts[g * 64 + lane] : 1.0;
    in.bc_old = a.bcg[(size_t)g * a.V + v];
    // target weight: 1, or in a folded sweep 1 + the folded members of the
    // lane's source, less one when v is itself one of them (lanes without a
    // source read slot 0 and are inactive)
    in.tau = 1.0;
    if (a.succ1) {
        const uint32_t m = in.srcl != ZG_NONE ? in.srcl : 0u;
        in.tau += (double)a.fcount[m] - (a.succ1[v] == m ? 1.0 : 0.0);
    }
    return in;
}

// delta / c / bc for lanes `active` of vertex v at level d, given S = sum of
// c over v's successors.
__device__ __forceinline__ void bc_finish(const BcDev &a, uint32_t g, uint32_t v, int lane, bool active,
                                          double S, int d, const BcFinishIn &in)
{
    double contrib = 0.0;
    if (active) {
        const size_t r = ROW(g, v) + lane;
        const double delta = in.sv * ldexp(S, -in.k);
        const double cv = (in.tau + delta) / in.sv;
        if (a.inplace && (uint32_t)d < a.lmcap) a.sig[r] = cv;
        else cc_buf(a, d & 1)[r] = cv;
        if (in.srcl != v) contrib = a.wts ? delta * in.wl : delta;
    }
    contrib = zg_wave_sum(contrib);
    // v's entry of the group accumulator has one writer per level (this
    // wavefront), and earlier levels finished before this launch
    if (lane == 0) a.bcg[(size_t)g * a.V + v] = in.bc_old + contrib;
}

// Backward level d over the level's successor-chunk entries.  Round trips:
// the entry; its chunk; everything that needs only v (first successors,
// hub slot, the finish inputs); the successors' level masks; their rows.
__global__ __launch_bounds__(256) void bc_backward_kernel(BcDev a, int d, GroupRange R)
{
    const uint32_t g = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const uint32_t wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= R.cnt[g]) return;
    const size_t idx = g * a.fcap + R.off[g] + wid;
    const uint32_t oc = a.flist[idx];
    const uint64_t mask = a.fmask[idx];
    const uint32_t v = a.och_v[oc], e0 = a.och_e0[oc], e1 = a.och_e1[oc];
    const bool active = (mask >> lane) & 1;
    // children finished at level d+1: in place below the table's end, else
    // in the parity buffer (zero in every lane not written at a level of
    // that parity: cleared before the backward pass of a batch that uses it)
    const bool exact_rows = a.inplace && (uint32_t)(d + 1) < a.lmcap;
    const double *cg = exact_rows ? a.sig : cc_buf(a, (d + 1) & 1);
    // (a sink's chunk is empty: no successor is read)
    uint32_t w_first = 0;
    if (e1 > e0) w_first = a.out_col[e0 + lane < e1 ? e0 + (uint32_t)lane : e1 - 1];
    const uint32_t hub = a.ohub_idx[v];
    const uint32_t first_ch = a.och_first[v];
    const BcFinishIn fin = bc_finish_load(a, g, v, lane, d, active);
    // A successor w is a BFS child of v in lane l iff l's level-(d+1)
    // frontier held w.  With the level's lane table, only child lanes read,
    // and they read c from w's sigma row, where it replaced sigma (when the
    // terms are kept in place).  Past the table, every active lane reads
    // the parity buffer row of successors that some lane reached at depth
    // >= d+1; those buffers were cleared before the backward pass, so a
    // lane for which w is no child reads zero.
    const bool exact = (uint32_t)(d + 1) < a.lmcap;
    double S = 0.0;
    uint32_t tree = 0;  // this lane's children read (counted only when a.tree is set)
    for (uint32_t base = e0; base < e1; base += 64) {
        const uint32_t e = base + lane;
        const uint32_t w = base == e0 ? w_first : a.out_col[e < e1 ? e : e1 - 1];
        uint64_t wm;
        if (exact) wm = a.lvl[LVLM(d + 1, g, w)] & mask;
        else wm = a.maxdep[MASK(g, w)] >= (uint32_t)(d + 1) ? mask : 0;
        if (e >= e1) wm = 0;
        uint64_t act = __ballot(wm != 0);
        // eight row reads in flight per round
        while (act) {
            uint32_t wq[8];
            uint64_t mq[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (act) {
                    const int j = __builtin_ctzll(act);
                    act &= act - 1;
                    wq[q] = zg_readlane32(w, j);
                    mq[q] = zg_readlane64(wm, j);
                } else {
                    wq[q] = 0;
                    mq[q] = 0;
                }
            }
            double val[8];
#pragma unroll
            for (int q = 0; q < 8; q++) val[q] = ((mq[q] >> lane) & 1) ? cg[ROW(g, wq[q]) + lane] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; q++) S += val[q];
            if (a.tree) {
#pragma unroll
                for (int q = 0; q < 8; q++) tree += val[q] != 0.0;  // c > 0 exactly for a child
            }
        }
    }
    if (a.tree) {
        // spread over slots by block: one counter per group would serialise
        const unsigned long long t = zg_wave_sum_u64(tree);
        if (lane == 0 && t) atomicAdd(&a.tree[(size_t)g * ZG_TREE_SLOTS + (blockIdx.x % ZG_TREE_SLOTS)], t);
    }
    if (hub != ZG_NONE) {
        // one slot per chunk, summed in chunk order by bc_hub_kernel
        if (active) a.hubacc[((size_t)g * a.n_ohub + hub + (oc - first_ch)) * 64 + lane] = S;
        return;
    }
    bc_finish(a, g, v, lane, active, S, d, fin);
}

// Backward level d for vertices with several successor chunks.
__global__ __launch_bounds__(256) void bc_hub_kernel(BcDev a, int d, GroupRange R)
{
    const uint32_t g = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const uint32_t wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= R.cnt[g]) return;
    const size_t idx = g * a.hcap + R.off[g] + wid;
    const uint32_t v = a.hlist[idx];
    const uint64_t mask = a.hmask[idx];
    const bool active = (mask >> lane) & 1;
    const BcFinishIn fin = bc_finish_load(a, g, v, lane, d, active);
    double S = 0.0;
    if (active) {
        const double *p = &a.hubacc[((size_t)g * a.n_ohub + a.ohub_idx[v]) * 64 + lane];
        const uint32_t k = a.och_first[v + 1] - a.och_first[v];
        for (uint32_t q = 0; q < k; q++) S += p[(size_t)q * 64];
    }
    bc_finish(a, g, v, lane, active, S, d, fin);
}

// Folded sweep, after a batch's backward pass: each lane's source m gets
// its folded members' dependency on it, w * (fcount[m] * (reached - 1) -
// folded members reached) (each member's paths all pass through m), one
// thread per group so repeated sources add in lane order.
__global__ void bc_fold_source_kernel(BcDev a)
{
    const uint32_t g = blockIdx.x;
    if (threadIdx.x != 0) return;
    for (int l = 0; l < 64; l++) {
        const uint32_t m = a.src[g * 64 + l];
        if (m == ZG_NONE) continue;
        const double w = a.wts ? a.wts[g * 64 + l] : 1.0;
        const double term = (double)a.fcount[m] * (double)a.rsum[g * 64 + l] - (double)a.fsum[g * 64 + l];
        a.bcg[(size_t)g * a.V + m] += w * term;
    }
}

// out[v] = sum over groups of the group accumulators (fixed order); with a
// view (perm non-null) the accumulators are indexed by the view's ids and
// perm[v] is v's.
__global__ void bc_reduce_kernel(uint32_t V, uint32_t G, const double *bcg, const uint8_t *live, const uint32_t *perm,
                                 double *out)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t x = perm ? perm[v] : v;
    double s = 0.0;
    for (uint32_t g = 0; g < G; g++) s += bcg[(size_t)g * V + x];
    out[v] = live[v] ? s : 0.0;
}

// Depth / sigma export of one group: [lane][v] planes; vertices outside the
// group's span [vlo, vhi) are unreached.
__global__ void bc_export_kernel(uint32_t V, uint32_t Vloc, uint32_t vlo, uint32_t vhi, uint32_t g,
                                 const uint32_t *dep, const double *sig, uint32_t *dep_out, double *sig_out)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t l = blockIdx.y;
    if (v >= V) return;
    uint32_t dd = ZG_NONE;
    double sv = 0.0;
    if (v >= vlo && v < vhi) {
        const size_t r = ((size_t)g * Vloc + (v - vlo)) * 64 + l;
        dd = dep[r];
        sv = dd == ZG_NONE ? 0.0 : sig[r];
    }
    dep_out[(size_t)l * V + v] = dd;
    sig_out[(size_t)l * V + v] = sv;
}

int zg_bc_ws_release(zg_graph *g)
{
    zg_bc_ws &w = g->bc;
    zg_dfree(w.sig); zg_dfree(w.cc); zg_dfree(w.hubacc); zg_dfree(w.bcg); zg_dfree(w.dep); zg_dfree(w.maxdep); zg_dfree(w.lvl);
    zg_dfree(w.vis); zg_dfree(w.fr); zg_dfree(w.stats); zg_dfree(w.tree); zg_dfree(w.touched); zg_dfree(w.flist); zg_dfree(w.fmask);
    zg_dfree(w.hlist); zg_dfree(w.hmask); zg_dfree(w.emax); zg_dfree(w.foff); zg_dfree(w.hoff); zg_dfree(w.modes); zg_dfree(w.src); zg_dfree(w.cnt);
    zg_dfree(w.wts); zg_dfree(w.dsum); zg_dfree(w.rsum); zg_dfree(w.fsum);
    zg_dfree(w.cur);
    if (w.level_args) (void)hipFree(w.level_args);
    if (w.h_level_args) (void)hipHostFree(w.h_level_args);
    if (w.graph) (void)hipGraphExecDestroy(w.graph);
    w.level_args = w.h_level_args = nullptr;
    w.graph = nullptr;
    memset(w.graph_dims, 0, sizeof(w.graph_dims));
    w.cmax = 0;
    if (w.h_cnt) (void)hipHostFree(w.h_cnt);
    w.h_cnt = nullptr;
    w.groups = 0;
    w.lcap = 0;
    w.lmcap = 0;
    w.vloc = 0;
    w.fcap = w.hcap = 0;
    w.with_depth = false;
    return ZG_OK;
}

// Vertex span of a group: the components its sources belong to (components
// are contiguous id ranges; a single graph is one component).
// The span ends at the last graph's last vertex with an arc, or after the
// group's last source if that comes later: the vertices past it have no
// arc, so no search reaches them and no row gathers them (the chunk, hub
// and arc tables at that end equal those at the graph's end).
static void bc_group_span(const zg_graph *g, const uint32_t *src, uint32_t n, uint32_t *vlo, uint32_t *vhi)
{
    const std::vector<uint32_t> &co = g->h_comp_off;
    const std::vector<uint32_t> &act = g->h_comp_act;
    uint32_t lo = ZG_NONE, hi = 0;
    for (uint32_t l = 0; l < n; l++) {
        if (src[l] == ZG_NONE) continue;  // an empty slot
        lo = std::min(lo, src[l]);
        hi = std::max(hi, src[l]);
    }
    if (co.size() <= 2) {
        *vlo = 0;
        *vhi = lo == ZG_NONE || act.size() < 2 ? g->V : std::max(act[0], hi + 1);
        return;
    }
    if (lo == ZG_NONE) {
        *vlo = *vhi = 0;
        return;
    }
    const size_t c_lo = std::upper_bound(co.begin(), co.end(), lo) - co.begin() - 1;
    const size_t c_hi = std::upper_bound(co.begin(), co.end(), hi) - co.begin() - 1;
    *vlo = co[c_lo];
    *vhi = act.size() == co.size() ? std::max(act[c_hi], hi + 1) : co[c_hi + 1];
}

// Sizes the lane-engine state for the sweep over h_sources[0..n): each group
// holds its span's vertices (Vloc of them) and list entries for its span's
// successor chunks and multi-chunk vertices.
static int bc_ws_ensure(zg_graph *g, bool with_depth, const uint32_t *h_sources, uint64_t n)
{
    zg_bc_ws &w = g->bc;
    uint32_t vloc = 1, och = 0, hubs = 0, cmax = 0;
    for (uint64_t b0 = 0; b0 < n; b0 += 64) {
        uint32_t lo, hi;
        bc_group_span(g, h_sources + b0, (uint32_t)std::min<uint64_t>(64, n - b0), &lo, &hi);
        vloc = std::max(vloc, hi - lo);
        const size_t clo = zg_comp_at(g, lo), chi = zg_comp_at(g, hi);
        cmax = std::max(cmax, g->h_cb_ich[chi] - g->h_cb_ich[clo]);
        och = std::max(och, g->h_cb_och[chi] - g->h_cb_och[clo]);
        hubs = std::max(hubs, g->h_cb_hub[chi] - g->h_cb_hub[clo]);
    }
    const size_t need_f = 64 * (size_t)och + 64 * (size_t)g->max_och_per_v;
    const size_t need_h = 64 * (size_t)hubs + 64;
    if (w.groups == g->groups && (w.with_depth || !with_depth) && w.vloc >= vloc && w.fcap >= need_f &&
        w.hcap >= need_h && w.cmax >= cmax)
        return ZG_OK;
    zg_bc_ws_release(g);
    const size_t G = g->groups, V = g->V;
    const size_t VL = vloc;
    w.fcap = need_f;
    w.hcap = need_h;
    w.lcap = 256;
    int rc = ZG_OK;
    if (rc == ZG_OK) rc = zg_dmalloc(&w.sig, G * VL * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.cc, 2 * G * VL * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.hubacc, G * (size_t)g->n_ohub * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.bcg, G * V);
    if (rc == ZG_OK && with_depth) rc = zg_dmalloc(&w.dep, G * VL * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.maxdep, G * VL);
    // per-level lane masks for the backward pass, up to ZG_LVL_LEVELS levels
    // and ZG_LVL_BYTES in all (levels past the table use maxdep alone; when
    // the table cannot be allocated, every level does)
    const size_t lvl_row = G * VL * sizeof(uint64_t);
    w.lmcap = lvl_row ? (uint32_t)std::min<size_t>(ZG_LVL_LEVELS, ZG_LVL_BYTES / lvl_row) : 0;
    if (w.lmcap < 3) w.lmcap = 0;
    if (rc == ZG_OK && w.lmcap && zg_dmalloc(&w.lvl, (size_t)w.lmcap * G * VL) != ZG_OK) w.lmcap = 0;
    // the table starts empty, so a lane bit read there is always a child lane
    if (rc == ZG_OK && w.lmcap) rc = zg_hip_code(hipMemset(w.lvl, 0, sizeof(uint64_t) * (size_t)w.lmcap * G * VL));
    for (int q = 0; q < ZG_MAX_GROUPS; q++) w.lvl_span[q] = 0;
    if (rc == ZG_OK) rc = zg_dmalloc(&w.vis, G * VL);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.fr, 3 * G * VL);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.stats, G * ZG_FIN_BLOCKS * ZG_BC_STATS);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.tree, G * ZG_TREE_SLOTS);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.touched, G * VL);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.flist, G * w.fcap);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.fmask, G * w.fcap);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.hlist, G * w.hcap);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.hmask, G * w.hcap);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.emax, (size_t)w.lcap * G * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.foff, (size_t)w.lcap * G);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.hoff, (size_t)w.lcap * G);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.modes, (size_t)w.lcap * G);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.src, G * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.wts, G * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.dsum, G * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.rsum, G * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.fsum, G * 64);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.cur, 1);
    if (rc == ZG_OK) rc = zg_hip_code(hipMalloc(&w.level_args, sizeof(BcLevelArgs)));
    if (rc == ZG_OK) rc = zg_hip_code(hipHostMalloc(&w.h_level_args, sizeof(BcLevelArgs), hipHostMallocDefault));
    if (rc == ZG_OK) rc = zg_dmalloc(&w.cnt, G);
    if (rc == ZG_OK) {
        hipError_t e = hipHostMalloc((void **)&w.h_cnt, G * sizeof(zg_bc_ws::Counters), hipHostMallocDefault);
        if (e != hipSuccess) rc = zg_hip_code(e);
    }
    if (rc == ZG_OK) {
        hipError_t e = hipMemset(w.hubacc, 0, sizeof(double) * (G * (size_t)g->n_ohub * 64 ? G * (size_t)g->n_ohub * 64 : 1));
        if (e == hipSuccess) e = hipMemset(w.emax, 0, sizeof(int) * (size_t)w.lcap * G * 64);
        if (e != hipSuccess) rc = zg_hip_code(e);
    }
    if (rc != ZG_OK) {
        zg_bc_ws_release(g);
        return rc;
    }
    w.groups = (int)G;
    w.vloc = vloc;
    w.cmax = cmax;
    w.with_depth = with_depth;
    w.last_levels = w.lcap;
    return ZG_OK;
}

// Grows the per-level tables (shifts, list offsets, directions) to hold
// level `need`, keeping their contents.
template <typename T>
static int bc_grow_table(T **tab, size_t old_rows, size_t new_rows, size_t row, hipStream_t s)
{
    T *nt = nullptr;
    ZG_TRY_RC(zg_dmalloc(&nt, new_rows * row));
    ZG_TRY(hipMemsetAsync(nt, 0, sizeof(T) * new_rows * row, s));
    ZG_TRY(hipMemcpyAsync(nt, *tab, sizeof(T) * old_rows * row, hipMemcpyDeviceToDevice, s));
    ZG_TRY(hipStreamSynchronize(s));
    zg_dfree(*tab);
    *tab = nt;
    return ZG_OK;
}

static int bc_grow_levels(zg_graph *g, uint32_t need, hipStream_t s)
{
    zg_bc_ws &w = g->bc;
    if (need < w.lcap) return ZG_OK;
    uint32_t nl = w.lcap;
    while (nl <= need) nl *= 2;
    const size_t G = (size_t)w.groups;
    ZG_TRY_RC(bc_grow_table(&w.emax, w.lcap, nl, G * 64, s));
    ZG_TRY_RC(bc_grow_table(&w.foff, w.lcap, nl, G, s));
    ZG_TRY_RC(bc_grow_table(&w.hoff, w.lcap, nl, G, s));
    ZG_TRY_RC(bc_grow_table(&w.modes, w.lcap, nl, G, s));
    w.lcap = nl;
    return ZG_OK;
}

#define ZG_GRAPH_LEVELS 4  // levels per captured graph; level chunks are multiples of it
static_assert(ZG_LEVEL_CHUNK0 % ZG_GRAPH_LEVELS == 0 && ZG_LEVEL_CHUNK_MAX % ZG_GRAPH_LEVELS == 0,
              "level chunks replay whole graphs");

// Captures ZG_GRAPH_LEVELS forward levels (level, push, pull, sparse and dense
// finalize kernels) reading their arguments from w.level_args; the grids are
// fixed by the workspace (dims: push, pull, sparse finalize, dense finalize
// blocks per group, groups), so one graph serves every chunk of every batch.
static int bc_capture_levels(zg_graph *g, const uint32_t dims[5])
{
    zg_bc_ws &w = g->bc;
    if (w.graph && !memcmp(dims, w.graph_dims, sizeof(w.graph_dims))) return ZG_OK;
    if (w.graph) (void)hipGraphExecDestroy(w.graph);
    w.graph = nullptr;
    if (!g->cap_stream) ZG_TRY(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking));
    hipStream_t cs = g->cap_stream;
    const BcLevelArgs *p = (const BcLevelArgs *)w.level_args;
    const uint32_t G = dims[4];
    ZG_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < ZG_GRAPH_LEVELS; k++) {
        hipLaunchKernelGGL(bc_level_kernel_g, dim3(1), dim3(64), 0, cs, p);
        hipLaunchKernelGGL(bc_push_kernel_g, dim3(dims[0], G), dim3(256), 0, cs, p);
        hipLaunchKernelGGL(bc_forward_kernel_g, dim3(dims[1], G), dim3(256), 0, cs, p);
        hipLaunchKernelGGL(bc_finalize_sparse_kernel_g, dim3(dims[2], G), dim3(ZG_FIN_THREADS), 0, cs, p);
        hipLaunchKernelGGL(bc_finalize_kernel_g, dim3(dims[3], G), dim3(ZG_FIN_THREADS), 0, cs, p);
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&w.graph, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        w.graph = nullptr;
        return zg_hip_code(e);
    }
    memcpy(w.graph_dims, dims, sizeof(w.graph_dims));
    return ZG_OK;
}

// Per-batch callback used by the depth/sigma export.
typedef int (*bc_batch_hook)(zg_graph *g, hipStream_t s, uint64_t first, uint32_t n_in_batch,
                             uint32_t groups_used, void *user);

// Options of one sweep: per-source weights (device, sweep order), per-source
// dependency totals (host, sweep order), per-batch hook.
struct BcSweepOpt {
    const double *d_weights = nullptr;
    double *h_totals = nullptr;
    bool fold = false;          // folded sweep (zg_bc_folded)
    bc_batch_hook hook = nullptr;
    void *user = nullptr;
};

static BcDev bc_dev(zg_graph *g, uint32_t groups_used, const GroupSpan &sp, const BcSweepOpt &opt)
{
    zg_bc_ws &w = g->bc;
    BcDev a;
    a.V = g->V;
    a.Vloc = w.vloc;
    for (int q = 0; q < ZG_MAX_GROUPS; q++) a.vlo[q] = sp.vlo[q];
    a.G = groups_used;
    a.Gcap = (uint32_t)w.groups;
    a.n_ich = g->n_ich;
    a.in_ptr = g->in_ptr; a.in_col = g->in_col; a.out_ptr = g->out_ptr; a.out_col = g->out_col;
    a.ich_v = g->ich_v; a.ich_e0 = g->ich_e0; a.ich_e1 = g->ich_e1;
    a.och_first = g->och_first; a.och_v = g->och_v; a.och_e0 = g->och_e0; a.och_e1 = g->och_e1;
    a.ohub_idx = g->ohub_idx;
    a.live = g->live;
    a.sig = w.sig; a.cc = w.cc; a.hubacc = w.hubacc; a.bcg = w.bcg; a.dep = w.dep; a.maxdep = w.maxdep;
    a.lvl = w.lvl; a.lmcap = w.lvl ? w.lmcap : 0;
    a.inplace = opt.hook ? 0u : 1u;
    a.vis = w.vis; a.fr = w.fr; a.stats = w.stats; a.touched = w.touched;
    a.tree = g->count_tree ? w.tree : nullptr;
    a.flist = w.flist; a.fmask = w.fmask; a.hlist = w.hlist; a.hmask = w.hmask;
    a.emax = w.emax; a.src = w.src; a.cnt = w.cnt;
    a.wts = opt.d_weights ? w.wts : nullptr;
    a.dsum = opt.h_totals ? w.dsum : nullptr;
    a.succ1 = opt.fold ? g->fold_succ1 : nullptr;
    a.fcount = opt.fold ? g->fold_count : nullptr;
    a.rsum = opt.fold ? w.rsum : nullptr;
    a.fsum = opt.fold ? w.fsum : nullptr;
    a.foff = w.foff; a.hoff = w.hoff; a.modes = w.modes;
    a.push_alpha = g->push_alpha;
    for (int q = 0; q < ZG_MAX_GROUPS; q++)
        a.Eg[q] = sp.vhi[q] > sp.vlo[q] ? g->h_cb_in[zg_comp_at(g, sp.vhi[q])] - g->h_cb_in[zg_comp_at(g, sp.vlo[q])] : 0;
    a.fcap = w.fcap; a.hcap = w.hcap;
    a.n_ohub = g->n_ohub;
    return a;
}

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Runs the batched sweep over d_sources[0..n) accumulating into the group
// accumulators; the caller reduces them.
static int bc_sweep(zg_graph *g, const uint32_t *d_sources, const uint32_t *h_sources, uint64_t n, hipStream_t s,
                    zg_bc_info *inf, const BcSweepOpt &opt)
{
    zg_bc_ws &w = g->bc;
    const uint32_t V = g->V;
    const uint32_t Gcap = (uint32_t)w.groups;
    const bool use_graph = g->use_graph && !g->profiling;
    ZG_TRY(hipMemsetAsync(w.bcg, 0, sizeof(double) * (size_t)Gcap * V, s));
    // per-block work statistics accumulate over all batches
    ZG_TRY(hipMemsetAsync(w.stats, 0, sizeof(unsigned long long) * Gcap * ZG_FIN_BLOCKS * ZG_BC_STATS, s));
    ZG_TRY(hipMemsetAsync(w.tree, 0, sizeof(unsigned long long) * Gcap * ZG_TREE_SLOTS, s));
    std::vector<std::vector<uint32_t>> foff(Gcap), hoff(Gcap);
    double span_used = 0.0, span_grid = 0.0;  // group spans swept, and the launch grids' span capacity
    for (uint64_t b0 = 0; b0 < n; b0 += (uint64_t)Gcap * 64) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>((uint64_t)Gcap * 64, n - b0);
        const uint32_t G = cdiv(nb, 64);
        // per-group spans: the components its sources belong to
        GroupSpan sp;
        memset(&sp, 0, sizeof(sp));
        uint32_t max_v = 0, max_c = 0;
        for (uint32_t q = 0; q < G; q++) {
            const uint32_t slots = std::min<uint32_t>(64, nb - 64 * q);
            bc_group_span(g, h_sources + b0 + 64 * q, slots, &sp.vlo[q], &sp.vhi[q]);
            // a padded group's empty slots are its last ones
            sp.used[q] = slots;
            while (sp.used[q] > 0 && h_sources[b0 + 64 * q + sp.used[q] - 1] == ZG_NONE) sp.used[q]--;
            if (sp.vhi[q] - sp.vlo[q] > w.vloc) return ZG_ERR_INVALID_ARG;  // workspace sized for other sources
            sp.clo[q] = g->h_cb_ich[zg_comp_at(g, sp.vlo[q])];
            sp.chi[q] = g->h_cb_ich[zg_comp_at(g, sp.vhi[q])];
            max_v = std::max(max_v, sp.vhi[q] - sp.vlo[q]);
            max_c = std::max(max_c, sp.chi[q] - sp.clo[q]);
        }
        BcDev a = bc_dev(g, G, sp, opt);
        inf->group_slots += 64ull * G;
        {
            uint64_t span_sum = 0;
            for (uint32_t q = 0; q < G; q++) span_sum += sp.vhi[q] - sp.vlo[q];
            span_used += (double)span_sum;
            span_grid += (double)G * (double)max_v;
        }
        // batch reset
        ZG_TRY(hipMemsetAsync(w.src, 0xFF, sizeof(uint32_t) * Gcap * 64, s));
        ZG_TRY(hipMemcpyAsync(w.src, d_sources + b0, sizeof(uint32_t) * nb, hipMemcpyDeviceToDevice, s));
        if (opt.d_weights)
            ZG_TRY(hipMemcpyAsync(w.wts, opt.d_weights + b0, sizeof(double) * nb, hipMemcpyDeviceToDevice, s));
        if (opt.h_totals) ZG_TRY(hipMemsetAsync(w.dsum, 0, sizeof(unsigned long long) * Gcap * 64, s));
        if (opt.fold) {
            ZG_TRY(hipMemsetAsync(w.rsum, 0, sizeof(unsigned long long) * Gcap * 64, s));
            ZG_TRY(hipMemsetAsync(w.fsum, 0, sizeof(unsigned long long) * Gcap * 64, s));
        }
        ZG_TRY(hipMemsetAsync(w.cnt, 0, sizeof(zg_bc_ws::Counters) * Gcap, s));
        ZG_TRY(hipMemsetAsync(w.emax, 0, sizeof(int) * (size_t)std::min(w.last_levels + 2, w.lcap) * Gcap * 64, s));
        const uint32_t nlev = a.lmcap ? std::min(w.last_levels + 2, a.lmcap) : 0;
        if (nlev) {
            // rows an earlier, larger span left bits in (the reset below covers
            // this batch's spans)
            LvlExtent ext;
            uint32_t most = 0;
            for (uint32_t q = 0; q < Gcap; q++) {
                const uint32_t cur = q < G ? sp.vhi[q] - sp.vlo[q] : 0u;
                ext.from[q] = q < G ? cur : 0u;
                ext.to[q] = q < G ? std::max(cur, w.lvl_span[q]) : 0u;
                most = std::max(most, ext.to[q] - ext.from[q]);
                if (q < G) w.lvl_span[q] = cur;
            }
            if (most)
                ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bc_lvl_clear_kernel, dim3(cdiv(most, 256), Gcap), dim3(256), 0, s, a,
                               ext, nlev);
        }
        if (max_v) {
            ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bc_reset_kernel, dim3(cdiv(max_v, 256), G), dim3(256), 0, s, a, sp, nlev);
            ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bc_zero_rows_kernel, dim3(cdiv(max_v, 4), G), dim3(256), 0, s, a, sp);
        }
        ZG_TRY(hipMemsetAsync(w.foff, 0, sizeof(uint32_t) * Gcap, s));  // level 0 starts at 0
        ZG_TRY(hipMemsetAsync(w.hoff, 0, sizeof(uint32_t) * Gcap, s));
        ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bc_source_kernel, dim3(1, G), dim3(64), 0, s, a);
        ZG_TRY(hipGetLastError());
        // Forward sweep, device-driven: each level opens with bc_level_kernel
        // (offsets, direction, counter resets); push, pull and both finalize
        // variants are queued and all but the chosen ones return at once.
        // The host syncs once per chunk of levels to find the last one.
        // Replayed from a captured graph, the grids follow the workspace's
        // largest span (the same for every batch); launched directly, they
        // follow this batch's largest span.
        const uint32_t span_c = use_graph ? w.cmax : max_c, span_v = use_graph ? w.vloc : max_v;
        uint32_t fwd_blocks = std::max<uint32_t>(1, cdiv(span_c, 4));
        if (g->pull_grid) fwd_blocks = std::min(fwd_blocks, g->pull_grid);
        const uint32_t fin_blocks = std::min<uint32_t>(ZG_FIN_BLOCKS, std::max<uint32_t>(1, cdiv(span_c, ZG_FIN_CHUNKS)));
        const uint32_t sfin_blocks =
            std::min<uint32_t>(ZG_SPARSE_FIN_BLOCKS, std::max<uint32_t>(1, cdiv(span_v, ZG_SPARSE_FIN_VERTS)));
        const uint32_t push_blocks = std::min<uint32_t>(ZG_PUSH_GRID, std::max<uint32_t>(1, cdiv(span_v, 32)));
        if (use_graph) {
            const uint32_t dims[5] = {push_blocks, fwd_blocks, sfin_blocks, fin_blocks, Gcap};
            ZG_TRY_RC(bc_capture_levels(g, dims));
            ZG_TRY(hipMemsetAsync(w.cur, 0xFF, sizeof(int), s));  // level counter = -1
        }
        std::vector<uint32_t> h_off, h_hoff, h_modes;  // [L][Gcap] copies
        int d = 0, maxd = -1;
        uint32_t chunk = ZG_LEVEL_CHUNK0;
        while (maxd < 0) {
            ZG_TRY_RC(bc_grow_levels(g, (uint32_t)d + chunk + 2, s));
            a = bc_dev(g, G, sp, opt);
            if (use_graph) {
                // the previous copy from the staging buffer has run: the stream
                // was synchronised after it
                BcLevelArgs *h = (BcLevelArgs *)w.h_level_args;
                h->a = a;
                h->sp = sp;
                h->cur = w.cur;
                ZG_TRY(hipMemcpyAsync(w.level_args, h, sizeof(BcLevelArgs), hipMemcpyHostToDevice, s));
                for (uint32_t k = 0; k < chunk; k += ZG_GRAPH_LEVELS) ZG_TRY(hipGraphLaunch(w.graph, s));
            }
            for (uint32_t k = 0; k < (use_graph ? 0u : chunk); k++) {
                const int L = d + (int)k;
                ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bc_level_kernel, dim3(1), dim3(64), 0, s, a, L);
                ZG_PROF_LAUNCH(ZG_PROF_BC_PUSH, bc_push_kernel, dim3(push_blocks, G), dim3(256), 0, s, a, L);
                if (max_c)
                    ZG_PROF_LAUNCH(ZG_PROF_BC_FORWARD, bc_forward_kernel, dim3(fwd_blocks, G), dim3(256), 0, s, a, L, sp);
                ZG_PROF_LAUNCH(ZG_PROF_BC_FINALIZE, bc_finalize_sparse_kernel, dim3(sfin_blocks, G),
                               dim3(ZG_FIN_THREADS), 0, s, a, L);
                ZG_PROF_LAUNCH(ZG_PROF_BC_FINALIZE, bc_finalize_kernel, dim3(fin_blocks, G), dim3(ZG_FIN_THREADS), 0,
                               s, a, L, 0, sp);
            }
            ZG_TRY(hipGetLastError());
            // offsets of levels up to d+chunk, directions of levels d..d+chunk-1
            const size_t rows = (size_t)d + chunk + 1;
            h_off.resize(rows * Gcap);
            h_hoff.resize(rows * Gcap);
            h_modes.resize(rows * Gcap);
            ZG_TRY(hipMemcpyAsync(h_off.data(), w.foff, sizeof(uint32_t) * rows * Gcap, hipMemcpyDeviceToHost, s));
            ZG_TRY(hipMemcpyAsync(h_hoff.data(), w.hoff, sizeof(uint32_t) * rows * Gcap, hipMemcpyDeviceToHost, s));
            ZG_TRY(hipMemcpyAsync(h_modes.data(), w.modes, sizeof(uint32_t) * rows * Gcap, hipMemcpyDeviceToHost, s));
            ZG_TRY(hipMemcpyAsync(w.h_cnt, w.cnt, sizeof(zg_bc_ws::Counters) * G, hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            for (uint32_t q = 0; q < G; q++) {
                if (w.h_cnt[q].err & ZG_BC_ERR_SOURCE) return ZG_ERR_NODE_NOT_FOUND;
                if (w.h_cnt[q].err) return ZG_ERR_HIP;  // work-list capacity exceeded
            }
            // level L holds entries [off[L], off[L+1]); the level after the
            // chunk ends at fcount
            bool tail_busy = false;
            for (uint32_t q = 0; q < G; q++) tail_busy |= w.h_cnt[q].fcount > h_off[(size_t)(d + chunk) * Gcap + q];
            if (tail_busy) {
                d += (int)chunk;
                chunk = std::min<uint32_t>(chunk * 2, ZG_LEVEL_CHUNK_MAX);
                continue;
            }
            for (int L = d + (int)chunk - 1; L >= 0 && maxd < 0; L--)
                for (uint32_t q = 0; q < G; q++)
                    if (h_off[(size_t)(L + 1) * Gcap + q] > h_off[(size_t)L * Gcap + q]) maxd = L;
            if (maxd < 0) maxd = 0;
        }
        for (int L = 0; L <= maxd; L++) {
            bool push = false, pull = false;
            for (uint32_t q = 0; q < G; q++) {
                push |= h_modes[(size_t)L * Gcap + q] == ZG_MODE_PUSH;
                pull |= h_modes[(size_t)L * Gcap + q] == ZG_MODE_PULL;
            }
            inf->push_levels += push;
            inf->pull_levels += pull;
        }
        for (uint32_t q = 0; q < G; q++) {
            foff[q].resize((size_t)maxd + 2);
            hoff[q].resize((size_t)maxd + 2);
            for (int L = 0; L <= maxd + 1; L++) {
                foff[q][L] = h_off[(size_t)L * Gcap + q];
                hoff[q][L] = h_hoff[(size_t)L * Gcap + q];
            }
        }
        inf->max_levels = std::max<uint32_t>(inf->max_levels, (uint32_t)maxd + 1);
        w.last_levels = (uint32_t)maxd + 2;
        // the parity buffers, read where a level's children lie past the lane
        // table (or for every level when the terms do not stay in place), start
        // at zero: a lane is read there only as a child or at zero
        if (!a.inplace || (uint32_t)maxd + 1 >= a.lmcap)
            ZG_TRY(hipMemsetAsync(w.cc, 0, sizeof(double) * 2 * (size_t)Gcap * w.vloc * 64, s));
        for (int L = maxd; L >= 1; L--) {
            GroupRange R, H;
            memset(&R, 0, sizeof(R));
            memset(&H, 0, sizeof(H));
            uint32_t mr = 0, mh = 0;
            for (uint32_t q = 0; q < G; q++) {
                R.off[q] = foff[q][L];
                R.cnt[q] = foff[q][L + 1] - foff[q][L];
                H.off[q] = hoff[q][L];
                H.cnt[q] = hoff[q][L + 1] - hoff[q][L];
                mr = std::max(mr, R.cnt[q]);
                mh = std::max(mh, H.cnt[q]);
            }
            if (mr) ZG_PROF_LAUNCH(ZG_PROF_BC_BACKWARD, bc_backward_kernel, dim3(cdiv(mr, 4), G), dim3(256), 0, s, a, L, R);
            if (mh) ZG_PROF_LAUNCH(ZG_PROF_BC_HUB, bc_hub_kernel, dim3(cdiv(mh, 4), G), dim3(256), 0, s, a, L, H);
        }
        if (opt.fold) ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, bc_fold_source_kernel, dim3(G), dim3(64), 0, s, a);
        ZG_TRY(hipGetLastError());
        inf->batches++;
        if (opt.h_totals) {
            std::vector<unsigned long long> t(nb);
            ZG_TRY(hipMemcpyAsync(t.data(), w.dsum, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            for (uint32_t i = 0; i < nb; i++) opt.h_totals[b0 + i] = (double)t[i];
        }
        if (opt.hook) ZG_TRY_RC(opt.hook(g, s, b0, nb, G, opt.user));
    }
    {
        std::vector<unsigned long long> st((size_t)Gcap * ZG_FIN_BLOCKS * ZG_BC_STATS), tr((size_t)Gcap * ZG_TREE_SLOTS);
        ZG_TRY(hipMemcpyAsync(st.data(), w.stats, sizeof(unsigned long long) * st.size(), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipMemcpyAsync(tr.data(), w.tree, sizeof(unsigned long long) * tr.size(), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        for (size_t i = 0; i < st.size(); i += ZG_BC_STATS) {
            inf->arcs_traversed += st[i];
            inf->nodes_reached += st[i + 1];
            inf->row_visits += st[i + 2];
            inf->row_arcs += st[i + 3];
        }
        for (unsigned long long t : tr) inf->tree_arcs += t;
    }
    inf->n_sources = n;
    inf->span_balance = span_grid > 0.0 ? span_used / span_grid : 1.0;
    return ZG_OK;
}

// One level of a reverse BFS from the hub core: a vertex joins level d+1 if
// one of its successors sits at level d.
__global__ void bc_coredist_kernel(uint32_t V, const uint32_t *out_ptr, const uint32_t *out_col, uint32_t *dist,
                                   uint32_t d, uint32_t *changed)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || dist[v] != ZG_NONE) return;
    for (uint32_t e = out_ptr[v], e1 = out_ptr[v + 1]; e < e1; e++) {
        if (dist[out_col[e]] == d) {
            dist[v] = d + 1;
            *changed = 1;
            return;
        }
    }
}

#define ZG_CORE_SIZE 16
#define ZG_CORE_MAX_LEVELS 64

// Component c of vertex v: co[c] <= v < co[c + 1].
__device__ __forceinline__ uint32_t bc_comp_of(const uint32_t *co, uint32_t ncomp, uint32_t v)
{
    uint32_t lo = 0, hi = ncomp;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (co[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Sort keys for the choice of the hub core: live vertices by decreasing
// in-degree, then the vertices that are not live; a second stable sort by
// component (keys = components of the first sort's output) groups them by
// graph.
__global__ void bc_core_key_kernel(uint32_t V, const uint32_t *in_ptr, const uint8_t *live, uint32_t *key,
                                   uint32_t *val)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const uint32_t indeg = in_ptr[v + 1] - in_ptr[v];
        key[v] = live[v] ? 0xFFFFFFFEu - min(indeg, 0xFFFFFFFEu) : 0xFFFFFFFFu;
        val[v] = v;
    }
}

__global__ void bc_core_comp_kernel(uint32_t V, uint32_t ncomp, const uint32_t *co, const uint32_t *vtx, uint32_t *key)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < V; j += gridDim.x * blockDim.x)
        key[j] = bc_comp_of(co, ncomp, vtx[j]);
}

// The first `per` live vertices of each component in key order are the core
// (level 0).
__global__ void bc_core_mark_kernel(uint32_t ncomp, uint32_t per, const uint32_t *co, const uint32_t *order,
                                    const uint8_t *live, uint32_t *dist)
{
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncomp; c += gridDim.x * blockDim.x)
        for (uint32_t j = co[c]; j < co[c + 1] && j - co[c] < per; j++) {
            const uint32_t v = order[j];
            if (!live[v]) break;  // the component's not-live vertices come last
            dist[v] = 0;
        }
}

// Distance of every vertex to the hub core -- the ZG_CORE_SIZE highest
// in-degree live vertices of a single graph, the highest one of each graph
// of a batch -- capped at ZG_CORE_MAX_LEVELS levels (unreachable = ZG_NONE);
// computed on the device, kept on the host with the out-degrees for the
// sweep order of later calls.
static int bc_core_distances(zg_graph *g, hipStream_t s)
{
    if (!g->h_coredist.empty() || g->V == 0) return ZG_OK;
    const uint32_t V = g->V;
    const std::vector<uint32_t> &co = g->h_comp_off;
    const uint32_t ncomp = (uint32_t)(co.size() - 1);
    const uint32_t per = ncomp == 1 ? ZG_CORE_SIZE : 1;
    ZgScratch sc;
    uint32_t *d_co, *d_dist, *d_changed, *key, *key2, *val, *val2, *order;
    ZG_TRY_RC(sc.get(&d_co, co.size()));
    ZG_TRY_RC(sc.get(&d_dist, V));
    ZG_TRY_RC(sc.get(&d_changed, 1));
    ZG_TRY_RC(sc.get(&key, V));
    ZG_TRY_RC(sc.get(&key2, V));
    ZG_TRY_RC(sc.get(&val, V));
    ZG_TRY_RC(sc.get(&val2, V));
    ZG_TRY(hipMemcpyAsync(d_co, co.data(), sizeof(uint32_t) * co.size(), hipMemcpyHostToDevice, s));
    const dim3 grd(std::min<uint32_t>(cdiv(V, 256), 8192));
    hipLaunchKernelGGL(bc_core_key_kernel, grd, dim3(256), 0, s, V, g->in_ptr, g->live, key, val);
    ZG_TRY(hipGetLastError());
    ZG_TRY_RC(zg_sort_pairs(key, key2, val, val2, V, 32, s, sc));
    order = val2;
    if (ncomp > 1) {
        hipLaunchKernelGGL(bc_core_comp_kernel, grd, dim3(256), 0, s, V, ncomp, d_co, val2, key);
        ZG_TRY(hipGetLastError());
        unsigned bits = 1;
        while (bits < 32 && ((ncomp - 1) >> bits)) bits++;
        ZG_TRY_RC(zg_sort_pairs(key, key2, val2, val, V, bits, s, sc));
        order = val;
    }
    ZG_TRY(hipMemsetAsync(d_dist, 0xFF, sizeof(uint32_t) * V, s));
    hipLaunchKernelGGL(bc_core_mark_kernel, dim3(std::min<uint32_t>(cdiv(ncomp, 256), 8192)), dim3(256), 0, s, ncomp,
                       per, d_co, order, g->live, d_dist);
    ZG_TRY(hipGetLastError());
    for (uint32_t d = 0; d < ZG_CORE_MAX_LEVELS; d++) {
        uint32_t changed = 0;
        ZG_TRY(hipMemsetAsync(d_changed, 0, sizeof(uint32_t), s));
        hipLaunchKernelGGL(bc_coredist_kernel, dim3(cdiv(V, 256)), dim3(256), 0, s, V, g->out_ptr, g->out_col, d_dist,
                           d, d_changed);
        ZG_TRY(hipGetLastError());
        ZG_TRY(hipMemcpyAsync(&changed, d_changed, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        if (!changed) break;
        if (d + 1 == ZG_CORE_MAX_LEVELS) g->high_diameter = true;  // still growing at the cap
    }
    std::vector<uint32_t> dist(V), ptr((size_t)V + 1);
    ZG_TRY(hipMemcpyAsync(dist.data(), d_dist, sizeof(uint32_t) * V, hipMemcpyDeviceToHost, s));
    ZG_TRY(hipMemcpyAsync(ptr.data(), g->out_ptr, sizeof(uint32_t) * ptr.size(), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    g->h_out_deg.resize(V);
    for (uint32_t v = 0; v < V; v++) g->h_out_deg[v] = ptr[(size_t)v + 1] - ptr[v];
    g->h_coredist.swap(dist);
    return ZG_OK;
}

// Folding of single-successor sources.  A source s whose only successor
// (self-loops aside; parallel arcs counted once) is w reaches everything
// through w, so its dependencies follow from w's sweep: with p_wt(v) the share
// of shortest w-t paths through v, delta_s(v) = delta_w(v) - p_ws(v) for
// v != s, w, and delta_s(w) = (reached from w) - 1 - [w reaches s].  Summed
// over the c_w sources folded into w, w's group {w} + folded contributes
// sum_t tau(t) p_wt(v) with tau(t) = 1 + c_w - [t folded into w] -- a Brandes
// dependency pass with per-target weights -- and, at w itself,
// c_w ((reached from w) - 1) - (folded members w reaches).  A source is
// folded only into a vertex with several successors, so groups do not nest.

// class[v]: 0 = no successor but itself, 1 = exactly one other, 2 = more;
// first[v] = that one successor (class 1).  Stops at the second distinct
// successor, so a row is read only as far as it must be.
__global__ void fold_class_kernel(uint32_t V, const uint32_t *out_ptr, const uint32_t *out_col, const uint8_t *live,
                                  uint8_t *cls, uint32_t *first)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint8_t c = 0;
    uint32_t w1 = ZG_NONE;
    if (live[v]) {
        for (uint32_t e = out_ptr[v]; e < out_ptr[v + 1]; e++) {
            const uint32_t w = out_col[e];
            if (w == v || w == w1) continue;
            if (w1 == ZG_NONE) {
                w1 = w;
                c = 1;
            } else {
                c = 2;
                break;
            }
        }
    }
    cls[v] = c;
    first[v] = c == 1 ? w1 : ZG_NONE;
}

__global__ void fold_map_kernel(uint32_t V, const uint8_t *cls, const uint32_t *first, uint32_t *succ1, uint32_t *count)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint32_t m = ZG_NONE;
    if (cls[v] == 1 && cls[first[v]] == 2) m = first[v];
    succ1[v] = m;
    if (m != ZG_NONE) atomicAdd(&count[m], 1u);
}

static int bc_fold_build(zg_graph *g, hipStream_t s)
{
    if (g->fold_built) return ZG_OK;
    const uint32_t V = g->V;
    uint8_t *d_cls = nullptr;
    uint32_t *d_first = nullptr;
    int rc = zg_dmalloc(&d_cls, V);
    if (rc == ZG_OK) rc = zg_dmalloc(&d_first, V);
    if (rc == ZG_OK && !g->fold_succ1) rc = zg_dmalloc(&g->fold_succ1, V);
    if (rc == ZG_OK && !g->fold_count) rc = zg_dmalloc(&g->fold_count, V);
    if (rc == ZG_OK && V) {
        rc = zg_hip_code(hipMemsetAsync(g->fold_count, 0, sizeof(uint32_t) * V, s));
        if (rc == ZG_OK) {
            hipLaunchKernelGGL(fold_class_kernel, dim3(cdiv(V, 256)), dim3(256), 0, s, V, g->out_ptr, g->out_col,
                               g->live, d_cls, d_first);
            hipLaunchKernelGGL(fold_map_kernel, dim3(cdiv(V, 256)), dim3(256), 0, s, V, d_cls, d_first, g->fold_succ1,
                               g->fold_count);
            rc = zg_hip_code(hipGetLastError());
        }
    }
    if (rc == ZG_OK) {
        g->h_fold_class.resize(V);
        g->h_fold_first.resize(V);
        g->h_fold_succ1.resize(V);
        g->h_fold_count.resize(V);
        if (V) {
            rc = zg_hip_code(hipMemcpyAsync(g->h_fold_class.data(), d_cls, V, hipMemcpyDeviceToHost, s));
            if (rc == ZG_OK)
                rc = zg_hip_code(hipMemcpyAsync(g->h_fold_first.data(), d_first, 4ull * V, hipMemcpyDeviceToHost, s));
            if (rc == ZG_OK)
                rc = zg_hip_code(hipMemcpyAsync(g->h_fold_succ1.data(), g->fold_succ1, 4ull * V, hipMemcpyDeviceToHost, s));
            if (rc == ZG_OK)
                rc = zg_hip_code(hipMemcpyAsync(g->h_fold_count.data(), g->fold_count, 4ull * V, hipMemcpyDeviceToHost, s));
        }
        if (rc == ZG_OK) rc = zg_hip_code(hipStreamSynchronize(s));
    }
    zg_dfree(d_cls);
    zg_dfree(d_first);
    if (rc == ZG_OK) g->fold_built = true;
    return rc;
}

extern "C" int zg_fold_map(zg_graph *g, uint8_t *class_out, uint32_t *succ_out, uint32_t *folded_into_out,
                           uint32_t *count_out)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    ZG_TRY_RC(bc_fold_build(g, g->stream));
    const size_t V = g->V;
    if (class_out && V) memcpy(class_out, g->h_fold_class.data(), V);
    if (succ_out && V) memcpy(succ_out, g->h_fold_first.data(), 4 * V);
    if (folded_into_out && V) memcpy(folded_into_out, g->h_fold_succ1.data(), 4 * V);
    if (count_out && V) memcpy(count_out, g->h_fold_count.data(), 4 * V);
    return ZG_OK;
}

// Discovery rank of every vertex in a breadth-first search from the hub core
// over predecessor arcs (the search that fills the core distances; siblings
// get neighbouring ranks); vertices it does not reach follow in id order.
static int bc_bfs_rank(zg_graph *g, hipStream_t s)
{
    if (!g->h_bfsrank.empty() || g->V == 0) return ZG_OK;
    ZG_TRY_RC(bc_core_distances(g, s));
    const uint32_t V = g->V;
    std::vector<uint32_t> col(g->E);
    if (g->E) {
        ZG_TRY(hipMemcpyAsync(col.data(), g->in_col, sizeof(uint32_t) * g->E, hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
    }
    std::vector<uint32_t> ip((size_t)V + 1);
    ZG_TRY(hipMemcpyAsync(ip.data(), g->in_ptr, sizeof(uint32_t) * ip.size(), hipMemcpyDeviceToHost, s));
    ZG_TRY(hipStreamSynchronize(s));
    std::vector<uint32_t> rank(V, ZG_NONE), queue;
    queue.reserve(V);
    for (uint32_t v = 0; v < V; v++)
        if (g->h_coredist[v] == 0) {
            rank[v] = (uint32_t)queue.size();
            queue.push_back(v);
        }
    for (size_t h = 0; h < queue.size(); h++) {
        const uint32_t w = queue[h];
        for (uint32_t e = ip[w]; e < ip[w + 1]; e++) {
            const uint32_t u = col[e];
            if (rank[u] == ZG_NONE && g->h_live[u]) {
                rank[u] = (uint32_t)queue.size();
                queue.push_back(u);
            }
        }
    }
    uint32_t next = (uint32_t)queue.size();
    for (uint32_t v = 0; v < V; v++)
        if (rank[v] == ZG_NONE) rank[v] = next++;
    g->h_bfsrank.swap(rank);
    return ZG_OK;
}

// Order in which the sources are swept: graphs (components) by decreasing
// size, and inside one graph the sources at the same distance from its hub
// core together (then larger out-degree first, then id), or (mode 2) in the
// order a search from the hub core discovers them.  Returns the positions of
// src in sweep order.
static void bc_order_perm(const zg_graph *g, const std::vector<uint32_t> &src, std::vector<uint64_t> &perm)
{
    const std::vector<uint32_t> &cd = g->h_coredist;
    const std::vector<uint32_t> &co = g->h_comp_off;
    auto comp = [&](uint32_t x) { return std::upper_bound(co.begin(), co.end(), x) - co.begin(); };
    perm.resize(src.size());
    for (size_t i = 0; i < perm.size(); i++) perm[i] = i;
    const bool by_rank = g->source_order == 2;
    std::stable_sort(perm.begin(), perm.end(), [&](uint64_t i, uint64_t j) {
        const uint32_t x = src[i], y = src[j];
        if (co.size() > 2) {
            // largest graph first (longest-processing-time order): the groups
            // of one batch then come from graphs of about the same size, so
            // one large graph does not hold a batch of small ones
            const auto cx = comp(x), cy = comp(y);
            if (cx != cy) {
                const uint32_t nx = co[cx] - co[cx - 1], ny = co[cy] - co[cy - 1];
                return nx != ny ? nx > ny : cx < cy;
            }
        }
        if (by_rank) return g->h_bfsrank[x] < g->h_bfsrank[y];
        if (cd[x] != cd[y]) return cd[x] < cd[y];
        if (g->h_out_deg[x] != g->h_out_deg[y]) return g->h_out_deg[x] > g->h_out_deg[y];
        return x < y;
    });
}

// Group layout of the lane engine on a resource of several graphs: a group
// takes the next sources of `list` while its span (the components from its
// first to its last source, which are contiguous id ranges) stays within the
// widest component that holds a source; otherwise the group is closed and its
// remaining slots stay empty (ZG_NONE).  Sources that alternate between far
// apart graphs thus never make a group's state span the graphs between them.
// slot[i] = position of list[i] in the padded list.
static void bc_pad_groups(const zg_graph *g, const std::vector<uint32_t> &list, std::vector<uint32_t> &padded,
                          std::vector<uint64_t> &slot)
{
    const std::vector<uint32_t> &co = g->h_comp_off;
    auto comp = [&](uint32_t x) { return (size_t)(std::upper_bound(co.begin(), co.end(), x) - co.begin() - 1); };
    uint32_t widest = 1;
    for (uint32_t x : list) widest = std::max(widest, co[comp(x) + 1] - co[comp(x)]);
    padded.clear();
    padded.reserve(list.size());
    slot.resize(list.size());
    size_t c_lo = 0, c_hi = 0, fill = 0;
    for (size_t i = 0; i < list.size(); i++) {
        const size_t c = comp(list[i]);
        if (fill > 0) {
            const size_t lo = std::min(c_lo, c), hi = std::max(c_hi, c);
            if (fill == 64 || co[hi + 1] - co[lo] > widest) {
                while (fill < 64) {
                    padded.push_back(ZG_NONE);
                    fill++;
                }
                fill = 0;
            } else {
                c_lo = lo;
                c_hi = hi;
            }
        }
        if (fill == 0) c_lo = c_hi = c;
        slot[i] = padded.size();
        padded.push_back(list[i]);
        fill++;
    }
}

static void bc_all_live_sources(zg_graph *g, std::vector<uint32_t> &out)
{
    out.clear();
    out.reserve(g->n_live);
    for (uint32_t v = 0; v < g->V; v++)
        if (g->h_live[v]) out.push_back(v);
}

// One Brandes call: sources (device list, or null = every live vertex in id
// order), optional per-source weights and dependency totals (host, in the
// caller's source order), optional output (device [V]).
struct BcCall {
    const uint32_t *d_sources = nullptr;
    uint64_t n_sources = 0;
    const double *h_weights = nullptr;
    double *h_totals = nullptr;
    double *d_out = nullptr;
    bool fold = false;          // every source stands for itself and the sources folded into it
};

static int zg_bc_device_impl(zg_graph *g, const BcCall &call, hipStream_t s, zg_bc_info *info)
{
    zg_bc_info inf;
    memset(&inf, 0, sizeof(inf));
    const uint32_t V = g->V;
    // the sweep needs the source list on the host (ordering, per-group spans)
    std::vector<uint32_t> list;
    if (!call.d_sources) {
        bc_all_live_sources(g, list);
    } else {
        list.resize(call.n_sources);
        if (call.n_sources) {
            ZG_TRY(hipMemcpyAsync(list.data(), call.d_sources, sizeof(uint32_t) * call.n_sources,
                                  hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
        }
        for (uint32_t x : list)
            if (x >= V || !g->h_live[x]) return ZG_ERR_NODE_NOT_FOUND;
    }
    const uint64_t n = list.size();
    const bool order = g->source_order && n > 64;
    // engine: the per-source queue engine for high-diameter graphs
    if (g->engine == ZG_ENGINE_AUTO) ZG_TRY_RC(bc_core_distances(g, s));
    // the folded sweep runs on the 64-lane engine
    const bool queue_engine = !call.fold &&
                              (g->engine == ZG_ENGINE_QUEUE || (g->engine == ZG_ENGINE_AUTO && g->high_diameter));
    if (call.fold) {
        ZG_TRY_RC(bc_fold_build(g, s));
        for (uint32_t x : list)
            if (g->h_fold_succ1[x] != ZG_NONE) return ZG_ERR_INVALID_ARG;  // counted with its successor
    }
    // the lane engine runs on the degree-ordered view when there is one; the
    // sources are then ordered and grouped by the view's own searches
    zg_graph *eg = g;
    if (!queue_engine && !call.fold && g->lane_order) {
        ZG_TRY_RC(zg_lane_view(g, s));
        if (g->lane) {
            eg = g->lane;
            eg->groups = g->groups;
            eg->push_alpha = g->push_alpha;
            eg->pull_grid = g->pull_grid;
            eg->use_graph = g->use_graph;
            eg->count_tree = g->count_tree;
            eg->profiling = g->profiling;
            eg->source_order = g->source_order;
            for (uint32_t &x : list) x = g->h_lane_perm[x];
        }
    }
    if (order) ZG_TRY_RC(bc_core_distances(eg, s));
    if (order && g->source_order == 2) ZG_TRY_RC(bc_bfs_rank(eg, s));
    std::vector<uint64_t> perm;
    if (order) {
        bc_order_perm(eg, list, perm);
        std::vector<uint32_t> sorted(n);
        for (uint64_t i = 0; i < n; i++) sorted[i] = list[perm[i]];
        list.swap(sorted);
    }
    // the lane engine on several graphs sweeps a padded list (no group spans
    // far apart graphs); slot maps the sorted list into it
    std::vector<uint64_t> slot;
    if (!queue_engine && g->h_comp_off.size() > 2) {
        std::vector<uint32_t> padded;
        bc_pad_groups(eg, list, padded, slot);
        list.swap(padded);
    }
    const uint64_t np = list.size();
    auto at = [&](uint64_t i) { return slot.empty() ? i : slot[i]; };
    std::vector<double> wsorted, tsorted;
    if (call.h_weights) {
        wsorted.assign(np, 0.0);
        for (uint64_t i = 0; i < n; i++) wsorted[at(i)] = call.h_weights[order ? perm[i] : i];
    }
    if (call.h_totals) tsorted.assign(np, 0.0);
    uint32_t *d_list = nullptr;
    double *d_w = nullptr, *d_acc = nullptr;
    int rc = zg_dmalloc(&d_list, np);
    if (rc == ZG_OK && np)
        rc = zg_hip_code(hipMemcpyAsync(d_list, list.data(), sizeof(uint32_t) * np, hipMemcpyHostToDevice, s));
    if (rc == ZG_OK && call.h_weights) {
        rc = zg_dmalloc(&d_w, np);
        if (rc == ZG_OK && np)
            rc = zg_hip_code(hipMemcpyAsync(d_w, wsorted.data(), sizeof(double) * np, hipMemcpyHostToDevice, s));
    }
    double *d_out = call.d_out;
    if (rc == ZG_OK && !d_out) {  // totals only: the scores go to scratch
        rc = zg_dmalloc(&d_acc, V);
        d_out = d_acc;
    }
    if (rc == ZG_OK) rc = zg_hip_code(hipStreamSynchronize(s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == ZG_OK) rc = zg_hip_code(hipEventCreate(&e0));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventCreate(&e1));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventRecord(e0, s));
    if (queue_engine) {
        if (rc == ZG_OK)
            rc = zg_bcq_sweep(g, d_list, n, d_w, call.h_totals ? tsorted.data() : nullptr, d_out, s, &inf);
    } else {
        if (rc == ZG_OK) rc = bc_ws_ensure(eg, false, list.data(), np);
        BcSweepOpt opt;
        opt.d_weights = d_w;
        opt.h_totals = call.h_totals ? tsorted.data() : nullptr;
        opt.fold = call.fold;
        if (rc == ZG_OK) rc = bc_sweep(eg, d_list, list.data(), np, s, &inf, opt);
        inf.n_sources = n;
        if (rc == ZG_OK && V) {
            const size_t zp = zg_prof_begin(eg, s);
            hipLaunchKernelGGL(bc_reduce_kernel, dim3(cdiv(V, 256)), dim3(256), 0, s, V, (uint32_t)eg->bc.groups,
                               eg->bc.bcg, g->live, eg != g ? g->lane_perm : nullptr, d_out);
            zg_prof_end(eg, s, ZG_PROF_BC_SETUP, zp);
            rc = zg_hip_code(hipGetLastError());
        }
    }
    if (rc == ZG_OK) rc = zg_hip_code(hipEventRecord(e1, s));
    if (rc == ZG_OK) rc = zg_hip_code(hipEventSynchronize(e1));
    if (rc == ZG_OK) {
        float ms = 0.f;
        rc = zg_hip_code(hipEventElapsedTime(&ms, e0, e1));
        inf.kernel_ms = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    if (rc == ZG_OK) rc = zg_prof_collect(g, s);
    if (eg != g) {
        const int prc = zg_prof_collect(eg, s);
        if (rc == ZG_OK) rc = prc;
        for (int k = 0; k < ZG_PROF_SLOTS; k++) {
            g->prof.ms[k] += eg->prof.ms[k];
            g->prof.launches[k] += eg->prof.launches[k];
        }
        memset(&eg->prof, 0, sizeof(eg->prof));
    }
    zg_dfree(d_list);
    zg_dfree(d_w);
    zg_dfree(d_acc);
    if (rc == ZG_OK && call.h_totals)
        for (uint64_t i = 0; i < n; i++) call.h_totals[order ? perm[i] : i] = tsorted[at(i)];
    if (rc == ZG_OK && info) *info = inf;
    return rc;
}

extern "C" int zg_bc_unweighted_device(zg_graph *g, const uint32_t *d_sources, uint64_t n_sources,
                                       double *d_out_bc, void *stream, zg_bc_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!d_out_bc && g->V) return ZG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    BcCall call;
    call.d_sources = d_sources;
    call.n_sources = n_sources;
    call.d_out = d_out_bc;
    return zg_bc_device_impl(g, call, s, info);
}

// Host-buffer form shared by zg_bc_unweighted and zg_bc_weighted.
static int bc_host_call(zg_graph *g, const uint32_t *sources, const double *weights, uint64_t n_sources,
                        double *out_bc, double *out_totals, zg_bc_info *info, bool fold = false)
{
    if (sources)
        for (uint64_t i = 0; i < n_sources; i++)
            if (sources[i] >= g->V || !g->h_live[sources[i]]) return ZG_ERR_NODE_NOT_FOUND;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = g->stream;
    uint32_t *d_src = nullptr;
    double *d_out = nullptr;
    int rc = zg_io_out(g, &d_out);  // totals only: the scores still land here
    if (rc == ZG_OK && sources) {
        rc = zg_io_src(g, n_sources, &d_src);
        if (rc == ZG_OK && n_sources)
            rc = zg_hip_code(hipMemcpyAsync(d_src, sources, sizeof(uint32_t) * n_sources, hipMemcpyHostToDevice, s));
    }
    BcCall call;
    call.d_sources = sources ? d_src : nullptr;
    call.n_sources = sources ? n_sources : 0;
    call.h_weights = weights;
    call.h_totals = out_totals;
    call.d_out = d_out;
    call.fold = fold;
    if (rc == ZG_OK) rc = zg_bc_device_impl(g, call, s, info);
    if (rc == ZG_OK && out_bc && g->V)
        rc = zg_hip_code(hipMemcpy(out_bc, d_out, sizeof(double) * g->V, hipMemcpyDeviceToHost));
    return rc;
}

extern "C" int zg_bc_unweighted(zg_graph *g, const uint32_t *sources, uint64_t n_sources, double *out_bc,
                                zg_bc_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out_bc && g->V) return ZG_ERR_INVALID_ARG;
    return bc_host_call(g, sources, nullptr, n_sources, out_bc, nullptr, info);
}

extern "C" int zg_bc_weighted(zg_graph *g, const uint32_t *sources, const double *weights, uint64_t n_sources,
                              double *out_bc, double *out_totals, zg_bc_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out_bc && !out_totals) return ZG_ERR_INVALID_ARG;
    if (weights && !sources) return ZG_ERR_INVALID_ARG;
    if (weights)
        for (uint64_t i = 0; i < n_sources; i++)
            if (!std::isfinite(weights[i])) return ZG_ERR_INVALID_ARG;
    return bc_host_call(g, sources, weights, n_sources, out_bc, out_totals, info);
}

extern "C" int zg_bc_folded(zg_graph *g, const uint32_t *sources, const double *weights, uint64_t n_sources,
                            double *out_bc, double *out_totals, zg_bc_info *info)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!out_bc && !out_totals) return ZG_ERR_INVALID_ARG;
    if (!sources && n_sources) return ZG_ERR_INVALID_ARG;
    if (weights)
        for (uint64_t i = 0; i < n_sources; i++)
            if (!std::isfinite(weights[i])) return ZG_ERR_INVALID_ARG;
    if (!sources) {
        // no sources: an empty sum
        if (out_bc && g->V) memset(out_bc, 0, sizeof(double) * g->V);
        if (info) memset(info, 0, sizeof(*info));
        return ZG_OK;
    }
    return bc_host_call(g, sources, weights, n_sources, out_bc, out_totals, info, true);
}

struct PathsCtx {
    const uint32_t *sources;
    uint32_t *depth_out;
    double *sigma_out;
    int32_t *shift_out;
    uint32_t max_levels;
    uint32_t *levels_out;
    uint32_t *d_dep;
    double *d_sig;
    std::vector<int> h_emax;
};

static int bc_paths_hook(zg_graph *g, hipStream_t s, uint64_t first, uint32_t nb, uint32_t G, void *user)
{
    PathsCtx *c = (PathsCtx *)user;
    const uint32_t V = g->V;
    zg_bc_ws &w = g->bc;
    const uint32_t lv = w.last_levels;  // levels of this batch + 1
    c->h_emax.resize((size_t)lv * w.groups * 64);
    ZG_TRY(hipMemcpyAsync(c->h_emax.data(), w.emax, sizeof(int) * c->h_emax.size(), hipMemcpyDeviceToHost, s));
    for (uint32_t q = 0; q < G; q++) {
        const uint32_t nl = std::min<uint32_t>(64, nb - 64 * q);
        if (V) {
            uint32_t vlo, vhi;
            bc_group_span(g, c->sources + first + 64ull * q, nl, &vlo, &vhi);
            hipLaunchKernelGGL(bc_export_kernel, dim3(cdiv(V, 256), 64), dim3(256), 0, s, V, w.vloc, vlo, vhi, q,
                               w.dep, w.sig, c->d_dep, c->d_sig);
            ZG_TRY(hipGetLastError());
        }
        const uint64_t i0 = first + 64ull * q;
        ZG_TRY(hipMemcpyAsync(c->depth_out + i0 * V, c->d_dep, sizeof(uint32_t) * (size_t)nl * V,
                              hipMemcpyDeviceToHost, s));
        ZG_TRY(hipMemcpyAsync(c->sigma_out + i0 * V, c->d_sig, sizeof(double) * (size_t)nl * V,
                              hipMemcpyDeviceToHost, s));
        ZG_TRY(hipStreamSynchronize(s));
        for (uint32_t l = 0; l < nl; l++) {
            const uint32_t *dp = c->depth_out + (i0 + l) * V;
            uint32_t mx = 0;
            for (uint32_t v = 0; v < V; v++)
                if (dp[v] != ZG_NONE && dp[v] + 1 > mx) mx = dp[v] + 1;
            c->levels_out[i0 + l] = mx;
            if (c->shift_out) {
                if (mx > c->max_levels) return ZG_ERR_INVALID_ARG;
                for (uint32_t L = 0; L < mx; L++)
                    c->shift_out[(i0 + l) * c->max_levels + L] =
                        L < lv ? c->h_emax[((size_t)L * w.groups + q) * 64 + l] : 0;
            }
        }
    }
    return ZG_OK;
}

extern "C" int zg_bc_paths(zg_graph *g, const uint32_t *sources, uint64_t n_sources, uint32_t *depth_out,
                           double *sigma_out, int32_t *shift_out, uint32_t max_levels, uint32_t *levels_out)
{
    if (!g) return ZG_ERR_INVALID_HANDLE;
    if (!sources || !depth_out || !sigma_out || !levels_out) return ZG_ERR_INVALID_ARG;
    for (uint64_t i = 0; i < n_sources; i++)
        if (sources[i] >= g->V || !g->h_live[sources[i]]) return ZG_ERR_NODE_NOT_FOUND;
    std::lock_guard<std::mutex> lk(g->mu);
    ZG_TRY(hipSetDevice(g->device));
    hipStream_t s = g->stream;
    ZG_TRY_RC(bc_ws_ensure(g, true, sources, n_sources));
    PathsCtx c;
    c.sources = sources;
    c.depth_out = depth_out;
    c.sigma_out = sigma_out;
    c.shift_out = shift_out;
    c.max_levels = max_levels;
    c.levels_out = levels_out;
    c.d_dep = nullptr;
    c.d_sig = nullptr;
    uint32_t *d_src = nullptr;
    int rc = zg_dmalloc(&c.d_dep, (size_t)64 * g->V);
    if (rc == ZG_OK) rc = zg_dmalloc(&c.d_sig, (size_t)64 * g->V);
    if (rc == ZG_OK) rc = zg_dmalloc(&d_src, n_sources);
    if (rc == ZG_OK && n_sources)
        rc = zg_hip_code(hipMemcpyAsync(d_src, sources, sizeof(uint32_t) * n_sources, hipMemcpyHostToDevice, s));
    zg_bc_info inf;
    memset(&inf, 0, sizeof(inf));
    BcSweepOpt opt;
    opt.hook = bc_paths_hook;
    opt.user = &c;
    if (rc == ZG_OK) rc = bc_sweep(g, d_src, sources, n_sources, s, &inf, opt);
    (void)hipStreamSynchronize(s);
    zg_dfree(c.d_dep);
    zg_dfree(c.d_sig);
    zg_dfree(d_src);
    return rc;
}
