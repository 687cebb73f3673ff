// Weighted Brandes betweenness (betweennessF64) on the gfx950.
//
// Reference: priv/zog/src/centrality.zig:529-667 with betweennessF64
// (:1355-1357): f64 arc weights, distances added in path order (a + b), a
// shortest path is one whose sum equals the node's distance exactly (`.eq`),
// path counts sigma add over every such arc (parallel arcs separately), and
// delta(u) += (sigma_u / sigma_v)(1 + delta_v) over the shortest-path DAG in
// reverse order; bc(v) += delta(v) for v != s, unnormalised.
//
// GPU form, per source (up to 2,048 sources per batch, each with its own
// rows; one wavefront per 64 entries of one source's list, a vertex with
// more than ZG_W_WIDE arcs walked by the whole wavefront):
//   1. distances by frontier relaxation: every vertex whose distance dropped
//      relaxes its arcs with a 64-bit atomic min on the distance's bit
//      pattern (non-negative doubles order as unsigned integers); a vertex
//      is queued at most once per round.  A vertex whose new distance is
//      above the source's bound waits in a far pile instead; when a source's
//      frontier runs dry the bound rises by delta (or to the least far
//      distance) and the far entries under it form the next frontier (the
//      near-far form of delta-stepping), which saves relaxing a vertex again
//      and again on the way down to its distance.  The far pile has two
//      tiers: a window below the bound + ZG_W_WINDOW steps, split at every
//      raise, and an overflow pile scanned only when the window runs dry or
//      the bound passes its end.  It ends at the least fixed point d(w) =
//      min over arcs of fl(d(v) + w(v,w)), the distances Dijkstra computes
//      with the same additions.  Every arc whose sum was at most its head's
//      distance when relaxed is marked a candidate tight arc.
//   2. the tight arcs (fl(d(v) + w) == d(w) bit for bit, v != w, w not the
//      source; with positive weights neither exception is ever tight) form the
//      shortest-path DAG: the candidates are checked against the final
//      distances and their in-arcs counted, then the DAG is peeled in rounds
//      (a vertex enters the round after its last tight in-arc was
//      processed), adding sigma(v) into every tight successor.
//   3. rounds in reverse: delta(v) = sigma(v) * sum over tight successors of
//      c(w), c(v) = (1 + delta(v)) / sigma(v); bc(v) += delta(v) (FP64
//      atomics across sources).
// Arc weights must be positive for the reference's semantics; a zero-weight
// cycle would never be peeled (its vertices get no sigma).  Everything one
// step touches for a (source, vertex) pair sits in one 32-byte record
// (zg_wrec), and the per-source kernels walk their chunks XCD by XCD
// (zg_q_walk).
#include "common.h"
#include "queue_map.h"

#include <algorithm>

#define ZG_W_MAX_SOURCES 2048  // sources per batch (the range tables are sized to the batch in LDS)
#define ZG_W_GRID 4096  // default workgroups of the per-round kernels ...
#define ZG_W_GRID_LARGE 12288  // ... and of batches with sources x vertices >= ZG_W_LARGE_SV
#define ZG_W_LARGE_SV ((size_t)1 << 28)
#define ZG_W_GRID_DAG 32768  // ... and of their tight-arc count and sigma peel
#define ZG_W_ROUNDS_PER_SYNC 16
#define ZG_W_WIDE 32  // vertices with more arcs are walked by a whole wavefront
#define ZG_W_HUB 4096  // ... and with more than this, cut into pieces for many wavefronts
#define ZG_W_PIECE 2048
#define ZG_W_HUB_CAP (1u << 20)  // hub pieces per relax round
#define ZG_W_UNSET 0xFFFFFFFFFFFFFFFFull
#define ZG_W_DONE 0xFFFFFFFEu  // aux of a record whose dependency pass is done (dist holds the dependency)
#define ZG_W_WINDOW 8  // the window pile spans this many bound steps

// far pile p, buffer b, source s
#define FARROW(p, b, s) (((size_t)((p) * 2 + (b)) * a.S + (s)) * a.V)
#define FARCNT(p, b, s) (((size_t)((p) * 2 + (b)) * a.S + (s)))

struct WbcDev {
    uint32_t V, S, lcap;
    const uint32_t *out_ptr, *out_col;
    const double *out_w;
    zg_wrec *rec;               // [S][V] distance bits (ZG_W_UNSET = unreached), sigma, c, the relax
                                // round that queued the vertex last, far-pile flag (2 for every
                                // reached vertex once the relax rounds end; then 2 + its tight
                                // in-arcs not yet processed)
    uint32_t *fr;               // [2][S][V] relax frontiers
    uint32_t *fcnt;             // [2][S]
    uint32_t *reach;            // [S][V] reached vertices
    uint32_t *rcnt;             // [S]
    uint32_t *order;            // [S][V] DAG order, round by round
    uint32_t *tbits;            // [S][tw] tight-arc bits in successor-CSR order
    size_t tw;                  // words per source
    uint32_t *qlen;             // [S]
    uint32_t *lstart;           // [lcap][S] round starts in order[]
    uint32_t *tot;              // [lcap] frontier sizes (relax) / entries added (peel)
    const uint32_t *src;        // [S]
    double *bc;                 // [V]
    unsigned long long *stats;  // [3] arcs, reached, arcs relaxed
    // near-far split of the relax frontier: a vertex whose new distance is
    // above its source's bound waits in a far pile until the bound passes it;
    // pile 0 (the window) holds distances below wlim, pile 1 the rest
    uint32_t *far;              // [2 piles][2 buffers][S][V] (fsel[p][s] is the live buffer)
    uint32_t *farcnt;           // [2][2][S]
    // far-pile flag in rec: 0 never far, 1 in the window, 3 in pile 1, 2 within the bound
    uint32_t *fsel;             // [2][S]
    uint32_t *split;            // [2][S] this round splits the window / scans pile 1
    unsigned long long *thr;    // [S] the bound (distance bits)
    unsigned long long *wlim;   // [S] the window's end (distance bits)
    unsigned long long *fmin;   // [2][S] least distance recorded for a pile's entry (bits)
    double delta;               // bound step
    uint32_t *hcnt;             // [1] hub pieces listed this round
    uint32_t *hub_v, *hub_a;    // [hcap] hub vertex; source << 21 | piece
    uint32_t hcap;
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

// Every record unreached: distance all ones, no paths, never queued, never far.
__global__ void wbc_fill_kernel(zg_wrec *rec, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        rec[i] = zg_wrec{ZG_W_UNSET, 0.0, 0.0, 0xFFFFFFFFu, 0u};
}

__global__ void wbc_init_kernel(WbcDev a)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.S) return;
    a.fcnt[a.S + s] = 0;
    a.lstart[s] = 0;
    for (int p = 0; p < 2; p++) {
        a.farcnt[FARCNT(p, 0, s)] = 0;
        a.farcnt[FARCNT(p, 1, s)] = 0;
        a.fsel[p * a.S + s] = 0;
        a.split[p * a.S + s] = 0;
        a.fmin[p * a.S + s] = ZG_W_UNSET;
    }
    a.thr[s] = zg_f64_to_bits(a.delta);
    a.wlim[s] = zg_f64_to_bits(a.delta * (1 + ZG_W_WINDOW));
    const uint32_t v = a.src[s];
    if (v == ZG_NONE) {
        a.fcnt[s] = 0;
        a.rcnt[s] = 0;
        a.qlen[s] = 0;
        a.lstart[a.S + s] = 0;
        return;
    }
    a.rec[WROW(s) + v] = zg_wrec{0ull, 1.0, 0.0, 0u, 2u};  // distance +0.0, queued for round 0
    a.fr[WROW(s)] = v;
    a.fcnt[s] = 1;
    a.reach[WROW(s)] = v;
    a.rcnt[s] = 1;
    a.order[WROW(s)] = v;
    a.qlen[s] = 1;
    a.lstart[a.S + s] = 1;
}

// Before relax round r, per source: the work left (frontier and far piles,
// summed into tot[r]), an empty buffer for the next frontier, and, for a
// source whose frontier is empty while a far pile is not, a raised bound
// max(bound + delta, least window distance) -- or, when the window is empty,
// max(bound + delta, least pile-1 distance) -- with a fresh window buffer for
// wbc_split_kernel; pile 1 is scanned too (and the window moved on by
// ZG_W_WINDOW steps past the bound) when the window is empty or the bound
// passes its end.  One block.
__global__ void wbc_round_kernel(WbcDev a, int r)
{
    __shared__ unsigned long long red[16];
    const int cur = r & 1, nxt = cur ^ 1;
    unsigned long long t = 0;
    if (threadIdx.x == 0) *a.hcnt = 0;
    for (uint32_t s = threadIdx.x; s < a.S; s += blockDim.x) {
        const uint32_t nf = a.fcnt[(size_t)cur * a.S + s];
        const uint32_t fw = a.fsel[s], fo = a.fsel[a.S + s];
        const uint32_t nw = a.farcnt[FARCNT(0, fw, s)], no = a.farcnt[FARCNT(1, fo, s)];
        t += (unsigned long long)nf + nw + no;
        a.fcnt[(size_t)nxt * a.S + s] = 0;
        bool sw = false, so = false;
        if (nf == 0 && (nw > 0 || no > 0)) {
            double T = zg_bits_to_f64(a.thr[s]) + a.delta;
            const unsigned long long fmw = a.fmin[s], fmo = a.fmin[a.S + s];
            if (nw > 0 && fmw != ZG_W_UNSET) T = fmax(T, zg_bits_to_f64(fmw));
            if (nw == 0 && fmo != ZG_W_UNSET) T = fmax(T, zg_bits_to_f64(fmo));
            a.thr[s] = zg_f64_to_bits(T);
            sw = nw > 0;
            so = no > 0 && (nw == 0 || zg_f64_to_bits(T) >= a.wlim[s]);
            if (so) {
                a.wlim[s] = zg_f64_to_bits(T + a.delta * ZG_W_WINDOW);
                a.fmin[a.S + s] = ZG_W_UNSET;
                a.fsel[a.S + s] = fo ^ 1;
                a.farcnt[FARCNT(1, fo ^ 1, s)] = 0;
            }
            a.fmin[s] = ZG_W_UNSET;
            a.fsel[s] = fw ^ 1;
            a.farcnt[FARCNT(0, fw ^ 1, s)] = 0;
        }
        a.split[s] = sw;
        a.split[a.S + s] = so;
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

// Least of a lane value over the wavefront (64-bit).
__device__ __forceinline__ unsigned long long zg_wave_min_u64(unsigned long long x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long y = __shfl_xor(x, off, 64);
        x = y < x ? y : x;
    }
    return x;
}

// Splits far pile p of the sources flagged by wbc_round_kernel.  Window
// (p = 0): entries still in it and within the raised bound join round r's
// frontier, the others stay.  Pile 1: entries within the bound join the
// frontier, those below the window's new end move to the window, the others
// stay.  Entries whose vertex has left the pile (come within the bound, or
// moved to the window) are dropped; each pile's least remaining distance is
// recorded.
__global__ __launch_bounds__(256) void wbc_split_kernel(WbcDev a, int r, int p)
{
    extern __shared__ uint32_t wbc_lds[];  // per-source range tables, 3 S + 1 words
    uint32_t *s_lo = wbc_lds, *s_hi = wbc_lds + a.S, *s_pre = wbc_lds + 2 * a.S;
    const int cur = r & 1;
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = 0;
            hi = a.split[(size_t)p * a.S + s] ? a.farcnt[FARCNT(p, a.fsel[(size_t)p * a.S + s] ^ 1, s)] : 0;
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63;
    const uint32_t mine = p == 0 ? 1u : 3u;
    for (ZgQWalk wk = zg_q_walk(total); wk.c < wk.end; wk.c += wk.stride) {
        const uint32_t c = wk.c;
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        const uint32_t fs = a.fsel[(size_t)p * a.S + s], fw = a.fsel[s];
        const unsigned long long Tb = a.thr[s], WL = a.wlim[s];
        zg_wrec *rec = a.rec + WROW(s);
        const uint32_t v = valid ? a.far[FARROW(p, fs ^ 1, s) + idx] : 0;
        bool move = false, to_window = false, keep = false;
        unsigned long long d = ZG_W_UNSET;
        if (valid && rec[v].flag == mine) {
            d = rec[v].dist;
            if (d <= Tb) move = atomicExch(&rec[v].flag, 2u) == mine;
            else if (p == 1 && d < WL) to_window = atomicCAS(&rec[v].flag, 3u, 1u) == 3u;
            else keep = true;
        }
        wbc_append(move, v, &a.fcnt[(size_t)cur * a.S + s], a.fr + ((size_t)cur * a.S + s) * a.V, lane);
        wbc_append(keep, v, &a.farcnt[FARCNT(p, fs, s)], a.far + FARROW(p, fs, s), lane);
        if (p == 1) wbc_append(to_window, v, &a.farcnt[FARCNT(0, fw, s)], a.far + FARROW(0, fw, s), lane);
        const unsigned long long mk = zg_wave_min_u64(keep ? d : ZG_W_UNSET);
        if (lane == 0 && mk != ZG_W_UNSET) atomicMin(&a.fmin[(size_t)p * a.S + s], mk);
        if (p == 1) {
            const unsigned long long mw = zg_wave_min_u64(to_window ? d : ZG_W_UNSET);
            if (lane == 0 && mw != ZG_W_UNSET) atomicMin(&a.fmin[s], mw);
        }
    }
}

__device__ __forceinline__ double zg_readlane_f64(double x, int j)
{
    return zg_bits_to_f64(zg_readlane64(zg_f64_to_bits(x), j));
}

// Walks the tight arcs of the arc range [e, e1) through the bit words that
// cover it.
struct WbcTightIt {
    uint32_t e, e1, base, bits;
};

__device__ __forceinline__ bool wbc_next_tight(WbcTightIt &it, const uint32_t *tb, uint32_t &arc)
{
    while (it.bits == 0) {
        if (it.e >= it.e1) return false;
        const uint32_t wi = it.e >> 5;
        const uint64_t wend = ((uint64_t)wi + 1) << 5;
        uint32_t b = tb[wi] & (~0u << (it.e & 31));
        if ((uint64_t)it.e1 < wend) b &= (1u << (it.e1 & 31)) - 1;
        it.base = wi << 5;
        it.bits = b;
        it.e = (uint32_t)std::min<uint64_t>(wend, it.e1);
    }
    arc = it.base + __builtin_ctz(it.bits);
    it.bits &= it.bits - 1;
    return true;
}

// Bit words covering a wide vertex's arcs [ej, ej + dj), and the range of
// word k of them (empty past the last word).
__device__ __forceinline__ uint32_t wbc_words(uint32_t ej, uint32_t dj)
{
    return ((ej + dj - 1) >> 5) - (ej >> 5) + 1;
}

__device__ __forceinline__ WbcTightIt wbc_word_share(uint32_t ej, uint32_t dj, uint32_t k)
{
    const uint32_t w0 = ej >> 5;
    WbcTightIt it{0, 0, 0, 0};
    if (k < wbc_words(ej, dj)) {
        const uint64_t lo = (uint64_t)(w0 + k) << 5;
        it.e = (uint32_t)std::max<uint64_t>(lo, ej);
        it.e1 = (uint32_t)std::min<uint64_t>(lo + 32, (uint64_t)ej + dj);
    }
    return it;
}

// One source's view of relax round r.
struct WbcRelax {
    zg_wrec *rec;
    uint32_t *ncnt, *nfr, *rcnt, *reach, *wcnt, *win, *ocnt, *ovf;
    uint32_t *tb;  // candidate bits: arcs that were tight when relaxed
    unsigned long long Tb, WL;
    uint32_t r1;
    unsigned long long fmw, fmo;  // least window / pile-1 distance this lane recorded

    __device__ __forceinline__ WbcRelax(const WbcDev &a, uint32_t s, int r)
    {
        const int nxt = (r & 1) ^ 1;
        rec = a.rec + WROW(s);
        ncnt = &a.fcnt[(size_t)nxt * a.S + s];
        nfr = a.fr + ((size_t)nxt * a.S + s) * a.V;
        rcnt = &a.rcnt[s];
        reach = a.reach + WROW(s);
        tb = a.tbits + s * a.tw;
        const uint32_t fw = a.fsel[s], fo = a.fsel[a.S + s];
        wcnt = &a.farcnt[FARCNT(0, fw, s)];
        win = a.far + FARROW(0, fw, s);
        ocnt = &a.farcnt[FARCNT(1, fo, s)];
        ovf = a.far + FARROW(1, fo, s);
        Tb = a.thr[s];
        WL = a.wlim[s];
        r1 = (uint32_t)(r + 1);
        fmw = fmo = ZG_W_UNSET;
    }

    // Lowers d(w) to fl(d + weight) for an arc whose head, weight and a read
    // of the head's distance are already loaded (a stale read is never below
    // the current one); reports a first reach, a queueing for the next round
    // (within the bound) and a first entry into a far pile.  Returns whether
    // the arc is a candidate tight arc: fl(d + weight) is at most the head's
    // distance once lowered.  Every arc that is tight at the end is a
    // candidate when its tail is relaxed at its final distance.
    __device__ __forceinline__ bool arc_loaded(double d, uint32_t w, double wt, unsigned long long dw, bool &first,
                                               bool &queue, bool &fwin, bool &fovf)
    {
        const unsigned long long nb = zg_f64_to_bits(d + wt);
        if (nb > dw) return false;
        if (nb == dw) return true;
        const unsigned long long old = atomicMin(&rec[w].dist, nb);
        if (nb < old) {
            first = old == ZG_W_UNSET;
            if (nb <= Tb) {
                queue = atomicExch(&rec[w].aux, r1) != r1;
                rec[w].flag = 2;
            } else if (nb < WL) {
                // into the window from "never far" or from pile 1; retried
                // when another lane moves the flag in between (0 -> 3 by a
                // larger sum that lowered d(w) first), so a vertex below the
                // window's end never stays in pile 1
                uint32_t f = rec[w].flag;
                while (f == 0u || f == 3u) {
                    const uint32_t o = atomicCAS(&rec[w].flag, f, 1u);
                    if (o == f) {
                        fwin = true;
                        break;
                    }
                    f = o;
                }
                fmw = nb < fmw ? nb : fmw;
            } else {
                fovf = atomicCAS(&rec[w].flag, 0u, 3u) == 0u;
                fmo = nb < fmo ? nb : fmo;
            }
        }
        return nb <= old;
    }


    // Marks the candidate arcs of a wavefront step over the arcs [k0, k0+64):
    // lane i's flag is arc k0+i's.  All lanes call it.
    __device__ __forceinline__ void mark_wave(bool cand, uint32_t k0, int lane)
    {
        const uint64_t M = __ballot(cand);
        if (!M) return;
        const uint32_t off = k0 & 31, wb = k0 >> 5;
        const uint64_t lo = M << off, hi = off ? M >> (64 - off) : 0;
        const uint32_t word = lane == 0 ? (uint32_t)lo : lane == 1 ? (uint32_t)(lo >> 32) : (uint32_t)hi;
        if (lane < 3 && word) atomicOr(&tb[wb + lane], word);
    }

    // The appends of one step; every lane of the wavefront calls it.
    __device__ __forceinline__ void append(bool first, bool queue, bool fwin, bool fovf, uint32_t w, int lane)
    {
        wbc_append(queue, w, ncnt, nfr, lane);
        wbc_append(fwin, w, wcnt, win, lane);
        wbc_append(fovf, w, ocnt, ovf, lane);
        wbc_append(first, w, rcnt, reach, lane);
    }

    // The whole wavefront over the arcs [e0, e1) of a vertex at distance d,
    // 256 arcs a step: each lane loads four heads and weights, then their
    // distances, so four loads of each kind are in flight per lane.
    __device__ __forceinline__ void wave_range(const WbcDev &a, double d, uint32_t e0, uint32_t e1, int lane)
    {
        for (uint32_t k0 = e0; k0 < e1; k0 += 256) {
            uint32_t w[4];
            double wt[4];
            unsigned long long dw[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t e = k0 + 64 * t + lane;
                w[t] = e < e1 ? a.out_col[e] : 0u;
                wt[t] = e < e1 ? a.out_w[e] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < 4; t++) dw[t] = k0 + 64 * t + lane < e1 ? rec[w[t]].dist : 0ull;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                bool first = false, queue = false, fwin = false, fovf = false, cand = false;
                if (k0 + 64 * t + lane < e1) cand = arc_loaded(d, w[t], wt[t], dw[t], first, queue, fwin, fovf);
                append(first, queue, fwin, fovf, w[t], lane);
                mark_wave(cand, k0 + 64 * t, lane);
            }
        }
    }

    // Folds the lanes' least far distance into the source's; all lanes call it.
    __device__ __forceinline__ void flush(const WbcDev &a, uint32_t s, int lane)
    {
        const unsigned long long mw = zg_wave_min_u64(fmw), mo = zg_wave_min_u64(fmo);
        if (lane == 0 && mw != ZG_W_UNSET) atomicMin(&a.fmin[s], mw);
        if (lane == 0 && mo != ZG_W_UNSET) atomicMin(&a.fmin[a.S + s], mo);
    }
};

// Relax round r over every source's frontier.  Vertices with more than
// ZG_W_WIDE arcs are walked by the whole wavefront, one at a time, and those
// with more than ZG_W_HUB arcs are cut into pieces of ZG_W_PIECE arcs for
// wbc_hub_kernel (inline when its list is full); the others are walked by
// their own lane.
__global__ __launch_bounds__(256) void wbc_relax_kernel(WbcDev a, int r)
{
    extern __shared__ uint32_t wbc_lds[];  // per-source range tables, 3 S + 1 words
    uint32_t *s_lo = wbc_lds, *s_hi = wbc_lds + a.S, *s_pre = wbc_lds + 2 * a.S;
    const int cur = r & 1;
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = 0;
            hi = a.fcnt[(size_t)cur * a.S + s];
        },
        s_lo, s_hi, s_pre);
    __shared__ unsigned long long red[4];
    const int lane = threadIdx.x & 63;
    unsigned long long scanned = 0;
    for (ZgQWalk wk = zg_q_walk(total); wk.c < wk.end; wk.c += wk.stride) {
        const uint32_t c = wk.c;
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        WbcRelax R(a, s, r);
        const uint32_t v = valid ? a.fr[((size_t)cur * a.S + s) * a.V + idx] : 0;
        const double dv = valid ? zg_bits_to_f64(R.rec[v].dist) : 0.0;  // latest value
        const uint32_t e0 = valid ? a.out_ptr[v] : 0, deg = valid ? a.out_ptr[v + 1] - e0 : 0;
        scanned += deg;
        for (uint64_t wide = __ballot(deg > ZG_W_WIDE); wide; wide &= wide - 1) {
            const int j = __builtin_ctzll(wide);
            const uint32_t vj = zg_readlane32(v, j), ej = zg_readlane32(e0, j), dj = zg_readlane32(deg, j);
            if (dj > ZG_W_HUB) {
                const uint32_t pieces = (dj + ZG_W_PIECE - 1) / ZG_W_PIECE;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(a.hcnt, pieces);
                base = zg_readlane32(base, 0);
                const bool fits = (uint64_t)base + pieces <= a.hcap;
                for (uint32_t k = lane; k < pieces && base + k < a.hcap; k += 64) {
                    a.hub_v[base + k] = fits ? vj : ZG_NONE;  // claimed slots of a list that is full stay empty
                    a.hub_a[base + k] = (s << 21) | k;  // s < 2^11, k < E / ZG_W_PIECE <= 2^21
                }
                if (fits) continue;
            }
            R.wave_range(a, zg_readlane_f64(dv, j), ej, ej + dj, lane);
        }
        const uint32_t mine = deg <= ZG_W_WIDE ? deg : 0;
        uint32_t maxdeg = mine;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) maxdeg = max(maxdeg, (uint32_t)__shfl_xor((int)maxdeg, off, 64));
        for (uint32_t k0 = 0; k0 < maxdeg; k0 += 4) {
            // heads, weights and distances of four arcs first, so the loads overlap
            uint32_t w[4];
            double wt[4];
            unsigned long long dw[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                w[t] = k0 + t < mine ? a.out_col[e0 + k0 + t] : 0u;
                wt[t] = k0 + t < mine ? a.out_w[e0 + k0 + t] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < 4; t++) dw[t] = k0 + t < mine ? R.rec[w[t]].dist : 0ull;
            // candidate bits of the four arcs: the word of arc e0+k0 and the next
            const uint32_t eb = e0 + k0;
            uint32_t m0 = 0, m1 = 0;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                bool first = false, queue = false, fwin = false, fovf = false;
                if (k0 + t < mine && R.arc_loaded(dv, w[t], wt[t], dw[t], first, queue, fwin, fovf)) {
                    const uint32_t e = eb + t;
                    if ((e >> 5) == (eb >> 5)) m0 |= 1u << (e & 31);
                    else m1 |= 1u << (e & 31);
                }
                R.append(first, queue, fwin, fovf, w[t], lane);
            }
            if (m0) atomicOr(&R.tb[eb >> 5], m0);
            if (m1) atomicOr(&R.tb[(eb >> 5) + 1], m1);
        }
        R.flush(a, s, lane);
    }
    scanned = zg_wave_sum_u64(scanned);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = scanned;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = red[0] + red[1] + red[2] + red[3];
        if (t) atomicAdd(&a.stats[2], t);
    }
}

// The hub pieces of relax round r: one wavefront per piece.
__global__ __launch_bounds__(256) void wbc_hub_kernel(WbcDev a, int r)
{
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t H = min(*a.hcnt, a.hcap);
    for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < H; i += waves) {
        const uint32_t v = a.hub_v[i];
        if (v == ZG_NONE) continue;
        const uint32_t item = a.hub_a[i], s = item >> 21, piece = item & ((1u << 21) - 1);
        WbcRelax R(a, s, r);
        const uint32_t b1 = a.out_ptr[v + 1];
        const uint32_t e0 = a.out_ptr[v] + piece * ZG_W_PIECE, e1 = min(b1, e0 + ZG_W_PIECE);
        R.wave_range(a, zg_bits_to_f64(R.rec[v].dist), e0, e1, lane);
        R.flush(a, s, lane);
    }
}

// Tight arcs among the candidates the relaxation marked, over every reached
// vertex: a candidate that is not tight at the final distances (v != w, w not
// the source, fl(d(v) + weight) == d(w) bit for bit) loses its bit; a tight
// one counts into its head's tight in-degree.  Also the sweep statistics.
__global__ __launch_bounds__(256) void wbc_count_kernel(WbcDev a)
{
    extern __shared__ uint32_t wbc_lds[];  // per-source range tables, 3 S + 1 words
    uint32_t *s_lo = wbc_lds, *s_hi = wbc_lds + a.S, *s_pre = wbc_lds + 2 * a.S;
    __shared__ unsigned long long red[2][4];
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = 0;
            hi = a.rcnt[s];
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long arcs = 0, reached = 0;
    for (ZgQWalk wk = zg_q_walk(total); wk.c < wk.end; wk.c += wk.stride) {
        const uint32_t c = wk.c;
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        zg_wrec *rec = a.rec + WROW(s);
        uint32_t *tb = a.tbits + s * a.tw;
        const uint32_t v = valid ? a.reach[WROW(s) + idx] : 0, src = a.src[s];
        const double dv = valid ? zg_bits_to_f64(rec[v].dist) : 0.0;
        const uint32_t e0 = valid ? a.out_ptr[v] : 0, deg = valid ? a.out_ptr[v + 1] - e0 : 0;
        arcs += deg;
        reached += valid ? 1 : 0;
        // up to four candidates per step, their loads overlapped
        auto walk = [&](WbcTightIt it, uint32_t vv, double d) {
            for (;;) {
                uint32_t e[4] = {0, 0, 0, 0};
                int n = 0;
                if (wbc_next_tight(it, tb, e[0])) {
                    n = 1;
                    if (wbc_next_tight(it, tb, e[1])) {
                        n = 2;
                        if (wbc_next_tight(it, tb, e[2])) n = wbc_next_tight(it, tb, e[3]) ? 4 : 3;
                    }
                }
                if (!n) break;
                uint32_t w[4];
                double wt[4];
                unsigned long long dw[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    w[t] = t < n ? a.out_col[e[t]] : 0u;
                    wt[t] = t < n ? a.out_w[e[t]] : 0.0;
                }
#pragma unroll
                for (int t = 0; t < 4; t++) dw[t] = t < n ? rec[w[t]].dist : 0ull;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (t < n) {
                        if (w[t] != vv && w[t] != src && zg_f64_to_bits(d + wt[t]) == dw[t])
                            atomicAdd(&rec[w[t]].flag, 1u);
                        else
                            atomicAnd(&tb[e[t] >> 5], ~(1u << (e[t] & 31)));
                    }
                }
            }
        };
        for (uint64_t wide = __ballot(deg > ZG_W_WIDE); wide; wide &= wide - 1) {
            const int j = __builtin_ctzll(wide);
            const uint32_t vj = zg_readlane32(v, j), ej = zg_readlane32(e0, j), dj = zg_readlane32(deg, j);
            const double dvj = zg_readlane_f64(dv, j);
            for (uint32_t k0 = 0, nw = wbc_words(ej, dj); k0 < nw; k0 += 64)
                walk(wbc_word_share(ej, dj, k0 + lane), vj, dvj);
        }
        walk(WbcTightIt{e0, deg <= ZG_W_WIDE ? e0 + deg : e0, 0, 0}, v, dv);
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
    extern __shared__ uint32_t wbc_lds[];  // per-source range tables, 3 S + 1 words
    uint32_t *s_lo = wbc_lds, *s_hi = wbc_lds + a.S, *s_pre = wbc_lds + 2 * a.S;
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = a.lstart[(size_t)r * a.S + s];
            hi = a.lstart[(size_t)(r + 1) * a.S + s];
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63;
    for (ZgQWalk wk = zg_q_walk(total); wk.c < wk.end; wk.c += wk.stride) {
        const uint32_t c = wk.c;
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        zg_wrec *rec = a.rec + WROW(s);
        uint32_t *qlen = &a.qlen[s], *order = a.order + WROW(s);
        const uint32_t *tb = a.tbits + s * a.tw;
        const uint32_t v = valid ? order[idx] : 0;
        const double sv = valid ? rec[v].sig : 0.0;
        const uint32_t e0 = valid ? a.out_ptr[v] : 0, deg = valid ? a.out_ptr[v + 1] - e0 : 0;
        // one tight arc per lane and step; every lane takes part in the append
        auto walk = [&](WbcTightIt it, double sg) {
            for (;;) {
                uint32_t e = 0;
                const bool has = wbc_next_tight(it, tb, e);
                if (!__ballot(has)) break;
                bool ready = false;
                uint32_t w = 0;
                if (has) {
                    w = a.out_col[e];
                    atomicAdd(&rec[w].sig, sg);
                    ready = atomicSub(&rec[w].flag, 1u) == 3u;  // its last tight in-arc
                }
                wbc_append(ready, w, qlen, order, lane);
            }
        };
        for (uint64_t wide = __ballot(deg > ZG_W_WIDE); wide; wide &= wide - 1) {
            const int j = __builtin_ctzll(wide);
            const uint32_t ej = zg_readlane32(e0, j), dj = zg_readlane32(deg, j);
            const double svj = zg_readlane_f64(sv, j);
            for (uint32_t k0 = 0, nw = wbc_words(ej, dj); k0 < nw; k0 += 64)
                walk(wbc_word_share(ej, dj, k0 + lane), svj);
        }
        walk(WbcTightIt{e0, deg <= ZG_W_WIDE ? e0 + deg : e0, 0, 0}, sv);
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
    extern __shared__ uint32_t wbc_lds[];  // per-source range tables, 3 S + 1 words
    uint32_t *s_lo = wbc_lds, *s_hi = wbc_lds + a.S, *s_pre = wbc_lds + 2 * a.S;
    const uint32_t total = zg_q_chunks(
        a.S,
        [&](uint32_t s, uint32_t &lo, uint32_t &hi) {
            lo = a.lstart[(size_t)r * a.S + s];
            hi = a.lstart[(size_t)(r + 1) * a.S + s];
        },
        s_lo, s_hi, s_pre);
    const int lane = threadIdx.x & 63;
    for (ZgQWalk wk = zg_q_walk(total); wk.c < wk.end; wk.c += wk.stride) {
        const uint32_t c = wk.c;
        const uint32_t s = zg_q_owner(s_pre, a.S, c);
        const uint32_t idx = s_lo[s] + (c - s_pre[s]) * 64 + lane;
        const bool valid = idx < s_hi[s];
        zg_wrec *rec = a.rec + WROW(s);
        const uint32_t *tb = a.tbits + s * a.tw;
        const uint32_t v = valid ? a.order[WROW(s) + idx] : 0, src = a.src[s];
        const uint32_t e0 = valid ? a.out_ptr[v] : 0, deg = valid ? a.out_ptr[v + 1] - e0 : 0;
        double S = 0.0;  // sum of c(w) over v's tight successors
        for (uint64_t wide = __ballot(deg > ZG_W_WIDE); wide; wide &= wide - 1) {
            const int j = __builtin_ctzll(wide);
            const uint32_t ej = zg_readlane32(e0, j), dj = zg_readlane32(deg, j);
            double part = 0.0;
            for (uint32_t k0 = 0, nw = wbc_words(ej, dj); k0 < nw; k0 += 64) {
                WbcTightIt it = wbc_word_share(ej, dj, k0 + lane);
                uint32_t e;
                while (wbc_next_tight(it, tb, e)) part += rec[a.out_col[e]].cc;
            }
            part = zg_wave_sum(part);
            if (lane == j) S = part;
        }
        if (deg <= ZG_W_WIDE) {
            WbcTightIt it{e0, e0 + deg, 0, 0};
            uint32_t e;
            while (wbc_next_tight(it, tb, e)) S += rec[a.out_col[e]].cc;
        }
        if (valid) {
            const double sv = rec[v].sig;
            const double delta = sv * S;
            rec[v].cc = (1.0 + delta) / sv;
            // the distance is not read again in this batch: it now holds the
            // dependency, which wbc_reduce_kernel adds in source order
            rec[v].dist = zg_f64_to_bits(v != src ? delta : 0.0);
            rec[v].aux = ZG_W_DONE;
        }
    }
}

// bc[v] += the batch's dependencies of v, source by source in slot order
// (records the dependency pass finished only).
__global__ void wbc_reduce_kernel(uint32_t V, uint32_t nb, const zg_wrec *rec, double *bc)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double acc = bc[v];
    for (uint32_t s = 0; s < nb; s++) {
        const zg_wrec &r = rec[(size_t)s * V + v];
        if (r.aux == ZG_W_DONE) acc += zg_bits_to_f64(r.dist);
    }
    bc[v] = acc;
}

__global__ void wbc_out_kernel(uint32_t V, const double *bc, const uint8_t *live, double *out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) out[v] = live[v] ? bc[v] : 0.0;
}

int zg_wbc_ws_release(zg_graph *g)
{
    zg_wbc_ws &w = g->wbc;
    zg_dfree(w.rec); zg_dfree(w.fr);
    zg_dfree(w.fcnt); zg_dfree(w.reach); zg_dfree(w.rcnt); zg_dfree(w.order); zg_dfree(w.qlen);
    zg_dfree(w.lstart); zg_dfree(w.tot); zg_dfree(w.src); zg_dfree(w.bc); zg_dfree(w.stats); zg_dfree(w.tbits);
    zg_dfree(w.far); zg_dfree(w.farcnt); zg_dfree(w.fsel); zg_dfree(w.split);
    zg_dfree(w.thr); zg_dfree(w.wlim); zg_dfree(w.fmin); zg_dfree(w.hcnt); zg_dfree(w.hub_v); zg_dfree(w.hub_a);
    w.S = 0;
    w.lcap = 0;
    return ZG_OK;
}

// Tight-arc bit words per source.
static size_t zg_wbc_words(const zg_graph *g)
{
    return (size_t)((g->E + 31) / 32) + 1;
}

static int wbc_ws_ensure(zg_graph *g, uint32_t S)
{
    zg_wbc_ws &w = g->wbc;
    if (w.S == S) return ZG_OK;
    zg_wbc_ws_release(g);
    const size_t V = g->V ? g->V : 1, SV = (size_t)S * V;
    w.lcap = 256;
    int rc = zg_dmalloc(&w.rec, SV);
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
    if (rc == ZG_OK) rc = zg_dmalloc(&w.stats, 3);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.tbits, (size_t)S * zg_wbc_words(g));
    if (rc == ZG_OK) rc = zg_dmalloc(&w.far, 4 * SV);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.farcnt, 4 * (size_t)S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.fsel, 2 * (size_t)S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.split, 2 * (size_t)S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.thr, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.wlim, S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.fmin, 2 * (size_t)S);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.hcnt, 1);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.hub_v, ZG_W_HUB_CAP);
    if (rc == ZG_OK) rc = zg_dmalloc(&w.hub_a, ZG_W_HUB_CAP);
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

// Bound step of the near frontier: the tuned value, else 0.3 x the mean arc
// weight / the mean out-degree (+inf when there is nothing to scale by).
// Smaller steps relax each vertex fewer times but run more rounds, each of
// which scans the far piles of the sources it raises the bound of; on R-MAT
// scale 20 with weights uniform in (0, 1] (2,048 sources per batch) the steps
// 0.0075 and 0.01 (= 0.3 x 0.5 / 16) were the fastest, 0.005 3% and 0.019 6%
// slower, 0.04 25% slower.
static double zg_wbc_delta(const zg_graph *g)
{
    if (g->wbc_delta > 0.0) return g->wbc_delta;
    const double live = (double)std::max<uint32_t>(g->n_live, 1);
    const double d = g->E ? 0.3 * g->wmean / ((double)g->E / live) : 0.0;
    return d > 0.0 ? d : HUGE_VAL;
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
    a.rec = w.rec; a.fr = w.fr;
    a.fcnt = w.fcnt; a.reach = w.reach; a.rcnt = w.rcnt; a.order = w.order; a.qlen = w.qlen;
    a.lstart = w.lstart; a.tot = w.tot; a.src = w.src; a.bc = w.bc; a.stats = w.stats;
    a.tbits = w.tbits;
    a.tw = zg_wbc_words(g);
    a.far = w.far; a.farcnt = w.farcnt; a.fsel = w.fsel; a.split = w.split;
    a.thr = w.thr; a.wlim = w.wlim; a.fmin = w.fmin;
    a.hcnt = w.hcnt; a.hub_v = w.hub_v; a.hub_a = w.hub_a; a.hcap = ZG_W_HUB_CAP;
    a.delta = zg_wbc_delta(g);
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
    // the tuned width (default ZG_W_MAX_SOURCES), and as many as half the
    // free device memory holds (64 bytes per vertex and 4 per 32 arcs of
    // state per source); a larger workspace from an earlier call is kept
    const uint64_t per_src = 64ull * V + 4ull * zg_wbc_words(g);
    const uint32_t cap = g->wbc_batch ? g->wbc_batch : ZG_W_MAX_SOURCES;
    const uint32_t want = (uint32_t)std::min<uint64_t>(cap, (n + 63) / 64 * 64);
    uint32_t S = g->wbc.S;
    if (S < want || S > cap) {
        size_t fr = 0, tot = 0;
        ZG_TRY(hipMemGetInfo(&fr, &tot));
        fr += (size_t)g->wbc.S * per_src;
        const uint64_t fit = (uint64_t)(0.5 * (double)fr) / per_src;
        S = (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(want, fit) / 64 * 64);
    }
    ZG_TRY_RC(wbc_ws_ensure(g, S));
    zg_wbc_ws &w = g->wbc;
    const size_t SV = (size_t)S * V;
    const uint32_t grid = g->wbc_grid ? g->wbc_grid : SV >= ZG_W_LARGE_SV ? ZG_W_GRID_LARGE : ZG_W_GRID;
    // the far-pile split does little per chunk, so every block's O(S) range
    // prefix dominates it: it keeps the smaller grid
    const uint32_t gsplit = std::min<uint32_t>(grid, ZG_W_GRID);
    // the tight-arc count and the sigma peel gain from more workgroups still
    const uint32_t gdag = g->wbc_grid || SV < ZG_W_LARGE_SV ? grid : ZG_W_GRID_DAG;
    const size_t lds = sizeof(uint32_t) * (3 * (size_t)S + 1);  // range tables of the per-source kernels
    ZG_TRY(hipMemsetAsync(w.bc, 0, sizeof(double) * (V ? V : 1), s));
    ZG_TRY(hipMemsetAsync(w.stats, 0, sizeof(unsigned long long) * 3, s));
    std::vector<uint32_t> tot(ZG_W_ROUNDS_PER_SYNC);
    for (uint64_t b0 = 0; b0 < n; b0 += S) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(S, n - b0);
        ZG_TRY(hipMemsetAsync(w.src, 0xFF, sizeof(uint32_t) * S, s));
        ZG_TRY(hipMemcpyAsync(w.src, d_sources + b0, sizeof(uint32_t) * nb, hipMemcpyDeviceToDevice, s));
        ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, wbc_fill_kernel, dim3(4096), dim3(256), 0, s, w.rec, SV);
        ZG_TRY(hipMemsetAsync(w.tbits, 0, sizeof(uint32_t) * S * zg_wbc_words(g), s));
        WbcDev a = wbc_dev(g);
        ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, wbc_init_kernel, dim3((S + 255) / 256), dim3(256), 0, s, a);
        ZG_TRY(hipGetLastError());
        // 1. distances: relax rounds, queued ZG_W_ROUNDS_PER_SYNC at a time
        uint32_t r = 0;
        for (bool done = false; !done;) {
            ZG_TRY_RC(wbc_grow(g, r + ZG_W_ROUNDS_PER_SYNC + 2, r + 1, s));
            a = wbc_dev(g);
            for (uint32_t k = 0; k < ZG_W_ROUNDS_PER_SYNC; k++) {
                ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, wbc_round_kernel, dim3(1), dim3(1024), 0, s, a, (int)(r + k));
                ZG_PROF_LAUNCH(ZG_PROF_BC_FORWARD, wbc_split_kernel, dim3(gsplit), dim3(256), lds, s, a, (int)(r + k), 0);
                ZG_PROF_LAUNCH(ZG_PROF_BC_FORWARD, wbc_split_kernel, dim3(gsplit), dim3(256), lds, s, a, (int)(r + k), 1);
                ZG_PROF_LAUNCH(ZG_PROF_BC_FORWARD, wbc_relax_kernel, dim3(grid), dim3(256), lds, s, a, (int)(r + k));
                ZG_PROF_LAUNCH(ZG_PROF_BC_FORWARD, wbc_hub_kernel, dim3(grid), dim3(256), 0, s, a, (int)(r + k));
            }
            ZG_TRY(hipGetLastError());
            ZG_TRY(hipMemcpyAsync(tot.data(), w.tot + r, sizeof(uint32_t) * ZG_W_ROUNDS_PER_SYNC,
                                  hipMemcpyDeviceToHost, s));
            ZG_TRY(hipStreamSynchronize(s));
            for (uint32_t k = 0; k < ZG_W_ROUNDS_PER_SYNC && !done; k++) done = tot[k] == 0;
            r += ZG_W_ROUNDS_PER_SYNC;
        }
        // 2. the tight DAG: in-arc counts, then peel rounds
        ZG_PROF_LAUNCH(ZG_PROF_BC_DAG, wbc_count_kernel, dim3(gdag), dim3(256), lds, s, a);
        ZG_TRY(hipGetLastError());
        int last = -1;
        for (uint32_t p = 0; last < 0;) {
            ZG_TRY_RC(wbc_grow(g, p + ZG_W_ROUNDS_PER_SYNC + 2, p + 2, s));
            a = wbc_dev(g);
            for (uint32_t k = 0; k < ZG_W_ROUNDS_PER_SYNC; k++) {
                ZG_PROF_LAUNCH(ZG_PROF_BC_DAG, wbc_peel_kernel, dim3(gdag), dim3(256), lds, s, a, (int)(p + k));
                ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, wbc_close_kernel, dim3(1), dim3(1024), 0, s, a, (int)(p + k + 1));
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
            ZG_PROF_LAUNCH(ZG_PROF_BC_BACKWARD, wbc_backward_kernel, dim3(grid), dim3(256), lds, s, a, q);
        if (V)
            ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, wbc_reduce_kernel, dim3((V + 255) / 256), dim3(256), 0, s, V, nb, w.rec,
                           w.bc);
        ZG_TRY(hipGetLastError());
        inf->max_levels = std::max<uint32_t>(inf->max_levels, (uint32_t)last + 1);
        inf->push_levels += r;
        inf->batches++;
    }
    unsigned long long st[3] = {0, 0, 0};
    ZG_TRY(hipMemcpyAsync(st, w.stats, sizeof(st), hipMemcpyDeviceToHost, s));
    if (V) {
        ZG_PROF_LAUNCH(ZG_PROF_BC_SETUP, wbc_out_kernel, dim3((V + 255) / 256), dim3(256), 0, s, V, w.bc, g->live, d_out);
        ZG_TRY(hipGetLastError());
    }
    ZG_TRY(hipStreamSynchronize(s));
    inf->arcs_traversed += st[0];
    inf->nodes_reached += st[1];
    inf->arcs_relaxed += st[2];
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
    int rc = zg_io_src(g, list.size(), &d_src);
    if (rc == ZG_OK) rc = zg_io_out(g, &d_out);
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
    if (rc == ZG_OK) rc = zg_prof_collect(g, s);
    if (rc == ZG_OK && g->V) rc = zg_hip_code(hipMemcpy(out_bc, d_out, sizeof(double) * g->V, hipMemcpyDeviceToHost));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == ZG_OK && info) *info = inf;
    return rc;
}
