// Work mapping of the per-source engines: each source s has a range
// [lo, hi) of entries (queue positions); a wavefront takes 64 consecutive
// entries of one source.  Every block computes the per-source chunk counts
// and their prefix in LDS (all threads must call zg_q_chunks).
#pragma once
#include "common.h"

// Per-source ranges of this kernel's level into LDS, and the prefix of
// 64-pair chunk counts; returns the total number of chunks.
template <typename RangeFn>
__device__ __forceinline__ uint32_t zg_q_chunks(uint32_t S, RangeFn range, uint32_t *s_lo, uint32_t *s_hi,
                                                uint32_t *s_pre)
{
    __shared__ uint32_t wtot[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < S; base += blockDim.x) {
        const uint32_t s = base + tid;
        uint32_t lo = 0, hi = 0;
        if (s < S) range(s, lo, hi);
        const uint32_t cnt = hi > lo ? (hi - lo + 63) / 64 : 0;
        if (s < S) {
            s_lo[s] = lo;
            s_hi[s] = hi;
        }
        const uint32_t incl = zg_wave_incl_scan(cnt, lane);
        if (lane == 63) wtot[wv] = incl;
        __syncthreads();
        uint32_t off = carry;
        for (int w = 0; w < wv; w++) off += wtot[w];
        if (s < S) s_pre[s] = off + incl - cnt;
        uint32_t tot = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) tot += wtot[w];
        carry += tot;
        __syncthreads();
    }
    if (tid == 0) s_pre[S] = carry;
    __syncthreads();
    return s_pre[S];
}

// Source owning chunk c: the last s with s_pre[s] <= c.
__device__ __forceinline__ uint32_t zg_q_owner(const uint32_t *s_pre, uint32_t S, uint32_t c)
{
    uint32_t lo = 0, hi = S;  // s_pre[lo] <= c < s_pre[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_pre[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}


// XCD-aware walk over the chunks [0, total): workgroups are dispatched
// round-robin over the chip's 8 XCDs, so block b runs on XCD b % 8.  The
// chunk space is cut into 8 contiguous equal slices and slice x is walked
// only by the wavefronts of XCD x's blocks; a source's chunks are contiguous,
// so its records stay in one XCD's L2 (a grid not a multiple of 8 walks the
// chunks plainly).
struct ZgQWalk {
    uint32_t c, end, stride;
};

__device__ __forceinline__ ZgQWalk zg_q_walk(uint32_t total)
{
    const uint32_t wpb = blockDim.x >> 6, wv = threadIdx.x >> 6;
    const uint32_t nx = 8;
    if (gridDim.x % nx != 0) return ZgQWalk{blockIdx.x * wpb + wv, total, gridDim.x * wpb};
    const uint32_t x = blockIdx.x % nx, lb = blockIdx.x / nx, nb = gridDim.x / nx;
    const uint32_t c0 = (uint32_t)((uint64_t)total * x / nx), c1 = (uint32_t)((uint64_t)total * (x + 1) / nx);
    return ZgQWalk{c0 + lb * wpb + wv, c1, nb * wpb};
}
