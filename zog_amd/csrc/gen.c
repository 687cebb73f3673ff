/*
 * Seeded synthetic-graph generators that are too sequential for numpy.
 *
 * Barabasi-Albert, the method of lib/zog/generator.ex:131-198: nodes
 * 0..n-1; a clique on 0..m-1; a target list holding every node once per unit
 * of degree (clique nodes m-1 times each); each new node u draws m distinct
 * targets uniformly from that list (redrawing repeats), links to them in
 * ascending order, then appends u, the targets, and u another m-1 times.
 * Undirected: every edge is emitted as the arcs (u, v), (v, u).
 *
 * R-MAT (Chakrabarti, Zhan & Faloutsos): each arc picks one quadrant per
 * bit level with probabilities a, b, c, 1-a-b-c (source bit set in the lower
 * half, target bit in the right half); arcs are drawn until n_arcs distinct
 * (source, target) pairs exist and are kept in first-drawn order;
 * self-loops are kept.
 *
 * Random numbers: xoshiro256** (Blackman & Vigna) seeded by SplitMix64.
 */
#include <stdint.h>
#include <stdlib.h>

typedef struct {
    uint64_t s[4];
} zgen_rng;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(zgen_rng *r, uint64_t seed)
{
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&seed);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(zgen_rng *r)
{
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform integer in [0, bound) by Lemire's multiply-shift with rejection */
static uint64_t rng_below(zgen_rng *r, uint64_t bound)
{
    for (;;) {
        uint64_t x = rng_next(r);
        __uint128_t m = (__uint128_t)x * bound;
        uint64_t lo = (uint64_t)m;
        if (lo >= bound || lo >= (-bound) % bound) return (uint64_t)(m >> 64);
    }
}

/* Number of arcs zgen_barabasi_albert writes. */
uint64_t zgen_ba_arcs(uint32_t n, uint32_t m)
{
    if (m < 1 || n <= m) return 0;
    return 2ull * ((uint64_t)m * (m - 1) / 2 + (uint64_t)(n - m) * m);
}

/* Returns the number of arcs written, or 0 for invalid parameters. */
uint64_t zgen_barabasi_albert(uint32_t n, uint32_t m, uint64_t seed, uint32_t *src, uint32_t *dst)
{
    if (m < 1 || n <= m) return 0;
    zgen_rng r;
    rng_seed(&r, seed);
    uint64_t k = 0;
    for (uint32_t u = 0; u + 1 < m; u++)
        for (uint32_t v = u + 1; v < m; v++) {
            src[k] = u; dst[k] = v; k++;
            src[k] = v; dst[k] = u; k++;
        }
    const uint64_t cap = (uint64_t)m * (m - 1) + (uint64_t)(n - m) * 2 * m + 1;
    uint32_t *deg = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    uint32_t *chosen = (uint32_t *)malloc(sizeof(uint32_t) * m);
    if (!deg || !chosen) {
        free(deg);
        free(chosen);
        return 0;
    }
    uint64_t size = 0;
    if (m > 1) {
        for (uint32_t u = 0; u < m; u++)
            for (uint32_t t = 0; t + 1 < m; t++) deg[size++] = u;
    } else {
        deg[size++] = 0;
    }
    for (uint32_t u = m; u < n; u++) {
        uint32_t c = 0;
        while (c < m) {
            uint32_t node = deg[rng_below(&r, size)];
            int dup = 0;
            for (uint32_t i = 0; i < c; i++) dup |= chosen[i] == node;
            if (!dup) chosen[c++] = node;
        }
        /* ascending target order (insertion sort, m is small) */
        for (uint32_t i = 1; i < m; i++) {
            uint32_t x = chosen[i];
            uint32_t j = i;
            while (j > 0 && chosen[j - 1] > x) {
                chosen[j] = chosen[j - 1];
                j--;
            }
            chosen[j] = x;
        }
        for (uint32_t i = 0; i < m; i++) {
            src[k] = u; dst[k] = chosen[i]; k++;
            src[k] = chosen[i]; dst[k] = u; k++;
        }
        deg[size++] = u;
        for (uint32_t i = 0; i < m; i++) deg[size++] = chosen[i];
        for (uint32_t t = 0; t + 1 < m; t++) deg[size++] = u;
    }
    free(deg);
    free(chosen);
    return k;
}

/* Writes n_arcs distinct R-MAT arcs on 2^scale nodes; returns n_arcs, or 0
 * for invalid parameters or when memory runs out. */
uint64_t zgen_rmat(uint32_t scale, uint64_t n_arcs, uint64_t seed, double a, double b, double c, uint32_t *src,
                   uint32_t *dst)
{
    if (scale < 1 || scale > 31 || !(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return 0;
    if (n_arcs > (1ull << scale) * (1ull << scale)) return 0;
    /* open-addressing set of keys src<<32|dst (+1 so 0 marks empty) */
    uint64_t cap = 1024;
    while (cap < 2 * n_arcs + 1024) cap <<= 1;
    uint64_t *set = (uint64_t *)calloc(cap, sizeof(uint64_t));
    if (!set) return 0;
    zgen_rng r;
    rng_seed(&r, seed);
    /* quadrant thresholds on a 32-bit uniform */
    const double two32 = 4294967296.0;
    const uint64_t ta = (uint64_t)(a * two32), tab = (uint64_t)((a + b) * two32), tabc = (uint64_t)((a + b + c) * two32);
    uint64_t k = 0;
    while (k < n_arcs) {
        uint32_t s = 0, d = 0;
        for (uint32_t lvl = 0; lvl < scale; lvl += 2) {
            const uint64_t x = rng_next(&r);
            for (uint32_t h = 0; h < 2 && lvl + h < scale; h++) {
                const uint64_t u = (x >> (32 * h)) & 0xFFFFFFFFull;
                const uint32_t bit = scale - 1 - (lvl + h);
                const uint32_t sb = u >= tab, db = (u >= ta && u < tab) || u >= tabc;
                s |= sb << bit;
                d |= db << bit;
            }
        }
        const uint64_t key = ((uint64_t)s << 32 | d) + 1;
        uint64_t h = (key * 0x9E3779B97F4A7C15ull) >> 17;
        for (;;) {
            h &= cap - 1;
            if (set[h] == key) break;
            if (set[h] == 0) {
                set[h] = key;
                src[k] = s;
                dst[k] = d;
                k++;
                break;
            }
            h++;
        }
    }
    free(set);
    return k;
}
