// Device helpers of the resource build: scratch allocations freed on every
// exit path, and rocPRIM's exclusive scan and stable radix sorts with their
// temporary storage taken from that scratch (defined in graph_build.hip, the
// one file that instantiates rocPRIM).
#pragma once
#include "common.h"

#include <vector>

struct ZgScratch {
    std::vector<void *> p;
    template <typename T>
    int get(T **out, size_t n)
    {
        int rc = zg_dmalloc(out, n);
        if (rc == ZG_OK) p.push_back(*out);
        return rc;
    }
    ~ZgScratch()
    {
        for (void *q : p) (void)hipFree(q);
    }
};

// out[0..n) = exclusive prefix sums of in[0..n) (pass n = count + 1 with
// in[count] = 0 to get the total in out[count]).
int zg_exclusive_scan(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s, ZgScratch &sc);
// The same with caller-held temporary storage: zg_scan_bytes(n) bytes serve
// every scan of at most n entries (ZG_ERR_INVALID_ARG if cap falls short).
size_t zg_scan_bytes(size_t n);
int zg_exclusive_scan_tmp(const uint32_t *in, uint32_t *out, size_t n, void *tmp, size_t cap, hipStream_t s);

// Stable sort of (key, value) pairs by the key's bits [0, bits).
int zg_sort_pairs(const uint32_t *kin, uint32_t *kout, const uint32_t *vin, uint32_t *vout, uint64_t n, unsigned bits,
                  hipStream_t s, ZgScratch &sc);
