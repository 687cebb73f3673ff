// Host form of the resource's CSR build (no HIP calls, so it also builds and
// runs under the CPU sanitizers).  graph.hip uploads the arrays.
#pragma once
#include <cstdint>
#include <vector>

struct ZgHostCsr {
    std::vector<uint32_t> out_ptr, out_col;  // successor CSR, newest arc first in every row
    std::vector<double> out_w;               // arc weights in successor order (weighted builds)
    std::vector<uint32_t> in_ptr, in_col;    // predecessor CSR, each row ascending by source
    std::vector<uint32_t> pr_order;          // vertices by decreasing out-degree (stable)
    std::vector<uint32_t> pr_id, pr_col;     // PageRank id of each vertex; in_col as PageRank ids
    double wsum = 0.0;                       // sum of the live arcs' weights
    bool weights_ok = true;                  // every live weight finite and non-negative
};

// Builds the CSR of arcs from[e] -> to[e] (e < m) over V ids, keeping arcs
// whose endpoints are both live (live[v] != 0).  Returns 0, or -2
// (ZG_ERR_NODE_NOT_FOUND) for an id >= V.
int zg_host_csr(uint32_t V, const uint8_t *live, const uint32_t *from, const uint32_t *to, const double *weight,
                uint64_t m, ZgHostCsr &out);
