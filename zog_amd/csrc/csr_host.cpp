// Host CSR build of a resource (see csr_host.h).
//
// Reference: buildGraph (lib/zog/resource_graph.ex:120-129) appends all nodes
// then calls addEdge (priv/zog/src/models/array_graph.zig:88-108) in array
// order; addEdge prepends, so successors() yields arcs in reverse insertion
// order (array_graph.zig:94-104, 166-180) and skips arcs to or from
// tombstoned nodes.  PageRank's in-lists are filled by scanning live sources
// in id order (centrality.zig:881-899): ascending by source, duplicates kept.
#include "csr_host.h"

#include <algorithm>
#include <cmath>

#include "../../include/zog_hip.h"

int zg_host_csr(uint32_t V, const uint8_t *live, const uint32_t *from, const uint32_t *to, const double *weight,
                uint64_t m, ZgHostCsr &c)
{
    c.out_ptr.assign((size_t)V + 1, 0);
    for (uint64_t e = 0; e < m; e++) {
        const uint32_t f = from[e], t = to[e];
        if (f >= V || t >= V) return ZG_ERR_NODE_NOT_FOUND;
        if (live[f] && live[t]) c.out_ptr[(size_t)f + 1]++;
    }
    for (uint32_t i = 0; i < V; i++) c.out_ptr[(size_t)i + 1] += c.out_ptr[i];
    const uint64_t E = c.out_ptr[V];
    c.out_col.assign(E, 0);
    c.out_w.assign(weight ? E : 0, 0.0);
    c.wsum = 0.0;
    c.weights_ok = true;
    {
        // filling each row from its end in insertion order leaves the newest
        // arc first
        std::vector<uint32_t> fill(c.out_ptr.begin() + 1, c.out_ptr.end());
        for (uint64_t e = 0; e < m; e++) {
            const uint32_t f = from[e], t = to[e];
            if (!(live[f] && live[t])) continue;
            const uint32_t k = --fill[f];
            c.out_col[k] = t;
            if (weight) {
                c.out_w[k] = weight[e];
                c.wsum += weight[e];
                // NaN, negative or infinite: no shortest paths under the reference's semantics
                if (!(weight[e] >= 0.0) || weight[e] == HUGE_VAL) c.weights_ok = false;
            }
        }
    }
    c.in_ptr.assign((size_t)V + 1, 0);
    for (uint64_t k = 0; k < E; k++) c.in_ptr[(size_t)c.out_col[k] + 1]++;
    for (uint32_t i = 0; i < V; i++) c.in_ptr[(size_t)i + 1] += c.in_ptr[i];
    c.in_col.assign(E, 0);
    {
        std::vector<uint32_t> fill(c.in_ptr.begin(), c.in_ptr.end() - 1);
        for (uint32_t v = 0; v < V; v++)
            for (uint32_t k = c.out_ptr[v]; k < c.out_ptr[(size_t)v + 1]; k++) c.in_col[fill[c.out_col[k]]++] = v;
    }
    c.pr_order.resize(V);
    for (uint32_t v = 0; v < V; v++) c.pr_order[v] = v;
    std::stable_sort(c.pr_order.begin(), c.pr_order.end(), [&](uint32_t x, uint32_t y) {
        return c.out_ptr[(size_t)x + 1] - c.out_ptr[x] > c.out_ptr[(size_t)y + 1] - c.out_ptr[y];
    });
    c.pr_id.assign(V, 0);
    for (uint32_t i = 0; i < V; i++) c.pr_id[c.pr_order[i]] = i;
    c.pr_col.assign(E, 0);
    for (size_t e = 0; e < E; e++) c.pr_col[e] = c.pr_id[c.in_col[e]];
    return ZG_OK;
}
