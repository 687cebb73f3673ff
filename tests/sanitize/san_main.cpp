// Drives the library's host code and the CPU oracle under AddressSanitizer
// and UndefinedBehaviorSanitizer (built by tests/sanitize/Makefile, run by
// tests/test_sanitize.py).  Usage: san_main <scratch dir> <golden edge list>.
// Exits 0 when every check holds; the sanitizers abort on the first error.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/zog_hip.h"
#include "../../zog_amd/csrc/csr_host.h"

extern "C" {
struct zo_graph;
int zo_graph_create(uint32_t n_cap, const uint8_t *live_mask, const uint32_t *from, const uint32_t *to, uint64_t m,
                    zo_graph **out);
void zo_graph_destroy(zo_graph *g);
uint64_t zo_successors(const zo_graph *g, uint32_t v, uint32_t *out, uint64_t cap);
int zo_bc_unweighted(const zo_graph *g, const uint32_t *sources, uint64_t n_sources, int mode, int n_threads,
                     double *out_bc, uint64_t *arcs_per_source);
int zo_bfs_paths(const zo_graph *g, uint32_t s, uint32_t *depth_out, uint64_t *sigma_u_out, double *sigma_f_out,
                 int32_t *shift_out, uint64_t *arcs_scanned, uint32_t *levels, int *overflow);
int zo_bfs_paths_exp(const zo_graph *g, uint32_t s, uint32_t *depth_out, double *mant_out, int32_t *exp_out);
int zo_pagerank_mt(const zo_graph *g, double damping, uint64_t max_iterations, double tolerance, int n_threads,
                   double *out, uint64_t *iters_out, double *l1_out);
int zo_core_numbers(const zo_graph *g, uint32_t *out);
int zo_graph_set_weights(zo_graph *g, const double *w, uint64_t m);
int zo_bc_f64(const zo_graph *g, const uint32_t *sources, uint64_t n_sources, double *out);
int zo_eigenvector(const zo_graph *g, uint64_t max_iterations, double tolerance, double *out, uint64_t *iters_out,
                   int *outcome_out);
int zo_katz_alpha(const zo_graph *g, double alpha, double beta, int has_beta, double start, uint64_t max_iterations,
                  double tolerance, double *out, uint64_t *iters_out, double *l1_out);
}

static int failures = 0;

#define CHECK(cond)                                                           \
    do {                                                                      \
        if (!(cond)) {                                                        \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            failures++;                                                       \
        }                                                                     \
    } while (0)

static void write_file(const std::string &path, const std::string &text)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(2);
    }
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
}

// Reads one file through the C ABI in every mode and walks every output.
static void edgelist_case(const std::string &path, uint64_t expect_arcs_directed_int = UINT64_MAX,
                          int int_rc = ZG_OK)
{
    for (int directed = 0; directed < 2; directed++)
        for (int ints = 0; ints < 2; ints++) {
            zg_edgelist *el = nullptr;
            const int want = ints ? int_rc : ZG_OK;
            int rc = zg_edgelist_read(path.c_str(), directed, ints, &el);
            if (rc != want) std::fprintf(stderr, "read rc %d (directed %d, integer labels %d)\n", rc, directed, ints);
            CHECK(rc == want);
            if (rc != ZG_OK) continue;
            const uint32_t *from, *to;
            const double *w;
            uint64_t m;
            uint32_t n;
            CHECK(zg_edgelist_arcs(el, &from, &to, &w, &m, &n) == ZG_OK);
            double acc = 0.0;
            for (uint64_t i = 0; i < m; i++) {
                CHECK(from[i] < n && to[i] < n);
                acc += w[i];
            }
            (void)acc;
            if (directed && ints && expect_arcs_directed_int != UINT64_MAX) CHECK(m == expect_arcs_directed_int);
            const char *blob;
            const uint64_t *off;
            uint32_t nl;
            CHECK(zg_edgelist_labels(el, &blob, &off, &nl) == ZG_OK);
            if (!ints) {
                CHECK(nl == n);
                for (uint32_t i = 0; i < nl; i++) CHECK(off[i] <= off[i + 1]);
            }
            zg_edgelist_free(el);
        }
}

static void edgelist_cases(const std::string &dir, const std::string &golden)
{
    edgelist_case(golden);
    const std::string p = dir + "/san_edges.txt";
    write_file(p, "# comment\nsource target\n0 1 1.5\n1,2,0x1.8p1\n2\t0\tinf\n\n   \n3 3\n1 0 -2\n0 1\n7 x\n");
    edgelist_case(p);
    write_file(p, "");
    edgelist_case(p, 0);
    // the largest u32 id needs 2^32 nodes (too many); the second id does not fit u32
    write_file(p, "4294967295 0\n0 4294967296\n");
    edgelist_case(p, UINT64_MAX, ZG_ERR_TOO_LARGE);
    write_file(p, "a b\nb c\nc a\na b 2.0");  // no trailing newline
    edgelist_case(p);
    std::string big;
    std::mt19937_64 rng(3);
    for (int i = 0; i < 20000; i++)
        big += std::to_string(rng() % 500) + ((i & 1) ? "," : " ") + std::to_string(rng() % 500) +
               ((i % 3) ? "\n" : " 0.25\n");
    write_file(p, big);
    edgelist_case(p, 20000);
    std::string longline(100000, 'x');
    write_file(p, longline + " y\n1 2\n");
    edgelist_case(p);
    zg_edgelist *el = nullptr;
    CHECK(zg_edgelist_read((dir + "/no_such_file").c_str(), 1, 0, &el) == ZG_ERR_IO);
}

// The host CSR against a direct construction from the arc list.
static void csr_cases()
{
    std::mt19937_64 rng(7);
    for (int t = 0; t < 40; t++) {
        const uint32_t V = 1 + (uint32_t)(rng() % 300);
        const uint64_t m = rng() % (6 * (uint64_t)V + 1);
        std::vector<uint32_t> from(m), to(m);
        std::vector<double> w(m);
        std::vector<uint8_t> live(V, 1);
        for (uint32_t v = 0; v < V; v++) live[v] = (rng() % 10) != 0;
        for (uint64_t e = 0; e < m; e++) {
            from[e] = (uint32_t)(rng() % V);
            to[e] = (uint32_t)(rng() % V);
            w[e] = (double)(rng() % 1000) / 100.0;
        }
        ZgHostCsr c;
        CHECK(zg_host_csr(V, live.data(), from.data(), to.data(), (t & 1) ? w.data() : nullptr, m, c) == ZG_OK);
        // row v: live arcs of v, newest first
        for (uint32_t v = 0; v < V; v++) {
            std::vector<uint32_t> row;
            for (uint64_t e = m; e-- > 0;)
                if (from[e] == v && live[from[e]] && live[to[e]]) row.push_back(to[e]);
            CHECK(c.out_ptr[v + 1] - c.out_ptr[v] == row.size());
            for (size_t k = 0; k < row.size() && k < c.out_ptr[v + 1] - c.out_ptr[v]; k++)
                CHECK(c.out_col[c.out_ptr[v] + k] == row[k]);
            for (uint32_t k = c.in_ptr[v] + 1; k < c.in_ptr[v + 1]; k++) CHECK(c.in_col[k - 1] <= c.in_col[k]);
        }
        CHECK(c.in_ptr[V] == c.out_ptr[V]);
        for (uint32_t i = 0; i < V; i++) CHECK(c.pr_order[c.pr_id[i]] == i);
    }
    // an id past V is NodeNotFound
    std::vector<uint32_t> f{0, 5}, tt{1, 1};
    std::vector<uint8_t> live(3, 1);
    ZgHostCsr c;
    CHECK(zg_host_csr(3, live.data(), f.data(), tt.data(), nullptr, 2, c) == ZG_ERR_NODE_NOT_FOUND);
}

static void oracle_cases()
{
    std::mt19937_64 rng(11);
    for (int t = 0; t < 24; t++) {
        const uint32_t V = (t == 0) ? 0 : (t == 1 ? 1 : 2 + (uint32_t)(rng() % 120));
        const uint64_t m = V ? rng() % (5 * (uint64_t)V + 1) : 0;
        std::vector<uint32_t> from(m), to(m);
        std::vector<double> w(m);
        std::vector<uint8_t> live(V ? V : 1, 1);
        for (uint32_t v = 0; v < V; v++) live[v] = (rng() % 8) != 0;
        for (uint64_t e = 0; e < m; e++) {
            from[e] = (uint32_t)(rng() % V);
            to[e] = (uint32_t)(rng() % V);
            w[e] = (double)(1 + rng() % 5);
        }
        zo_graph *g = nullptr;
        CHECK(zo_graph_create(V, (t & 1) ? live.data() : nullptr, from.data(), to.data(), m, &g) == 0);
        if (!g) continue;
        std::vector<double> bc(V ? V : 1), bc3(V ? V : 1);
        for (int mode = 0; mode < 3; mode++) {
            CHECK(zo_bc_unweighted(g, nullptr, 0, mode, 1, bc.data(), nullptr) >= 0);
            CHECK(zo_bc_unweighted(g, nullptr, 0, mode, 3, bc3.data(), nullptr) >= 0);
            for (uint32_t v = 0; v < V; v++) CHECK(std::fabs(bc[v] - bc3[v]) <= 1e-9 * (1.0 + std::fabs(bc[v])));
        }
        if (V) {
            std::vector<uint32_t> depth(V);
            std::vector<uint64_t> su(V);
            std::vector<double> sf(V), mant(V);
            std::vector<int32_t> shift(V + 1), ex(V);
            uint64_t arcs;
            uint32_t levels;
            int ovf;
            for (uint32_t s = 0; s < V; s += 7) {
                if ((t & 1) && !live[s]) continue;
                CHECK(zo_bfs_paths(g, s, depth.data(), su.data(), sf.data(), shift.data(), &arcs, &levels, &ovf) == 0);
                CHECK(zo_bfs_paths_exp(g, s, depth.data(), mant.data(), ex.data()) == 0);
            }
            std::vector<uint32_t> succ(m + 1);
            for (uint32_t v = 0; v < V; v++) CHECK(zo_successors(g, v, succ.data(), succ.size()) <= m);
        }
        std::vector<double> pr1(V ? V : 1), pr4(V ? V : 1);
        uint64_t it1, it4;
        double l1a, l1b;
        CHECK(zo_pagerank_mt(g, 0.85, 50, 1e-9, 1, pr1.data(), &it1, &l1a) == 0);
        CHECK(zo_pagerank_mt(g, 0.85, 50, 1e-9, 4, pr4.data(), &it4, &l1b) == 0);
        CHECK(it1 == it4 && std::memcmp(pr1.data(), pr4.data(), sizeof(double) * V) == 0);
        std::vector<uint32_t> core(V ? V : 1);
        CHECK(zo_core_numbers(g, core.data()) == 0);
        CHECK(zo_graph_set_weights(g, w.data(), m) == 0);
        CHECK(zo_bc_f64(g, nullptr, 0, bc.data()) == 0);
        uint64_t it;
        int oc;
        double l1;
        CHECK(zo_eigenvector(g, 30, 1e-6, pr1.data(), &it, &oc) == 0);
        CHECK(zo_katz_alpha(g, 0.01, 1.0, 1, 1.0, 30, 1e-9, pr1.data(), &it, &l1) == 0);
        CHECK(zo_katz_alpha(g, 0.01, 0.0, 0, 1.0, 30, 1e-9, pr1.data(), &it, &l1) == 0);
        zo_graph_destroy(g);
    }
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: san_main <scratch dir> <golden edge list>\n");
        return 2;
    }
    edgelist_cases(argv[1], argv[2]);
    csr_cases();
    oracle_cases();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitized host code: all checks passed\n");
    return 0;
}
