// Native edge-list reader: nif_read_edgelist and addParsedEdges
// (lib/zog/resource_graph.ex:990-1084, :953-988).  Host code only.
//
//   * lines split at '\n', trimmed of " \t\r"; empty lines and lines that
//     start with '#' are skipped;
//   * tokens separated by any of ' ', '\t', ',' (tokenizeAny); a line needs
//     two; a line whose first two tokens are "source" and "target" (any
//     case) is a header and skipped;
//   * the third token is the weight (std.fmt.parseFloat, 1.0 when absent or
//     not a number);
//   * integer labels: both ids parsed as u32 in base 10 (std.fmt.parseInt;
//     the line is skipped when either fails), node count = largest id + 1;
//     otherwise labels get ids in order of first appearance;
//   * directed: arcs in file order; undirected: a stable sort by
//     (min(u, v), max(u, v)), the first arc of each pair kept, added in both
//     directions (once for a self-loop) with its weight.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/zog_hip.h"

struct zg_edgelist {
    std::vector<uint32_t> from, to;
    std::vector<double> weight;
    uint32_t n_nodes = 0;
    std::string blob;               // labels, concatenated in id order
    std::vector<uint64_t> offsets;  // [n_labels + 1]
};

namespace {

bool is_sep(char c) { return c == ' ' || c == '\t' || c == ','; }
bool is_digit(char c) { return c >= '0' && c <= '9'; }
bool is_hex(char c) { return is_digit(c) || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F'); }

bool ieq(std::string_view a, const char *b)
{
    const size_t n = strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; i++) {
        char x = a[i];
        if (x >= 'A' && x <= 'Z') x = (char)(x - 'A' + 'a');
        if (x != b[i]) return false;
    }
    return true;
}

// std.fmt.parseInt(u32, s, 10): an optional '+', decimal digits, '_'
// allowed inside (not first or last), no overflow.
bool parse_u32(std::string_view s, uint32_t &out)
{
    size_t i = 0;
    if (!s.empty() && s[0] == '+') i = 1;
    if (i >= s.size() || s[i] == '_' || s.back() == '_') return false;
    uint64_t v = 0;
    bool any = false;
    for (; i < s.size(); i++) {
        const char c = s[i];
        if (c == '_') continue;
        if (!is_digit(c)) return false;
        v = v * 10 + (uint64_t)(c - '0');
        if (v > 0xFFFFFFFFull) return false;
        any = true;
    }
    if (!any) return false;
    out = (uint32_t)v;
    return true;
}

// std.fmt.parseFloat(f64, s): decimal or hexadecimal (0x, p exponent)
// numbers with an optional sign, inf / infinity / nan in any case, '_'
// allowed only between two digits; the whole token must be a number.
bool parse_f64(std::string_view s, double &out)
{
    if (s.empty()) return false;
    std::string t;
    t.reserve(s.size());
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] == '_') {
            if (i == 0 || i + 1 >= s.size() || !is_hex(s[i - 1]) || !is_hex(s[i + 1])) return false;
            continue;
        }
        t.push_back(s[i]);
    }
    std::string_view body(t);
    if (!body.empty() && (body[0] == '+' || body[0] == '-')) body.remove_prefix(1);
    if (ieq(body, "inf") || ieq(body, "infinity")) {
        out = (t[0] == '-') ? -HUGE_VAL : HUGE_VAL;
        return true;
    }
    if (ieq(body, "nan")) {
        out = (double)NAN;
        return true;
    }
    // strtod also takes forms the Zig parser does not ("nan(...)", a leading
    // space): the checks above and below keep to plain numbers
    for (char c : body)
        if (!(is_hex(c) || c == '.' || c == 'x' || c == 'X' || c == 'p' || c == 'P' || c == '+' || c == '-'))
            return false;
    char *end = nullptr;
    errno = 0;
    const double v = strtod(t.c_str(), &end);
    if (end != t.c_str() + t.size() || end == t.c_str()) return false;
    out = v;  // overflow gives +-inf as in the Zig parser
    return true;
}

struct Parsed {
    uint32_t u, v;
    double w;
};

// Label -> id, open addressing over views into the file buffer.
class LabelTable {
  public:
    explicit LabelTable(size_t hint)
    {
        size_t c = 1024;
        while (c < 2 * hint) c <<= 1;
        slots_.assign(c, Slot{nullptr, 0, 0});
    }
    // id of `key`, inserting next_id when new (then *is_new = true)
    uint32_t get(std::string_view key, uint32_t next_id, bool *is_new)
    {
        if (2 * (count_ + 1) > slots_.size()) grow();
        const uint64_t h = hash(key);
        size_t i = h & (slots_.size() - 1);
        for (;;) {
            Slot &sl = slots_[i];
            if (!sl.p) {
                sl = Slot{key.data(), (uint32_t)key.size(), next_id};
                count_++;
                *is_new = true;
                return next_id;
            }
            if (sl.len == key.size() && memcmp(sl.p, key.data(), key.size()) == 0) {
                *is_new = false;
                return sl.id;
            }
            i = (i + 1) & (slots_.size() - 1);
        }
    }

  private:
    struct Slot {
        const char *p;
        uint32_t len, id;
    };
    static uint64_t hash(std::string_view k)  // FNV-1a, then a multiplicative mix
    {
        uint64_t h = 1469598103934665603ull;
        for (unsigned char c : k) h = (h ^ c) * 1099511628211ull;
        return (h ^ (h >> 29)) * 0xbf58476d1ce4e5b9ull;
    }
    void grow()
    {
        std::vector<Slot> old;
        old.swap(slots_);
        slots_.assign(old.size() * 2, Slot{nullptr, 0, 0});
        for (const Slot &sl : old)
            if (sl.p) {
                size_t i = hash(std::string_view(sl.p, sl.len)) & (slots_.size() - 1);
                while (slots_[i].p) i = (i + 1) & (slots_.size() - 1);
                slots_[i] = sl;
            }
    }
    std::vector<Slot> slots_;
    size_t count_ = 0;
};

// Stable LSD radix sort of (key, idx) pairs by key (16-bit digits, as many
// passes as `bits` needs); keys and indices move together.
void radix_sort_pairs(std::vector<uint64_t> &key, std::vector<uint32_t> &idx, unsigned bits)
{
    const size_t n = key.size();
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> i2(n);
    std::vector<uint64_t> cnt(1 << 16);
    for (unsigned shift = 0; shift < bits; shift += 16) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t j = 0; j < n; j++) cnt[(key[j] >> shift) & 0xFFFF]++;
        uint64_t sum = 0;
        for (auto &c : cnt) {
            const uint64_t t = c;
            c = sum;
            sum += t;
        }
        for (size_t j = 0; j < n; j++) {
            const uint64_t at = cnt[(key[j] >> shift) & 0xFFFF]++;
            k2[at] = key[j];
            i2[at] = idx[j];
        }
        key.swap(k2);
        idx.swap(i2);
    }
}

}  // namespace

extern "C" int zg_edgelist_read(const char *path, int directed, int integer_labels, zg_edgelist **out)
{
    if (!path || !out) return ZG_ERR_INVALID_ARG;
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return ZG_ERR_IO;
    std::string buf;
    try {
        char tmp[1 << 16];
        size_t k;
        while ((k = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.append(tmp, k);
    } catch (const std::bad_alloc &) {
        fclose(f);
        return ZG_ERR_OOM;
    }
    const bool err = ferror(f) != 0;
    fclose(f);
    if (err) return ZG_ERR_IO;

    zg_edgelist *el = new (std::nothrow) zg_edgelist();
    if (!el) return ZG_ERR_OOM;
    try {
        std::vector<Parsed> edges;
        LabelTable ids(integer_labels ? 0 : buf.size() / 64);
        std::vector<std::string_view> labels;
        uint64_t next_id = 0;
        size_t pos = 0;
        const size_t n = buf.size();
        while (pos < n) {
            size_t eol = buf.find('\n', pos);
            if (eol == std::string::npos) eol = n;
            size_t a = pos, b = eol;
            pos = eol + 1;
            while (a < b && (buf[a] == ' ' || buf[a] == '\t' || buf[a] == '\r')) a++;
            while (b > a && (buf[b - 1] == ' ' || buf[b - 1] == '\t' || buf[b - 1] == '\r')) b--;
            if (a == b || buf[a] == '#') continue;
            std::string_view tok[3];
            int nt = 0;
            for (size_t i = a; i < b && nt < 3;) {
                while (i < b && is_sep(buf[i])) i++;
                if (i >= b) break;
                size_t j = i;
                while (j < b && !is_sep(buf[j])) j++;
                tok[nt++] = std::string_view(buf.data() + i, j - i);
                i = j;
            }
            if (nt < 2) continue;
            if (ieq(tok[0], "source") && ieq(tok[1], "target")) continue;
            double w = 1.0;
            if (nt == 3 && !parse_f64(tok[2], w)) w = 1.0;
            uint32_t u, v;
            if (integer_labels) {
                if (!parse_u32(tok[0], u) || !parse_u32(tok[1], v)) continue;
                const uint64_t mx = std::max(u, v);
                if (mx + 1 > next_id) next_id = mx + 1;
            } else {
                uint32_t idv[2];
                for (int t = 0; t < 2; t++) {
                    if (next_id >= 0xFFFFFFFFull) {
                        delete el;
                        return ZG_ERR_TOO_LARGE;
                    }
                    bool is_new = false;
                    idv[t] = ids.get(tok[t], (uint32_t)next_id, &is_new);
                    if (is_new) {
                        next_id++;
                        labels.push_back(tok[t]);
                    }
                }
                u = idv[0];
                v = idv[1];
            }
            edges.push_back({u, v, w});
        }
        if (next_id > 0xFFFFFFFEull) {  // ids are u32 and the node count must fit beside them
            delete el;
            return ZG_ERR_TOO_LARGE;
        }
        el->n_nodes = (uint32_t)next_id;
        if (directed) {
            el->from.reserve(edges.size());
            el->to.reserve(edges.size());
            el->weight.reserve(edges.size());
            for (const Parsed &e : edges) {
                el->from.push_back(e.u);
                el->to.push_back(e.v);
                el->weight.push_back(e.w);
            }
        } else {
            std::vector<uint32_t> order(edges.size());
            std::iota(order.begin(), order.end(), 0u);
            unsigned idbits = 1;
            while (idbits < 32 && (next_id >> idbits)) idbits++;
            // (min, max) as one key: max in the low idbits bits, min above
            std::vector<uint64_t> packed(edges.size());
            for (size_t i = 0; i < edges.size(); i++)
                packed[i] = ((uint64_t)std::min(edges[i].u, edges[i].v) << idbits) | std::max(edges[i].u, edges[i].v);
            radix_sort_pairs(packed, order, 2 * idbits);
            uint64_t prev = ~0ull;
            for (uint32_t i : order) {
                const Parsed &e = edges[i];
                const uint64_t key = ((uint64_t)std::min(e.u, e.v) << 32) | std::max(e.u, e.v);
                if (key == prev) continue;
                prev = key;
                el->from.push_back(e.u);
                el->to.push_back(e.v);
                el->weight.push_back(e.w);
                if (e.u != e.v) {
                    el->from.push_back(e.v);
                    el->to.push_back(e.u);
                    el->weight.push_back(e.w);
                }
            }
        }
        el->offsets.reserve(labels.size() + 1);
        el->offsets.push_back(0);
        for (std::string_view l : labels) {
            el->blob.append(l.data(), l.size());
            el->offsets.push_back(el->blob.size());
        }
    } catch (const std::bad_alloc &) {
        delete el;
        return ZG_ERR_OOM;
    }
    *out = el;
    return ZG_OK;
}

extern "C" int zg_edgelist_arcs(const zg_edgelist *el, const uint32_t **from, const uint32_t **to,
                                const double **weight, uint64_t *n_arcs, uint32_t *n_nodes)
{
    if (!el) return ZG_ERR_INVALID_HANDLE;
    if (from) *from = el->from.data();
    if (to) *to = el->to.data();
    if (weight) *weight = el->weight.data();
    if (n_arcs) *n_arcs = el->from.size();
    if (n_nodes) *n_nodes = el->n_nodes;
    return ZG_OK;
}

extern "C" int zg_edgelist_labels(const zg_edgelist *el, const char **blob, const uint64_t **offsets,
                                  uint32_t *n_labels)
{
    if (!el) return ZG_ERR_INVALID_HANDLE;
    if (blob) *blob = el->blob.data();
    if (offsets) *offsets = el->offsets.data();
    if (n_labels) *n_labels = (uint32_t)(el->offsets.size() - 1);
    return ZG_OK;
}

extern "C" void zg_edgelist_free(zg_edgelist *el)
{
    delete el;
}
