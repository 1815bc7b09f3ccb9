// Schedule compiler: validates the reference order.txt arrays and turns them
// into the device schedule of the walk kernel and the entry list of the
// embedding (shared-minimum) form.
//
// Semantics follow _Comp_Tree::Calc_sim (comp.h:332-375): node i reads child 1
// (leaf order_1 >= 0, else register order_1 + REG_SIZE), child 2 likewise,
// then writes register order_d + REG_SIZE. Registers are renamed here by value
// liveness; the dataflow (which write each read sees) is unchanged.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "cms_internal.h"

namespace cms {

static thread_local std::string g_err;
void set_global_error(const std::string &msg) { g_err = msg; }
const char *global_error() { return g_err.c_str(); }

namespace {

// floor(v * 10^s) and the rounding remainder of v = m * 2^k (m < 2^53,
// -127 <= k < 0, 0 <= s <= 20) in 128-bit integers: m * 10^s < 2^127.
struct Scaled {
    unsigned __int128 q, rem, half;  // q = floor, rem = exact remainder, half = 2^(-k-1)
};

Scaled scale(uint64_t m, int k, int s) {
    static const uint64_t p10[21] = {1ull,
                                     10ull,
                                     100ull,
                                     1000ull,
                                     10000ull,
                                     100000ull,
                                     1000000ull,
                                     10000000ull,
                                     100000000ull,
                                     1000000000ull,
                                     10000000000ull,
                                     100000000000ull,
                                     1000000000000ull,
                                     10000000000000ull,
                                     100000000000000ull,
                                     1000000000000000ull,
                                     10000000000000000ull,
                                     100000000000000000ull,
                                     1000000000000000000ull,
                                     10000000000000000000ull,
                                     0ull};
    unsigned __int128 P = (unsigned __int128)m * (s <= 19 ? p10[s] : p10[19]);
    if (s == 20) P *= 10u;
    const int sh = -k;
    Scaled r;
    r.q = P >> sh;
    r.rem = P - (r.q << sh);
    r.half = (unsigned __int128)1 << (sh - 1);
    return r;
}

// printf("%.6g", v) for v in [1e-15, 1e5): the six significant digits of
// the exact binary value rounded half to even (what glibc's printf does),
// then %g's choice of fixed or exponent notation with trailing zeros
// removed. Returns false outside that range (the caller uses snprintf).
bool fast_g6(double v, char *out, int &len) {
    if (!(v >= 1e-15 && v < 1e5)) return false;
    uint64_t bits;
    memcpy(&bits, &v, 8);
    const int be = (int)((bits >> 52) & 0x7ff);
    if (be == 0) return false;  // subnormal: below the range anyway
    const uint64_t m = (bits & ((1ull << 52) - 1)) | (1ull << 52);
    const int k = be - 1075;  // v = m * 2^k
    if (k >= 0 || k < -127) return false;
    // decimal exponent from the binary one (floor(e2 * log10 2), off by at
    // most one, corrected below)
    const int e2 = be - 1023;
    int e10 = (e2 * 78913) >> 18;
    Scaled sc{};
    for (int tries = 0; tries < 3; tries++) {
        const int sp = 5 - e10;
        if (sp < 0 || sp > 20) return false;
        sc = scale(m, k, sp);
        if (sc.q < 100000u) {
            e10--;
        } else if (sc.q >= 1000000u) {
            e10++;
        } else {
            break;
        }
    }
    if (sc.q < 100000u || sc.q >= 1000000u) return false;
    uint64_t N = (uint64_t)sc.q;
    if (sc.rem > sc.half || (sc.rem == sc.half && (N & 1u))) N++;
    int X = e10;
    if (N == 1000000u) {
        N = 100000u;
        X++;
    }
    char d[6];
    {
        const uint32_t hi = (uint32_t)(N / 1000u), lo = (uint32_t)(N % 1000u);
        d[0] = (char)('0' + hi / 100u);
        d[1] = (char)('0' + (hi / 10u) % 10u);
        d[2] = (char)('0' + hi % 10u);
        d[3] = (char)('0' + lo / 100u);
        d[4] = (char)('0' + (lo / 10u) % 10u);
        d[5] = (char)('0' + lo % 10u);
    }
    int nd = 6;
    while (nd > 1 && d[nd - 1] == '0') nd--;
    int n = 0;
    if (X < -4 || X >= 6) {
        out[n++] = d[0];
        if (nd > 1) {
            out[n++] = '.';
            for (int i = 1; i < nd; i++) out[n++] = d[i];
        }
        out[n++] = 'e';
        out[n++] = X < 0 ? '-' : '+';
        const int ax = X < 0 ? -X : X;
        if (ax >= 100) out[n++] = (char)('0' + ax / 100);
        out[n++] = (char)('0' + (ax / 10) % 10);
        out[n++] = (char)('0' + ax % 10);
    } else if (X >= 0) {
        for (int i = 0; i <= X; i++) out[n++] = i < nd ? d[i] : '0';
        if (nd > X + 1) {
            out[n++] = '.';
            for (int i = X + 1; i < nd; i++) out[n++] = d[i];
        }
    } else {
        out[n++] = '0';
        out[n++] = '.';
        for (int i = 0; i < -X - 1; i++) out[n++] = '0';
        for (int i = 0; i < nd; i++) out[n++] = d[i];
    }
    len = n;
    return true;
}

}  // namespace

int write_g6(char *out, double v) {
    int n = 0;
    if (!fast_g6(v, out, n)) n = snprintf(out, 32, "%.6g", v);
    return n;
}

void append_g6(std::string &s, double v) {
    char buf[64];
    const int n = write_g6(buf, v);
    s.append(buf, (size_t)n);
}

// The formatter's own check: the fast path's text and snprintf's, for the
// tests (returns 1 when they differ).
int g6_fast_differs(double v) {
    char a[64], b[64];
    int na = 0;
    if (!fast_g6(v, a, na)) return 0;
    const int nb = snprintf(b, sizeof b, "%.6g", v);
    return na != nb || memcmp(a, b, (size_t)na) != 0;
}

// 1 - dist with the MIN_DIST clamp, FP32 exactly as comp.h:341-344.
static inline float weight_of(float dist) {
    float w = 1 - dist;
    w = (w < 0) ? (float)0.00001 : w;
    return w;
}

int compile_tree(int leaf_n, int order_n, const int32_t *o1, const float *d1,
                 const int32_t *o2, const float *d2, const int32_t *od,
                 int leaf_chunk, CompiledTree &out, std::string &err) {
    char msg[256];
    if (leaf_n < 0 || order_n < 0) {
        err = "negative leaf_n/order_n";
        return CMS_E_INVAL;
    }
    if (order_n > 0 && (!o1 || !d1 || !o2 || !d2 || !od)) {
        err = "null schedule array";
        return CMS_E_INVAL;
    }
    if (leaf_chunk < 2) leaf_chunk = 2;
    out = CompiledTree();
    out.leaf_n = leaf_n;
    out.order_n = order_n;
    out.leaf_chunk = leaf_chunk;

    // Pass 1: range checks and value-id dataflow.
    std::vector<int32_t> cur(CMS_REG_SIZE, -1);  // value id held by each original register
    std::vector<int32_t> reads(order_n, 0), last_read(order_n, -1);
    std::vector<int32_t> child_val(2 * (size_t)order_n, -1);  // producer node or -1 for leaf
    std::vector<int32_t> leaf_uses(leaf_n, 0);
    for (int i = 0; i < order_n; i++) {
        const int32_t ch[2] = {o1[i], o2[i]};
        for (int j = 0; j < 2; j++) {
            int32_t c = ch[j];
            if (c >= 0) {
                if (c >= leaf_n) {
                    snprintf(msg, sizeof msg, "node %d child %d: leaf %d >= leaf_n %d", i, j + 1, c, leaf_n);
                    err = msg;
                    return CMS_E_SCHEDULE;
                }
                leaf_uses[c]++;
            } else {
                if (c < -CMS_REG_SIZE) {
                    snprintf(msg, sizeof msg, "node %d child %d: register %d out of range", i, j + 1, c);
                    err = msg;
                    return CMS_E_SCHEDULE;
                }
                int32_t v = cur[c + CMS_REG_SIZE];
                if (v < 0) {
                    snprintf(msg, sizeof msg, "node %d child %d: register %d read before write", i, j + 1, c);
                    err = msg;
                    return CMS_E_SCHEDULE;
                }
                child_val[2 * (size_t)i + j] = v;
                // A node reading the same value twice counts once for liveness
                // but twice for the tree test.
                reads[v]++;
                last_read[v] = i;
            }
        }
        if (od[i] >= 0 || od[i] < -CMS_REG_SIZE) {
            snprintf(msg, sizeof msg, "node %d: destination register %d out of range", i, od[i]);
            err = msg;
            return CMS_E_SCHEDULE;
        }
        cur[od[i] + CMS_REG_SIZE] = i;
    }
    bool is_tree = true;
    for (int l = 0; l < leaf_n; l++)
        if (leaf_uses[l] > 1) is_tree = false;
    for (int i = 0; i < order_n; i++)
        if (reads[i] > 1) is_tree = false;
    out.is_tree = is_tree;

    // Pass 2: liveness re-allocation (lowest free slot at definition).
    std::vector<int32_t> slot_of(order_n, -1);
    std::vector<char> busy(CMS_REG_SIZE, 0);
    int nslots = 0;
    out.nodes.resize(order_n);
    out.leaf_stream.clear();
    out.leaf_stream.reserve((size_t)order_n + 1);
    for (int i = 0; i < order_n; i++) {
        WalkNode nd{};
        nd.leafpos = (int32_t)out.leaf_stream.size();
        const int32_t ch[2] = {o1[i], o2[i]};
        uint32_t code = 0;
        for (int j = 0; j < 2; j++) {
            if (ch[j] >= 0) {
                code |= (j == 0) ? kC1Leaf : kC2Leaf;
                out.leaf_stream.push_back(ch[j]);
            } else {
                int32_t v = child_val[2 * (size_t)i + j];
                code |= (uint32_t)slot_of[v] << (8 * j);
            }
        }
        for (int j = 0; j < 2; j++) {
            if (ch[j] < 0) {
                int32_t v = child_val[2 * (size_t)i + j];
                if (last_read[v] == i && slot_of[v] >= 0 && busy[slot_of[v]]) busy[slot_of[v]] = 0;
            }
        }
        if (reads[i] > 0) {
            int s = 0;
            while (s < CMS_REG_SIZE && busy[s]) s++;
            if (s >= CMS_REG_SIZE) {  // cannot happen: the original used <= REG_SIZE live registers
                err = "register re-allocation exceeded REG_SIZE";
                return CMS_E_SCHEDULE;
            }
            busy[s] = 1;
            slot_of[i] = s;
            nslots = std::max(nslots, s + 1);
            code |= ((uint32_t)s << 16) | kDstLive;
        }
        nd.code = code;
        nd.w1 = weight_of(d1[i]);
        nd.w2 = weight_of(d2[i]);
        out.nodes[i] = nd;
    }
    out.nslots = nslots;

    // Leaf chunks for LDS staging.
    out.chunks.clear();
    {
        WalkChunk c{0, 0, 0, 0};
        for (int i = 0; i < order_n; i++) {
            int nl = ((out.nodes[i].code & kC1Leaf) ? 1 : 0) + ((out.nodes[i].code & kC2Leaf) ? 1 : 0);
            int lp = out.nodes[i].leafpos;
            if (lp + nl - c.leaf_begin > leaf_chunk) {
                c.node_end = i;
                c.leaf_end = lp;
                out.chunks.push_back(c);
                c = WalkChunk{i, i, lp, lp};
            }
        }
        c.node_end = order_n;
        c.leaf_end = (int32_t)out.leaf_stream.size();
        if (order_n > 0) out.chunks.push_back(c);
    }

    // Embedding entries: beta of a child edge is 1 - w when the parent's value
    // is consumed later, 1 when the parent is a root (its value is never read).
    out.ent_child.resize(2 * (size_t)order_n);
    out.ent_beta.resize(2 * (size_t)order_n);
    out.nonneg_beta = true;
    for (int i = 0; i < order_n; i++) {
        int lp = out.nodes[i].leafpos;
        for (int j = 0; j < 2; j++) {
            int32_t c = (j == 0) ? o1[i] : o2[i];
            size_t e = 2 * (size_t)i + j;
            if (c >= 0)
                out.ent_child[e] = lp++;
            else
                out.ent_child[e] = -1 - child_val[e];
            double w = (double)((j == 0) ? out.nodes[i].w1 : out.nodes[i].w2);
            double beta = (reads[i] > 0) ? 1.0 - w : 1.0;
            out.ent_beta[e] = beta;
            if (beta < 0) out.nonneg_beta = false;
        }
    }
    return CMS_OK;
}

} // namespace cms

extern "C" int cms_validate_tree(int leaf_n, int order_n, const int32_t *o1, const float *d1, const int32_t *o2,
                                 const float *d2, const int32_t *od, int *nslots, int *is_tree, int *nonneg_beta) {
    cms::CompiledTree t;
    std::string err;
    int rc = cms::compile_tree(leaf_n, order_n, o1, d1, o2, d2, od, 32, t, err);
    if (rc) {
        cms::set_global_error(err);
        return rc;
    }
    if (nslots) *nslots = t.nslots;
    if (is_tree) *is_tree = t.is_tree ? 1 : 0;
    if (nonneg_beta) *nonneg_beta = t.nonneg_beta ? 1 : 0;
    return CMS_OK;
}

namespace cms {

void plan_segments(CompiledTree &t, int seg_max) {
    const int n = t.order_n;
    t.seg_first.clear();
    t.seg_last.clear();
    t.seg_out.clear();
    t.top_nodes.clear();
    t.top_src.clear();
    std::vector<int32_t> size(n, 1), first(n), consumer(n, -1);
    std::vector<char> contig(n, 1);
    for (int i = 0; i < n; i++) {
        first[i] = i;
        for (int j = 0; j < 2; j++) {
            const int32_t c = t.ent_child[2 * (size_t)i + j];
            if (c >= 0) continue;
            const int p = -1 - c;
            size[i] += size[p];
            first[i] = std::min(first[i], first[p]);
            if (!contig[p]) contig[i] = 0;
            consumer[p] = i;
        }
        if (i - first[i] + 1 != size[i]) contig[i] = 0;
    }
    auto fits = [&](int i) { return t.is_tree && contig[i] && size[i] <= seg_max; };
    std::vector<int32_t> seg_of(n, -1);  // segment whose root is node i
    std::vector<char> in_seg(n, 0);
    for (int i = 0; i < n; i++) {
        if (!fits(i)) continue;
        if (consumer[i] >= 0 && fits(consumer[i])) continue;  // not maximal
        seg_of[i] = (int32_t)t.seg_first.size();
        t.seg_first.push_back(first[i]);
        t.seg_last.push_back(i);
        t.seg_out.push_back(consumer[i] >= 0 ? 1 : 0);
        for (int k = first[i]; k <= i; k++) in_seg[k] = 1;
    }
    for (int i = 0; i < n; i++) {
        if (in_seg[i]) continue;
        t.top_nodes.push_back(i);
        for (int j = 0; j < 2; j++) {
            const int32_t c = t.ent_child[2 * (size_t)i + j];
            t.top_src.push_back(c < 0 ? seg_of[-1 - c] : -1);
        }
    }
}

}  // namespace cms
