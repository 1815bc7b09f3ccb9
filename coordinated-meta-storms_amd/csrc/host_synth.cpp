// Seeded synthetic inputs: tree database (Yule topology, Sethi-Ullman
// order.txt), copy numbers and OTU count samples (Zipf popularity,
// Gamma-Poisson negative binomial counts). The reference ships no data
// (its databases are LFS pointers and its example table is missing), so every
// workload here is generated, reproducibly from a 64-bit seed.
//
// Published methods used: splitmix64 / xoshiro256** (Blackman & Vigna),
// Marsaglia & Tsang (2000) gamma variates, Knuth's multiplication method and
// Hoermann's PTRS (1993) for Poisson variates.
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cms_internal.h"

using cms::set_global_error;

namespace {

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }     // [0, 1)
    double uniform_pos() { return ((double)(next() >> 11) + 0.5) * 0x1.0p-53; }  // (0, 1)
    uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n); }
    double normal() {
        double u1 = uniform_pos(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // Marsaglia-Tsang; shape < 1 via Gamma(a+1) * U^(1/a).
    double gamma(double a, double scale) {
        if (a < 1.0) {
            double g = gamma(a + 1.0, 1.0);
            return scale * g * std::pow(uniform_pos(), 1.0 / a);
        }
        double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x = normal(), v = 1.0 + c * x;
            if (v <= 0) continue;
            v = v * v * v;
            double u = uniform_pos();
            if (std::log(u) < 0.5 * x * x + d - d * v + d * std::log(v)) return scale * d * v;
        }
    }
    int64_t poisson(double lam) {
        if (lam <= 0) return 0;
        if (lam < 10.0) {
            double L = std::exp(-lam), p = 1.0;
            int64_t k = 0;
            do { k++; p *= uniform(); } while (p > L);
            return k - 1;
        }
        // PTRS (Hoermann 1993)
        double slam = std::sqrt(lam), loglam = std::log(lam);
        double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
        double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
        for (;;) {
            double U = uniform() - 0.5, V = uniform();
            double us = 0.5 - std::fabs(U);
            int64_t k = (int64_t)std::floor((2.0 * a / us + b) * U + lam + 0.43);
            if (us >= 0.07 && V <= vr) return k;
            if (k < 0 || (us < 0.013 && V > us)) continue;
            if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
                -lam + (double)k * loglam - std::lgamma((double)k + 1.0))
                return k;
        }
    }
};

// Round to 6 decimals through text, the precision of the written order.txt;
// the float read back is the canonical branch length everywhere.
inline float round6(double x) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.6f", x);
    return strtof(buf, nullptr);
}

bool mkdirs(const std::string &path) {
    std::string cur;
    for (size_t i = 0; i < path.size(); i++) {
        cur += path[i];
        if ((path[i] == '/' && i > 0) || i + 1 == path.size()) {
            if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
        }
    }
    return true;
}

}  // namespace

extern "C" {

int cms_synth_tree(uint64_t seed, int leaf_n, double bl_mean, double bl_max, int32_t *o1, float *d1,
                   int32_t *o2, float *d2, int32_t *od, int *max_slots) {
    if (leaf_n < 2 || !o1 || !d1 || !o2 || !d2 || !od || bl_mean < 0 || bl_max < 0) {
        set_global_error("cms_synth_tree: need leaf_n >= 2 and output arrays");
        return CMS_E_INVAL;
    }
    Rng rng(seed);
    const int total = 2 * leaf_n - 1;
    // Yule process: split a uniformly chosen current leaf until leaf_n leaves.
    std::vector<int32_t> left(total, -1), right(total, -1);
    std::vector<int32_t> leaves;
    leaves.reserve(leaf_n);
    int next = 1;
    left[0] = next++;
    right[0] = next++;
    leaves.push_back(1);
    leaves.push_back(2);
    while ((int)leaves.size() < leaf_n) {
        size_t k = rng.below(leaves.size());
        int v = leaves[k];
        left[v] = next++;
        right[v] = next++;
        leaves[k] = left[v];
        leaves.push_back(right[v]);
    }
    // Leaf indices (id.txt order) are a random permutation of tree leaves.
    std::vector<int32_t> leaf_index(total, -1);
    {
        std::vector<int32_t> perm(leaf_n);
        for (int i = 0; i < leaf_n; i++) perm[i] = i;
        for (int i = leaf_n - 1; i > 0; i--) std::swap(perm[i], perm[rng.below((uint64_t)i + 1)]);
        for (int i = 0; i < leaf_n; i++) leaf_index[leaves[i]] = perm[i];
    }
    std::vector<float> bl(total, 0.0f);
    for (int v = 1; v < total; v++) {
        double x = -bl_mean * std::log(1.0 - rng.uniform());
        x = std::min(std::max(x, 0.0), bl_max);
        bl[v] = round6(x);
    }
    // Sethi-Ullman register need; children always have larger ids.
    std::vector<int32_t> need(total, 0);
    for (int v = total - 1; v >= 0; v--) {
        if (left[v] < 0) continue;
        int a = left[v], b = right[v];
        if (need[b] > need[a]) std::swap(a, b);
        int n = std::max(need[a], (left[a] >= 0 ? 1 : 0) + need[b]);
        need[v] = std::max(n, 1);
        left[v] = a;  // evaluate the heavier child first
        right[v] = b;
    }
    // Post-order emission with lowest-free register allocation.
    std::vector<int32_t> slot(total, -1);
    std::vector<char> busy(CMS_REG_SIZE, 0);
    int nslots = 0, line = 0;
    std::vector<std::pair<int32_t, int>> stack;
    stack.emplace_back(0, 0);
    while (!stack.empty()) {
        auto &top = stack.back();
        int v = top.first;
        if (left[v] < 0) { stack.pop_back(); continue; }
        if (top.second == 0) { top.second = 1; stack.emplace_back(left[v], 0); continue; }
        if (top.second == 1) { top.second = 2; stack.emplace_back(right[v], 0); continue; }
        stack.pop_back();
        int ch[2] = {left[v], right[v]};
        int32_t code[2];
        for (int j = 0; j < 2; j++) {
            int c = ch[j];
            if (left[c] < 0) code[j] = leaf_index[c];
            else {
                code[j] = -(slot[c] + 1);
                busy[slot[c]] = 0;
            }
        }
        int s = 0;
        while (s < CMS_REG_SIZE && busy[s]) s++;
        if (s >= CMS_REG_SIZE) {
            set_global_error("cms_synth_tree: register need exceeds REG_SIZE");
            return CMS_E_SCHEDULE;
        }
        busy[s] = 1;
        slot[v] = s;
        nslots = std::max(nslots, s + 1);
        o1[line] = code[0];
        d1[line] = bl[ch[0]];
        o2[line] = code[1];
        d2[line] = bl[ch[1]];
        od[line] = -(s + 1);
        line++;
    }
    if (max_slots) *max_slots = nslots;
    return CMS_OK;
}

int cms_synth_samples(const cms_synth_samples_opts *o, int64_t **rowptr_out, int32_t **col_out, float **cnt_out) {
    if (!o || !rowptr_out || !col_out || !cnt_out || o->n_samples < 0 || o->leaf_n < 1 || o->nnz_lo < 1 ||
        o->nnz_hi < o->nnz_lo || o->nb_mean <= 0 || o->nb_dispersion <= 0) {
        set_global_error("cms_synth_samples: bad options");
        return CMS_E_INVAL;
    }
    const int N = o->n_samples, L = o->leaf_n;
    Rng base(o->seed);
    // Popularity: a random rank order of leaves, weight (rank+1)^-s.
    std::vector<int32_t> rank_leaf(L);
    for (int i = 0; i < L; i++) rank_leaf[i] = i;
    for (int i = L - 1; i > 0; i--) std::swap(rank_leaf[i], rank_leaf[base.below((uint64_t)i + 1)]);
    std::vector<double> cum(L);
    double acc = 0;
    for (int r = 0; r < L; r++) { acc += std::pow((double)(r + 1), -o->zipf_s); cum[r] = acc; }
    const double wsum = acc;
    // Per-sample nnz first (sequential, deterministic), then parallel fill.
    std::vector<int32_t> nnz(N);
    for (int s = 0; s < N; s++) {
        int k;
        if (o->nnz_log_uniform) {
            double lo = std::log((double)o->nnz_lo), hi = std::log((double)o->nnz_hi);
            k = (int)std::lround(std::exp(lo + (hi - lo) * base.uniform()));
        } else {
            k = o->nnz_lo + (int)base.below((uint64_t)(o->nnz_hi - o->nnz_lo + 1));
        }
        nnz[s] = std::min(std::max(k, 1), L);
    }
    int64_t *rowptr = (int64_t *)malloc(sizeof(int64_t) * ((size_t)N + 1));
    if (!rowptr) { set_global_error("cms_synth_samples: out of memory"); return CMS_E_NOMEM; }
    rowptr[0] = 0;
    for (int s = 0; s < N; s++) rowptr[s + 1] = rowptr[s] + nnz[s];
    const size_t total = (size_t)rowptr[N];
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (total ? total : 1));
    float *cnt = (float *)malloc(sizeof(float) * (total ? total : 1));
    if (!col || !cnt) {
        free(rowptr); free(col); free(cnt);
        set_global_error("cms_synth_samples: out of memory");
        return CMS_E_NOMEM;
    }
    const uint64_t sample_seed = base.next();
#pragma omp parallel
    {
        std::vector<int32_t> stamp(L, -1);
        std::vector<int32_t> picked;
#pragma omp for schedule(dynamic, 16)
        for (int s = 0; s < N; s++) {
            Rng rng(sample_seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(s + 1)));
            picked.clear();
            const int k = nnz[s];
            if (k == L) {
                for (int l = 0; l < L; l++) picked.push_back(l);
            } else {
                while ((int)picked.size() < k) {
                    double u = rng.uniform() * wsum;
                    int r = (int)(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin());
                    if (r >= L) r = L - 1;
                    int32_t leaf = rank_leaf[r];
                    if (stamp[leaf] == s) continue;
                    stamp[leaf] = s;
                    picked.push_back(leaf);
                }
            }
            std::sort(picked.begin(), picked.end());
            const double shape = o->nb_dispersion, scale = o->nb_mean / o->nb_dispersion;
            int64_t off = rowptr[s];
            for (int i = 0; i < k; i++) {
                int64_t c = 0;
                while (c == 0) c = rng.poisson(rng.gamma(shape, scale));
                col[off + i] = picked[i];
                cnt[off + i] = (float)c;
            }
        }
    }
    *rowptr_out = rowptr;
    *col_out = col;
    *cnt_out = cnt;
    return CMS_OK;
}

int cms_synth_copy_numbers(uint64_t seed, int leaf_n, int cp_max, float *cp) {
    if (leaf_n < 0 || cp_max < 1 || (leaf_n && !cp)) {
        set_global_error("cms_synth_copy_numbers: bad argument");
        return CMS_E_INVAL;
    }
    Rng rng(seed);
    for (int i = 0; i < leaf_n; i++) cp[i] = (float)(1 + (int)rng.below((uint64_t)cp_max));
    return CMS_OK;
}

// Load_abd arithmetic (comp.h:252-275) over CSR rows whose columns are leaf
// indices in ascending (column) order.
int cms_counts_to_abundance_csr(int n, const int64_t *rowptr, const int32_t *col, const float *cnt,
                                const float *cp, float *abd_val) {
    if (n < 0 || (n && (!rowptr || !col || !cnt || !abd_val))) {
        set_global_error("cms_counts_to_abundance_csr: bad argument");
        return CMS_E_INVAL;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < n; s++) {
        float total = 0;
        for (int64_t k = rowptr[s]; k < rowptr[s + 1]; k++) {
            float c = cp ? cp[col[k]] : 1.0f;
            float x = cnt[k] > 0 ? cnt[k] / c : 0.0f;
            abd_val[k] = x;
            total += x;
        }
        total /= 100.0;
        for (int64_t k = rowptr[s]; k < rowptr[s + 1]; k++) abd_val[k] = abd_val[k] / total;
    }
    return CMS_OK;
}

int cms_write_db(const char *root, const char *name, int leaf_n, int order_n, const int32_t *o1, const float *d1,
                 const int32_t *o2, const float *d2, const int32_t *od, const float *cp) {
    if (!root || !name || leaf_n < 0 || order_n < 0 || (order_n && (!o1 || !d1 || !o2 || !d2 || !od))) {
        set_global_error("cms_write_db: bad argument");
        return CMS_E_INVAL;
    }
    std::string base = std::string(root) + "/databases/";
    std::string dir = base + name + "/";
    if (!mkdirs(dir + "tree")) {
        set_global_error("cms_write_db: cannot create " + dir + "tree");
        return CMS_E_IO;
    }
    FILE *f = fopen((base + "db.config").c_str(), "w");
    if (!f) { set_global_error("cms_write_db: cannot write db.config"); return CMS_E_IO; }
    fprintf(f, "#Synthetic database\n#Database id\nG\n#domain\n0\n#name\n%s\n#description\nSynthetic tree (%d leaves)\n"
               "#taxonomy\nN\n#tree\nY\n#copy number\n%s\n#function\nN\n",
            name, leaf_n, cp ? "Y" : "N");
    fclose(f);
    f = fopen((dir + "tree/id.txt").c_str(), "w");
    if (!f) { set_global_error("cms_write_db: cannot write id.txt"); return CMS_E_IO; }
    for (int i = 0; i < leaf_n; i++) fprintf(f, "T%d\n", i);
    fclose(f);
    f = fopen((dir + "tree/order.txt").c_str(), "w");
    if (!f) { set_global_error("cms_write_db: cannot write order.txt"); return CMS_E_IO; }
    for (int i = 0; i < order_n; i++)
        fprintf(f, "%d\t%.6f\t%d\t%.6f\t%d\n", o1[i], (double)d1[i], o2[i], (double)d2[i], od[i]);
    fclose(f);
    if (cp) {
        f = fopen((dir + "copy_number.txt").c_str(), "w");
        if (!f) { set_global_error("cms_write_db: cannot write copy_number.txt"); return CMS_E_IO; }
        fprintf(f, "ID\tcopy_number\n");
        for (int i = 0; i < leaf_n; i++) fprintf(f, "T%d\t%g\n", i, (double)cp[i]);
        fclose(f);
    }
    return CMS_OK;
}

int cms_write_table(const char *path, int n, int leaf_n, const int64_t *rowptr, const int32_t *col,
                    const float *cnt) {
    if (!path || n < 0 || leaf_n < 0 || (n && (!rowptr || !col || !cnt))) {
        set_global_error("cms_write_table: bad argument");
        return CMS_E_INVAL;
    }
    FILE *f = fopen(path, "w");
    if (!f) { set_global_error(std::string("cms_write_table: cannot write ") + path); return CMS_E_IO; }
    std::string s = "SampleID";
    for (int j = 0; j < leaf_n; j++) { s += "\tT"; s += std::to_string(j); }
    s += '\n';
    fwrite(s.data(), 1, s.size(), f);
    std::vector<float> row(leaf_n);
    for (int i = 0; i < n; i++) {
        std::fill(row.begin(), row.end(), 0.0f);
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) row[col[k]] = cnt[k];
        s = "S" + std::to_string(i);
        char buf[64];
        for (int j = 0; j < leaf_n; j++) {
            int m = snprintf(buf, sizeof buf, "\t%g", (double)row[j]);
            s.append(buf, (size_t)m);
        }
        s += '\n';
        fwrite(s.data(), 1, s.size(), f);
    }
    fclose(f);
    return CMS_OK;
}

int cms_write_table_csr(const char *path, int n, int leaf_n, const int64_t *rowptr, const int32_t *col,
                        const float *cnt) {
    if (!path || n < 0 || leaf_n < 0 || (n && (!rowptr || !col || !cnt))) {
        set_global_error("cms_write_table_csr: bad argument");
        return CMS_E_INVAL;
    }
    std::vector<std::string> fn((size_t)leaf_n), sn((size_t)n);
    std::vector<const char *> fp((size_t)leaf_n), sp((size_t)n);
    for (int j = 0; j < leaf_n; j++) {
        fn[j] = "T" + std::to_string(j);
        fp[j] = fn[j].c_str();
    }
    for (int i = 0; i < n; i++) {
        sn[i] = "S" + std::to_string(i);
        sp[i] = sn[i].c_str();
    }
    // counts as the text table prints them (%g), so both forms hold one table
    std::vector<float> c((size_t)std::max<int64_t>(n ? rowptr[n] : 0, 1));
    for (int64_t k = 0; n && k < rowptr[n]; k++) {
        char buf[64];
        snprintf(buf, sizeof buf, "%g", (double)cnt[k]);
        c[(size_t)k] = strtof(buf, nullptr);
    }
    cms_table *t = nullptr;
    int rc = cms_table_from_csr(n, leaf_n, sp.data(), fp.data(), rowptr, col, c.data(), &t);
    if (rc) return rc;
    rc = cms_table_save_csr(t, path);
    cms_table_free(t);
    return rc;
}

}  // extern "C"
