// Host loader: reference database configuration, tree files, copy numbers,
// OTU count tables, abundance normalisation and the .dist writer.
//
// Behaviour mirrors the reference loaders (cited per function); the code is a
// fresh implementation with a tokenizer instead of iostreams so that tables of
// 10^4 x 10^5 entries load in seconds, and with errors returned instead of
// exit().
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <thread>
#include <vector>

// the reference's hash container (hash.h), for the iteration order of the
// sample-file loader
#define _GLIBCXX_PERMIT_BACKWARD_HASH
#include <ext/hash_map>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "cms_internal.h"

using cms::set_global_error;

struct cms_db {
    std::string root, name, path;
    bool has_tree = false, has_cp = false;
    std::vector<std::string> ids;                 // tree/id.txt, one per non-empty line
    std::unordered_map<std::string, int> id_hash; // last index wins (comp.h:126-127)
    std::vector<int32_t> o1, o2, od;
    std::vector<float> d1, d2;
    std::unordered_map<std::string, float> cp;    // first occurrence wins (db.h:269-271)
};

struct cms_table {
    std::vector<std::string> features;
    std::vector<std::string> samples;
    std::vector<float> abd;  // n x f row-major (text tables)
    int skipped = 0;
    // sparse tables: row s = (column, count) pairs, columns ascending
    bool sparse = false;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> cols;
    std::vector<float> cnts;
    // visit row s's columns with a count in ascending column order: fn(j, v)
    template <class Fn>
    void for_row(size_t s, Fn fn) const {
        if (sparse) {
            for (int64_t k = rowptr[s]; k < rowptr[s + 1]; k++) fn((size_t)cols[(size_t)k], cnts[(size_t)k]);
        } else {
            const size_t F = features.size();
            const float *row = abd.data() + s * F;
            for (size_t j = 0; j < F; j++) fn(j, row[j]);
        }
    }
};

namespace {

bool read_file(const std::string &path, std::string &buf) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    buf.clear();
    char tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, n);
    fclose(f);
    return true;
}

// Split into lines like std::getline (the '\n' is dropped, a final line
// without '\n' still counts, a trailing '\n' does not create an empty line).
void split_lines(const std::string &buf, std::vector<std::pair<size_t, size_t>> &lines) {
    lines.clear();
    size_t start = 0, n = buf.size();
    while (start < n) {
        const void *p = memchr(buf.data() + start, '\n', n - start);
        size_t end = p ? (size_t)((const char *)p - buf.data()) : n;
        lines.emplace_back(start, end - start);
        start = end + 1;
    }
}

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Next whitespace-delimited token in [p, e); returns false at end.
inline bool next_token(const char *&p, const char *e, const char *&ts, const char *&te) {
    while (p < e && is_ws(*p)) p++;
    if (p >= e) return false;
    ts = p;
    while (p < e && !is_ws(*p)) p++;
    te = p;
    return true;
}

// Decimal float token as read by `stream >> float`: the longest prefix of the
// form [sign]digits[.digits][(e|E)[sign]digits]. Returns 1 on a full token,
// 2 when a valid prefix is followed by other characters (the value counts and
// the next extraction fails), 0 when no number can be read.
int parse_float_token(const char *ts, const char *te, float &v) {
    const char *p = ts;
    if (p < te && (*p == '+' || *p == '-')) p++;
    const char *ds = p;
    while (p < te && *p >= '0' && *p <= '9') p++;
    bool digits = p > ds;
    if (p < te && *p == '.') {
        p++;
        const char *fs = p;
        while (p < te && *p >= '0' && *p <= '9') p++;
        digits = digits || p > fs;
    }
    if (!digits) return 0;
    if (p < te && (*p == 'e' || *p == 'E')) {
        const char *q = p + 1;
        if (q < te && (*q == '+' || *q == '-')) q++;
        const char *es = q;
        while (q < te && *q >= '0' && *q <= '9') q++;
        if (q == es) return 0;  // dangling exponent: the stream conversion fails
        p = q;
    }
    char tmp[128];
    size_t len = (size_t)(p - ts);
    std::string big;
    const char *s;
    if (len < sizeof tmp) {
        memcpy(tmp, ts, len);
        tmp[len] = 0;
        s = tmp;
    } else {
        big.assign(ts, len);
        s = big.c_str();
    }
    v = strtof(s, nullptr);
    return p == te ? 1 : 2;
}

// `stream >> int` on a token (order.txt fields).
bool parse_int_token(const char *ts, const char *te, int32_t &v) {
    const char *p = ts;
    bool neg = false;
    if (p < te && (*p == '+' || *p == '-')) { neg = (*p == '-'); p++; }
    if (p >= te || *p < '0' || *p > '9') return false;
    long long x = 0;
    while (p < te && *p >= '0' && *p <= '9') { x = x * 10 + (*p - '0'); if (x > 4000000000LL) x = 4000000000LL; p++; }
    x = neg ? -x : x;
    if (x > INT32_MAX) x = INT32_MAX;
    if (x < INT32_MIN) x = INT32_MIN;
    v = (int32_t)x;
    return true;
}

// Check_OTU (utility.h:181-190): strip an "otu_"/"OTU_" prefix when longer than 4.
std::string check_otu(const std::string &otu) {
    if (otu.size() > 4) {
        if (otu.compare(0, 4, "otu_") == 0 || otu.compare(0, 4, "OTU_") == 0) return otu.substr(4);
    }
    return otu;
}

}  // namespace

extern "C" {

// _PMDB::Load_config (db.h:287-358) + Set_DB (db.h:32-67) + _Comp_Tree::Init
// (comp.h:122-192) + _PMDB::Load_Copy_Number (db.h:251-280).
int cms_db_open(const char *root, char db_letter, cms_db **out) {
    if (!out) { set_global_error("cms_db_open: out is NULL"); return CMS_E_INVAL; }
    *out = nullptr;
    std::string r;
    if (root) r = root;
    else {
        const char *env = getenv("ParallelMETA");
        if (!env) {
            set_global_error("Please set the environment variable \"ParallelMETA\" to the directory");
            return CMS_E_INVAL;
        }
        r = env;
    }
    std::string buf;
    if (!read_file(r + "/databases/db.config", buf)) {
        set_global_error("Cannot open reference database configuration: " + r + "/databases/db.config");
        return CMS_E_IO;
    }
    std::vector<std::pair<size_t, size_t>> lines;
    split_lines(buf, lines);
    std::map<char, std::vector<std::string>> config;
    size_t li = 0;
    while (li < lines.size()) {
        std::string first = buf.substr(lines[li].first, lines[li].second);
        li++;
        if (first.empty() || first[0] == '#') continue;
        std::vector<std::string> entry{first};
        for (int k = 0; k < 7; k++) {
            while (li < lines.size()) {
                std::string a = buf.substr(lines[li].first, lines[li].second);
                li++;
                if (a.empty() || a[0] == '#') continue;
                entry.push_back(a);
                break;
            }
        }
        if (entry.size() < 8) break;          // incomplete entry ends parsing
        if (config.count(entry[0][0])) continue; // duplicate id keeps the first
        config[entry[0][0]] = entry;
    }
    if (config.empty()) {
        set_global_error("Cannot open reference database configuration (no complete entry)");
        return CMS_E_IO;
    }
    // _Comp_Tree's member _PMDB is default-constructed with 'G' first
    // (db.h:24-26), which indexes the 'G' entry unconditionally.
    if (!config.count('G')) {
        set_global_error("db.config has no entry 'G' (the reference dereferences it unconditionally)");
        return CMS_E_IO;
    }
    char key = config.count(db_letter) ? db_letter : 'G';
    const std::vector<std::string> &cfg = config[key];
    auto yes = [](const std::string &s) { return !s.empty() && (s[0] == 'Y' || s[0] == 'y'); };
    cms_db *db = new cms_db();
    db->root = r;
    db->name = cfg[2];
    db->has_tree = yes(cfg[5]);
    db->has_cp = yes(cfg[6]);
    db->path = r + "/databases/" + db->name + "/";

    // tree/id.txt (comp.h:139-159)
    if (!read_file(db->path + "tree/id.txt", buf)) {
        set_global_error("Cannot open file : " + db->path + "tree/id.txt");
        delete db;
        return CMS_E_IO;
    }
    split_lines(buf, lines);
    for (auto &l : lines)
        if (l.second) db->ids.emplace_back(buf.substr(l.first, l.second));
    for (size_t i = 0; i < db->ids.size(); i++) db->id_hash[db->ids[i]] = (int)i;

    // tree/order.txt (comp.h:161-192)
    if (db->has_tree) {
        if (!read_file(db->path + "tree/order.txt", buf)) {
            set_global_error("Cannot open file : " + db->path + "tree/order.txt");
            delete db;
            return CMS_E_IO;
        }
        split_lines(buf, lines);
        for (auto &l : lines) {
            if (!l.second) continue;
            const char *p = buf.data() + l.first, *e = p + l.second, *ts, *te;
            int32_t a = 0, b = 0, d = 0;
            float x = 0, y = 0;
            bool ok = next_token(p, e, ts, te) && parse_int_token(ts, te, a);
            ok = ok && next_token(p, e, ts, te) && parse_float_token(ts, te, x) == 1;
            ok = ok && next_token(p, e, ts, te) && parse_int_token(ts, te, b);
            ok = ok && next_token(p, e, ts, te) && parse_float_token(ts, te, y) == 1;
            ok = ok && next_token(p, e, ts, te) && parse_int_token(ts, te, d);
            (void)ok;  // fields after the first failure stay 0, as with the stream
            db->o1.push_back(a);
            db->d1.push_back(x);
            db->o2.push_back(b);
            db->d2.push_back(y);
            db->od.push_back(d);
        }
    }

    // copy_number.txt (db.h:251-280): header skipped, "id cp" per line.
    if (db->has_cp) {
        if (read_file(db->path + "copy_number.txt", buf)) {
            split_lines(buf, lines);
            for (size_t k = 1; k < lines.size(); k++) {
                const char *p = buf.data() + lines[k].first, *e = p + lines[k].second, *ts, *te;
                if (!next_token(p, e, ts, te)) continue;
                std::string id(ts, te);
                float v;
                if (!next_token(p, e, ts, te) || parse_float_token(ts, te, v) == 0) continue;
                if (!db->cp.count(id)) db->cp[id] = v;
            }
        } else {
            fprintf(stderr, "Warning: Cannot open copy number table : %scopy_number.txt, copy number normalization is disabled\n",
                    db->path.c_str());
        }
    }
    *out = db;
    return CMS_OK;
}

void cms_db_close(cms_db *db) { delete db; }
int cms_db_leaf_n(const cms_db *db) { return db ? (int)db->ids.size() : CMS_E_INVAL; }
int cms_db_order_n(const cms_db *db) { return db ? (int)db->o1.size() : CMS_E_INVAL; }
int cms_db_has_tree(const cms_db *db) { return db ? (int)db->has_tree : CMS_E_INVAL; }
int cms_db_has_cp(const cms_db *db) { return db ? (int)db->has_cp : CMS_E_INVAL; }

int cms_db_order(const cms_db *db, const int32_t **o1, const float **d1, const int32_t **o2,
                 const float **d2, const int32_t **od) {
    if (!db) return CMS_E_INVAL;
    if (o1) *o1 = db->o1.data();
    if (d1) *d1 = db->d1.data();
    if (o2) *o2 = db->o2.data();
    if (d2) *d2 = db->d2.data();
    if (od) *od = db->od.data();
    return CMS_OK;
}

const char *cms_db_leaf_name(const cms_db *db, int i) {
    if (!db || i < 0 || i >= (int)db->ids.size()) return "NULL";  // Get_Id (comp.h:92-96)
    return db->ids[i].c_str();
}

namespace {

// Sparse table file (cms_table_save_csr): little-endian
//   "CMSTCSR1", u64 samples, u64 features, u64 entries,
//   features x (u32 length, name bytes), samples x (u32 length, name bytes),
//   (samples + 1) x i64 row offsets, entries x i32 columns (ascending per
//   row), entries x f32 counts.
const char kCsrMagic[8] = {'C', 'M', 'S', 'T', 'C', 'S', 'R', '1'};

int parse_csr_table(const std::string &buf, const char *path, cms_table *t) {
    size_t at = 8;
    auto fail = [&](const char *why) {
        set_global_error(std::string("sparse table ") + path + ": " + why);
        return CMS_E_IO;
    };
    auto get = [&](void *dst, size_t bytes) {
        if (bytes > buf.size() - at) return false;
        memcpy(dst, buf.data() + at, bytes);
        at += bytes;
        return true;
    };
    uint64_t n = 0, f = 0, nnz = 0;
    if (!get(&n, 8) || !get(&f, 8) || !get(&nnz, 8)) return fail("truncated header");
    if (n > (uint64_t)INT32_MAX || f > (uint64_t)INT32_MAX || nnz > n * f) return fail("sizes out of range");
    auto names = [&](uint64_t count, std::vector<std::string> &v) {
        v.resize(count);
        for (uint64_t i = 0; i < count; i++) {
            uint32_t len = 0;
            if (!get(&len, 4) || len > buf.size() - at) return false;
            v[i].assign(buf.data() + at, len);
            at += len;
        }
        return true;
    };
    if (!names(f, t->features) || !names(n, t->samples)) return fail("truncated names");
    t->rowptr.resize(n + 1);
    t->cols.resize(nnz);
    t->cnts.resize(nnz);
    if (!get(t->rowptr.data(), 8 * (n + 1)) || !get(t->cols.data(), 4 * nnz) || !get(t->cnts.data(), 4 * nnz))
        return fail("truncated arrays");
    if (at != buf.size()) return fail("trailing bytes");
    if (t->rowptr[0] != 0 || t->rowptr[n] != (int64_t)nnz) return fail("row offsets do not cover the entries");
    for (uint64_t i = 0; i < n; i++) {
        if (t->rowptr[i + 1] < t->rowptr[i]) return fail("row offsets decrease");
        for (int64_t k = t->rowptr[i]; k < t->rowptr[i + 1]; k++) {
            if (t->cols[k] < 0 || (uint64_t)t->cols[k] >= f) return fail("column out of range");
            if (k > t->rowptr[i] && t->cols[k] <= t->cols[k - 1]) return fail("columns not ascending in a row");
        }
    }
    t->sparse = true;
    return CMS_OK;
}

}  // namespace

// _Table_Format::Load_Table (table_format.h:110-165); a file that starts
// with the sparse-table magic is read as the same table in CSR form.
int cms_table_load(const char *path, cms_table **out) {
    if (!path || !out) { set_global_error("cms_table_load: null argument"); return CMS_E_INVAL; }
    *out = nullptr;
    std::string buf;
    if (!read_file(path, buf)) {
        set_global_error(std::string("Cannot open file : ") + path);
        return CMS_E_IO;
    }
    if (buf.size() >= 8 && memcmp(buf.data(), kCsrMagic, 8) == 0) {
        cms_table *t = new cms_table();
        const int rc = parse_csr_table(buf, path, t);
        if (rc) {
            delete t;
            return rc;
        }
        *out = t;
        return CMS_OK;
    }
    std::vector<std::pair<size_t, size_t>> lines;
    split_lines(buf, lines);
    cms_table *t = new cms_table();
    if (!lines.empty()) {
        const char *p = buf.data() + lines[0].first, *e = p + lines[0].second, *ts, *te;
        if (next_token(p, e, ts, te))  // title token ignored
            while (next_token(p, e, ts, te)) t->features.emplace_back(ts, te);
    }
    const size_t F = t->features.size();
    const size_t nl = lines.size() > 0 ? lines.size() - 1 : 0;
    // Parse rows in parallel, keep file order.
    std::vector<std::string> names(nl);
    std::vector<char> ok(nl, 0);
    std::vector<float> vals(nl * F);
#pragma omp parallel for schedule(dynamic, 16)
    for (long r = 0; r < (long)nl; r++) {
        const auto &l = lines[r + 1];
        const char *p = buf.data() + l.first, *e = p + l.second, *ts, *te;
        if (next_token(p, e, ts, te)) names[r].assign(ts, te);
        size_t cnt = 0;
        float *row = vals.data() + (size_t)r * F;
        while (next_token(p, e, ts, te)) {
            float v;
            int st = parse_float_token(ts, te, v);
            if (st == 0) break;
            if (cnt < F) row[cnt] = v;
            cnt++;
            if (st == 2) break;
        }
        ok[r] = (cnt == F);
    }
    for (size_t r = 0; r < nl; r++) {
        if (!ok[r]) {
            fprintf(stderr, "Error: Sample: %s does not have %zu features\n", names[r].c_str(), F);
            t->skipped++;
            continue;
        }
        t->samples.push_back(names[r]);
        t->abd.insert(t->abd.end(), vals.begin() + r * F, vals.begin() + (r + 1) * F);
    }
    *out = t;
    return CMS_OK;
}

int cms_table_from_dense(int n, int f, const char *const *sample_names,
                         const char *const *feature_names, const float *counts, cms_table **out) {
    if (!out || n < 0 || f < 0 || (n && !sample_names) || (f && !feature_names) || (n && f && !counts)) {
        set_global_error("cms_table_from_dense: bad argument");
        return CMS_E_INVAL;
    }
    cms_table *t = new cms_table();
    for (int j = 0; j < f; j++) t->features.emplace_back(feature_names[j]);
    for (int i = 0; i < n; i++) t->samples.emplace_back(sample_names[i]);
    t->abd.assign(counts, counts + (size_t)n * f);
    *out = t;
    return CMS_OK;
}

int cms_table_from_csr(int n, int f, const char *const *sample_names, const char *const *feature_names,
                       const int64_t *rowptr, const int32_t *col, const float *counts, cms_table **out) {
    if (!out || n < 0 || f < 0 || (n && (!sample_names || !rowptr)) || (f && !feature_names)) {
        set_global_error("cms_table_from_csr: bad argument");
        return CMS_E_INVAL;
    }
    *out = nullptr;
    const int64_t nnz = n ? rowptr[n] : 0;
    if ((n && rowptr[0] != 0) || nnz < 0 || (nnz && (!col || !counts))) {
        set_global_error("cms_table_from_csr: bad row offsets");
        return CMS_E_INVAL;
    }
    for (int i = 0; i < n; i++) {
        if (rowptr[i + 1] < rowptr[i]) { set_global_error("cms_table_from_csr: row offsets decrease"); return CMS_E_INVAL; }
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++)
            if (col[k] < 0 || col[k] >= f || (k > rowptr[i] && col[k] <= col[k - 1])) {
                set_global_error("cms_table_from_csr: columns out of range or not ascending");
                return CMS_E_INVAL;
            }
    }
    cms_table *t = new cms_table();
    for (int j = 0; j < f; j++) t->features.emplace_back(feature_names[j]);
    for (int i = 0; i < n; i++) t->samples.emplace_back(sample_names[i]);
    t->sparse = true;
    t->rowptr.assign(rowptr, rowptr + (n ? n + 1 : 1));
    if (!n) t->rowptr[0] = 0;
    t->cols.assign(col, col + nnz);
    t->cnts.assign(counts, counts + nnz);
    *out = t;
    return CMS_OK;
}

int cms_table_save_csr(const cms_table *t, const char *path) {
    if (!t || !path) { set_global_error("cms_table_save_csr: null argument"); return CMS_E_INVAL; }
    const size_t n = t->samples.size(), f = t->features.size();
    std::vector<int64_t> rp(n + 1, 0);
    std::vector<int32_t> cj;
    std::vector<float> cv;
    for (size_t s = 0; s < n; s++) {
        t->for_row(s, [&](size_t j, float v) {
            if (v != 0.0f) {
                cj.push_back((int32_t)j);
                cv.push_back(v);
            }
        });
        rp[s + 1] = (int64_t)cj.size();
    }
    FILE *fp = fopen(path, "wb");
    if (!fp) { set_global_error(std::string("cms_table_save_csr: cannot write ") + path); return CMS_E_IO; }
    bool ok = fwrite(kCsrMagic, 1, 8, fp) == 8;
    const uint64_t hdr[3] = {(uint64_t)n, (uint64_t)f, (uint64_t)cj.size()};
    ok = ok && fwrite(hdr, 8, 3, fp) == 3;
    auto put_names = [&](const std::vector<std::string> &v) {
        for (const std::string &x : v) {
            const uint32_t len = (uint32_t)x.size();
            ok = ok && fwrite(&len, 4, 1, fp) == 1 && fwrite(x.data(), 1, len, fp) == len;
        }
    };
    put_names(t->features);
    put_names(t->samples);
    ok = ok && fwrite(rp.data(), 8, n + 1, fp) == n + 1;
    ok = ok && fwrite(cj.data(), 4, cj.size(), fp) == cj.size() && fwrite(cv.data(), 4, cv.size(), fp) == cv.size();
    if (fclose(fp) != 0) ok = false;
    if (!ok) { set_global_error(std::string("cms_table_save_csr: write failed: ") + path); return CMS_E_IO; }
    return CMS_OK;
}

int cms_table_is_sparse(const cms_table *t) { return t ? (int)t->sparse : CMS_E_INVAL; }

void cms_table_free(cms_table *t) { delete t; }
int cms_table_n(const cms_table *t) { return t ? (int)t->samples.size() : CMS_E_INVAL; }
int cms_table_f(const cms_table *t) { return t ? (int)t->features.size() : CMS_E_INVAL; }
const char *cms_table_sample_name(const cms_table *t, int i) {
    return (t && i >= 0 && i < (int)t->samples.size()) ? t->samples[i].c_str() : nullptr;
}
const char *cms_table_feature_name(const cms_table *t, int j) {
    return (t && j >= 0 && j < (int)t->features.size()) ? t->features[j].c_str() : nullptr;
}
int cms_table_skipped(const cms_table *t) { return t ? t->skipped : CMS_E_INVAL; }

// _Comp_Tree::Load_abd(_Table_Format*, ...) (comp.h:243-279):
//   for each column with value > 0: otu = Check_OTU(name); cp from the copy
//   number table when correcting; if otu is a tree leaf: x[otu] = v / cp
//   (a later duplicate overwrites x but both add to the total), total += x
//   in column order (FP32); total = (float)((double)total / 100.0);
//   abd[leaf] = x / total.
int cms_table_abundance(const cms_db *db, const cms_table *t, int cp_correct, float *abd, int *mapped) {
    if (!db || !t || !abd) { set_global_error("cms_table_abundance: null argument"); return CMS_E_INVAL; }
    const size_t L = db->ids.size(), F = t->features.size(), N = t->samples.size();
    std::vector<int32_t> leaf_of(F, -1);
    std::vector<float> cp_of(F, 1.0f);
    for (size_t j = 0; j < F; j++) {
        std::string a = check_otu(t->features[j]);
        if (cp_correct) {
            auto it = db->cp.find(a);
            if (it != db->cp.end()) cp_of[j] = it->second;
        }
        auto it = db->id_hash.find(a);
        if (it != db->id_hash.end()) leaf_of[j] = it->second;
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (long s = 0; s < (long)N; s++) {
        float *out = abd + (size_t)s * L;
        memset(out, 0, L * sizeof(float));
        std::vector<int32_t> touched;
        std::vector<char> seen;  // lazily sized
        float total = 0;
        t->for_row((size_t)s, [&](size_t j, float v) {
            if (v > 0 && leaf_of[j] >= 0) {
                float x = v / cp_of[j];
                int32_t l = leaf_of[j];
                out[l] = x;  // later duplicate overwrites
                touched.push_back(l);
                total += x;
            }
        });
        total /= 100.0;
        int m = 0;
        if (seen.size() < L) seen.assign(L, 0);
        for (int32_t l : touched) {
            if (seen[l]) continue;
            seen[l] = 1;
            out[l] = out[l] / total;
            m++;
        }
        if (mapped) mapped[s] = m;
    }
    return CMS_OK;
}

// The same abundances as cms_table_abundance, as CSR over leaves (leaf ids
// ascending per sample; a leaf named by two columns keeps the later column's
// value, as the dense form does). Arrays are malloc'ed; free with cms_free.
int cms_table_abundance_csr(const cms_db *db, const cms_table *t, int cp_correct, int64_t **rowptr_out,
                            int32_t **col_out, float **val_out) {
    if (!db || !t || !rowptr_out || !col_out || !val_out) {
        set_global_error("cms_table_abundance_csr: null argument");
        return CMS_E_INVAL;
    }
    const size_t L = db->ids.size(), F = t->features.size(), N = t->samples.size();
    std::vector<int32_t> leaf_of(F, -1);
    std::vector<float> cp_of(F, 1.0f);
    for (size_t j = 0; j < F; j++) {
        std::string a = check_otu(t->features[j]);
        if (cp_correct) {
            auto it = db->cp.find(a);
            if (it != db->cp.end()) cp_of[j] = it->second;
        }
        auto it = db->id_hash.find(a);
        if (it != db->id_hash.end()) leaf_of[j] = it->second;
    }
    std::vector<std::vector<std::pair<int32_t, float>>> rows(N);
#pragma omp parallel for schedule(dynamic, 8)
    for (long s = 0; s < (long)N; s++) {
        std::vector<std::pair<int32_t, float>> &r = rows[(size_t)s];
        float total = 0;
        t->for_row((size_t)s, [&](size_t j, float v) {
            if (v > 0 && leaf_of[j] >= 0) {
                const float x = v / cp_of[j];
                r.push_back({leaf_of[j], x});
                total += x;  // a duplicated leaf counts twice, as in Load_abd
            }
        });
        total /= 100.0;
        // last column wins for a duplicated leaf: stable sort by leaf, keep the last
        std::stable_sort(r.begin(), r.end(), [](const std::pair<int32_t, float> &a, const std::pair<int32_t, float> &b) {
            return a.first < b.first;
        });
        size_t k = 0;
        for (size_t i = 0; i < r.size(); i++) {
            if (k > 0 && r[k - 1].first == r[i].first) k--;
            r[k++] = r[i];
        }
        r.resize(k);
        for (auto &e : r) e.second = e.second / total;
    }
    size_t nnz = 0;
    for (auto &r : rows) nnz += r.size();
    int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (N + 1));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * std::max<size_t>(nnz, 1));
    float *val = (float *)malloc(sizeof(float) * std::max<size_t>(nnz, 1));
    if (!rp || !col || !val) {
        free(rp);
        free(col);
        free(val);
        set_global_error("cms_table_abundance_csr: out of host memory");
        return CMS_E_NOMEM;
    }
    rp[0] = 0;
    for (size_t s = 0; s < N; s++) {
        size_t at = (size_t)rp[s];
        for (auto &e : rows[s]) {
            col[at] = e.first;
            val[at] = e.second;
            at++;
        }
        rp[s + 1] = (int64_t)at;
    }
    (void)L;
    *rowptr_out = rp;
    *col_out = col;
    *val_out = val;
    return CMS_OK;
}

// Output_Matrix (cms_cuda.cu:1305-1342, comp_sam.cpp:184-219).
int cms_write_dist(const char *path, int n, const char *const *names, const float *sim, int is_sim) {
    if (!path || n < 0 || (n && !names) || (n > 1 && !sim)) {
        set_global_error("cms_write_dist: bad argument");
        return CMS_E_INVAL;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_global_error(std::string("Cannot open output file : ") + path);
        return CMS_E_IO;
    }
    std::string hdr;
    for (int i = 0; i < n; i++) { hdr += '\t'; hdr += names[i]; }
    hdr += '\n';
    fwrite(hdr.data(), 1, hdr.size(), f);
    // rows are formatted in parallel batches; a writer thread writes each
    // finished batch while the next one is formatted
    const int batch = 256;
    std::vector<std::string> rows[2] = {std::vector<std::string>(batch), std::vector<std::string>(batch)};
    // the batch's entries below the diagonal, row i's at low[(i - r0) * r1 +
    // j] for j < min(i, r1): they sit in column i of the packed rows j, so
    // they are gathered by 64 x 64 blocks (each packed row's segment read
    // contiguously, written transposed) instead of one strided read per entry
    std::vector<float> low((size_t)batch * n);
    std::thread writer;
    bool write_ok = true;
    int cur = 0;
    auto packed = [n](long a, long b) { return a * (long)n + b - (1 + a + 1) * (a + 1) / 2; };  // a < b
    for (int r0 = 0; r0 < n; r0 += batch) {
        int r1 = r0 + batch < n ? r0 + batch : n;
        std::vector<std::string> &rb = rows[cur];
        const int nj = (r1 + 63) / 64;
#pragma omp parallel for schedule(dynamic, 1)
        for (int jb = 0; jb < nj; jb++) {
            const int j0 = jb * 64, j1 = std::min(j0 + 64, r1);
            for (int i0 = r0; i0 < r1; i0 += 64) {
                const int i1 = std::min(i0 + 64, r1);
                for (int j = j0; j < j1; j++) {
                    const int ib = std::max(i0, j + 1);  // entries of packed row j at columns ib..i1-1
                    if (ib >= i1) continue;
                    const float *src = sim + packed(j, ib);
                    for (int i = ib; i < i1; i++) low[(size_t)(i - r0) * r1 + j] = src[i - ib];
                }
            }
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = r0; i < r1; i++) {
            // the row is written straight into its buffer: a tab and at most
            // 13 bytes per value (%.6g of a double, e.g. -1.23457e+308), and
            // a margin for the formatter's 32-byte scratch at the end
            std::string &s = rb[i - r0];
            const size_t nl = strlen(names[i]);
            s.resize(nl + (size_t)n * 14 + 40);
            char *o = &s[0];
            memcpy(o, names[i], nl);
            o += nl;
            const float *lrow = low.data() + (size_t)(i - r0) * r1;
            for (int j = 0; j < i; j++) {
                *o++ = '\t';
                o += cms::write_g6(o, is_sim ? (double)lrow[j] : 1.0 - (double)lrow[j]);
            }
            *o++ = '\t';
            *o++ = is_sim ? '1' : '0';
            const float *urow = i + 1 < n ? sim + packed(i, i + 1) : nullptr;
            for (int j = i + 1; j < n; j++) {
                *o++ = '\t';
                o += cms::write_g6(o, is_sim ? (double)urow[j - i - 1] : 1.0 - (double)urow[j - i - 1]);
            }
            *o++ = '\n';
            s.resize((size_t)(o - s.data()));
        }
        if (writer.joinable()) writer.join();
        writer = std::thread([&rb, r1, r0, f, &write_ok]() {
            for (int i = 0; i < r1 - r0; i++)
                if (fwrite(rb[i].data(), 1, rb[i].size(), f) != rb[i].size()) write_ok = false;
        });
        cur ^= 1;
    }
    if (writer.joinable()) writer.join();
    if (fclose(f) != 0 || !write_ok) {
        set_global_error(std::string("write failed : ") + path);
        return CMS_E_IO;
    }
    return CMS_OK;
}

void cms_free(void *p) { free(p); }
const char *cms_version(void) { return "cms-mi355x 0.1"; }

}  // extern "C"

// ------------------------------------------------------------------ sample files (-i / -l)

// Classification files hold one sample each. The reference keeps a file's
// OTU counts in a __gnu_cxx::hash_map<string, int> with
// hash<const char *> over the id (hash.h:12-19) and sums the normalisation
// total in that container's iteration order (comp.h:194-236). The FP32 total
// depends on that order, so the same libstdc++ container is used here with
// the same sequence of container operations (an increment through
// operator[] can grow the table, so increments go through it too).

struct cms_list {
    std::vector<std::string> paths, names;
};

namespace {

struct OtuIdHash {
    size_t operator()(const std::string &x) const { return __gnu_cxx::hash<const char *>()(x.c_str()); }
};
typedef __gnu_cxx::hash_map<std::string, int, OtuIdHash> OtuCounts;

// Leading int of a token as sscanf("%d") reads it (0 when none: the
// reference's count variable keeps its initial 0).
int scan_int(const char *ts, const char *te) {
    std::string t(ts, te);
    char *end = nullptr;
    errno = 0;
    long v = strtol(t.c_str(), &end, 10);
    if (end == t.c_str()) return 0;
    return (int)v;
}

// _OTU_Parser::Load_file_to_hash (otu_parser.h:104-157): the title line
// selects the format. "#S..." (per-read listing, old format): every line is
// "<read id> <otu id> ...", one count per line. Otherwise (new format): every
// line is "<otu id> <count> ...", the last line of an id sets its count.
// Lines without the fields the format needs are skipped (the reference then
// reuses the previous line's buffer).
bool load_class_file(const std::string &path, OtuCounts &counts, std::string &err) {
    std::string buf;
    if (!read_file(path, buf)) {
        err = "Cannot open input file: " + path;
        return false;
    }
    std::vector<std::pair<size_t, size_t>> lines;
    split_lines(buf, lines);
    if (lines.empty()) return true;
    const bool old_format = lines[0].second > 1 && buf[lines[0].first + 1] == 'S';
    for (size_t k = 1; k < lines.size(); k++) {
        const char *p = buf.data() + lines[k].first, *e = p + lines[k].second, *ts, *te;
        if (old_format) {
            if (!next_token(p, e, ts, te) || !next_token(p, e, ts, te)) continue;
            const std::string id(ts, te);
            if (counts.count(id) == 0) counts[id] = 1;
            else counts[id]++;
        } else {
            if (!next_token(p, e, ts, te)) continue;
            const std::string id(ts, te);
            int c = 0;
            if (next_token(p, e, ts, te)) c = scan_int(ts, te);
            counts[id] = c;
        }
    }
    return true;
}

// _Comp_Tree::Load_abd(const char *, float *, bool) (comp.h:194-236):
// count / copy number per OTU, FP32 total in the hash map's order,
// total = (float)(total / 100.0), a = x / total for ids in the tree.
// Returns the number of the file's OTUs found in the tree. A file whose
// counts sum to zero gives an all-zero sample (the reference divides by 0).
bool sample_file_abundance(const cms_db *db, const std::string &path, bool cp_correct,
                           std::vector<std::pair<int32_t, float>> &out, int &mapped, std::string &err) {
    OtuCounts counts;
    out.clear();
    mapped = 0;
    if (!load_class_file(path, counts, err)) return false;
    std::vector<std::pair<const std::string *, float>> x;
    x.reserve(counts.size());
    float total = 0;
    for (OtuCounts::const_iterator it = counts.begin(); it != counts.end(); ++it) {
        float cp_no = 1.0f;
        if (cp_correct) {
            auto c = db->cp.find(it->first);
            if (c != db->cp.end()) cp_no = c->second;
        }
        const float v = (float)it->second / cp_no;
        x.emplace_back(&it->first, v);
        total += v;
    }
    total /= 100.0;
    for (auto &e : x) {
        auto id = db->id_hash.find(*e.first);
        if (id == db->id_hash.end()) continue;
        mapped++;
        const float a = total != 0.0f ? e.second / total : 0.0f;
        out.emplace_back(id->second, a);
    }
    std::sort(out.begin(), out.end());
    return true;
}

}  // namespace

extern "C" {

int cms_sample_file_abundance(const cms_db *db, const char *path, int cp_correct, float *abd, int *mapped) {
    if (!db || !path || !abd) {
        set_global_error("cms_sample_file_abundance: null argument");
        return CMS_E_INVAL;
    }
    std::vector<std::pair<int32_t, float>> v;
    int m = 0;
    std::string err;
    if (!sample_file_abundance(db, path, cp_correct != 0, v, m, err)) {
        set_global_error(err);
        return CMS_E_IO;
    }
    memset(abd, 0, db->ids.size() * sizeof(float));
    for (auto &e : v) abd[e.first] = e.second;
    if (mapped) *mapped = m;
    return CMS_OK;
}

int cms_sample_files_csr(const cms_db *db, int n, const char *const *paths, int cp_correct, int64_t **rowptr_out,
                         int32_t **col_out, float **val_out, int *mapped) {
    if (!db || n < 0 || (n > 0 && !paths) || !rowptr_out || !col_out || !val_out) {
        set_global_error("cms_sample_files_csr: null argument");
        return CMS_E_INVAL;
    }
    std::vector<std::vector<std::pair<int32_t, float>>> rows((size_t)n);
    std::vector<std::string> errs((size_t)n);
    std::vector<int> m((size_t)n, 0);
    std::vector<char> ok((size_t)n, 1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < n; s++) {
        ok[s] = paths[s] && sample_file_abundance(db, paths[s], cp_correct != 0, rows[s], m[s], errs[s]);
        if (!paths[s]) errs[s] = "cms_sample_files_csr: null path";
        // zero abundances are absent entries of the CSR form
        rows[s].erase(std::remove_if(rows[s].begin(), rows[s].end(),
                                     [](const std::pair<int32_t, float> &e) { return e.second == 0.0f; }),
                      rows[s].end());
    }
    for (int s = 0; s < n; s++)
        if (!ok[s]) {
            set_global_error(errs[s]);
            return CMS_E_IO;
        }
    size_t nnz = 0;
    for (auto &r : rows) nnz += r.size();
    int64_t *rowptr = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * std::max<size_t>(nnz, 1));
    float *val = (float *)malloc(sizeof(float) * std::max<size_t>(nnz, 1));
    if (!rowptr || !col || !val) {
        free(rowptr);
        free(col);
        free(val);
        set_global_error("cms_sample_files_csr: out of host memory");
        return CMS_E_NOMEM;
    }
    rowptr[0] = 0;
    for (int s = 0; s < n; s++) {
        int64_t k = rowptr[s];
        for (auto &e : rows[s]) {
            col[k] = e.first;
            val[k] = e.second;
            k++;
        }
        rowptr[s + 1] = k;
        if (mapped) mapped[s] = m[s];
    }
    *rowptr_out = rowptr;
    *col_out = col;
    *val_out = val;
    return CMS_OK;
}

// Load_List(list, files, ids, prefix) (utility.h:344-391): per non-empty
// line, "<id> <file> ..." when it has two or more fields, else the whole
// line is the file and the id is the directory that holds it (the last
// path component before the final '/', or everything before a single '/',
// or the whole line without '/'). The prefix is prepended to every file.
int cms_list_load(const char *path, const char *prefix, cms_list **out) {
    if (!path || !out) {
        set_global_error("cms_list_load: null argument");
        return CMS_E_INVAL;
    }
    *out = nullptr;
    std::string buf;
    if (!read_file(path, buf)) {
        set_global_error(std::string("Cannot open file: ") + path);
        return CMS_E_IO;
    }
    const std::string pre = prefix ? prefix : "";
    std::vector<std::pair<size_t, size_t>> lines;
    split_lines(buf, lines);
    cms_list *l = new cms_list();
    for (auto &ln : lines) {
        if (ln.second == 0) continue;
        const std::string line = buf.substr(ln.first, ln.second);
        const char *p = line.data(), *e = p + line.size(), *ts, *te;
        std::vector<std::string> f;
        while (next_token(p, e, ts, te)) f.emplace_back(ts, te);
        if (f.size() >= 2) {
            l->names.push_back(f[0]);
            l->paths.push_back(pre + f[1]);
            continue;
        }
        l->paths.push_back(pre + line);
        const size_t first = line.find_first_of('/'), last = line.find_last_of('/');
        if (last == std::string::npos) {
            l->names.push_back(line);
        } else if (first == last) {
            l->names.push_back(line.substr(0, last));
        } else {
            const size_t b = line.rfind('/', last - 1) + 1;
            l->names.push_back(line.substr(b, last - b));
        }
    }
    *out = l;
    return CMS_OK;
}

int cms_list_n(const cms_list *l) { return l ? (int)l->paths.size() : CMS_E_INVAL; }

const char *cms_list_path(const cms_list *l, int i) {
    return (l && i >= 0 && i < (int)l->paths.size()) ? l->paths[(size_t)i].c_str() : nullptr;
}

const char *cms_list_name(const cms_list *l, int i) {
    return (l && i >= 0 && i < (int)l->names.size()) ? l->names[(size_t)i].c_str() : nullptr;
}

void cms_list_free(cms_list *l) { delete l; }

}  // extern "C"
