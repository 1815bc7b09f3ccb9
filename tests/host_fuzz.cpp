// Host-side loaders under AddressSanitizer / UBSan: the golden inputs and
// seeded mutations of them (truncations, deletions, stray delimiters, bytes,
// long tokens) through the database, table, sample-file and list loaders, the
// .dist writer, the Newick parser and the schedule validator. Every call may
// fail with an error code; none may touch memory out of bounds, leak or hit
// undefined behaviour. Built and run by tests/test_host_sanitized.py.
//
// usage: host_fuzz <dir with unpacked golden cases> <core.tre> <scratch dir>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include "../../include/cms.h"
namespace cms { const char *global_error(); }
extern "C" const char *cms_last_error(const cms_ctx *) { return cms::global_error(); }

static std::string slurp(const char *p) {
    FILE *f = fopen(p, "rb"); if (!f) return "";
    std::string s; char b[65536]; size_t n;
    while ((n = fread(b, 1, sizeof b, f)) > 0) s.append(b, n);
    fclose(f); return s;
}
static void spit(const std::string &p, const std::string &s) { FILE *f = fopen(p.c_str(), "wb"); fwrite(s.data(), 1, s.size(), f); fclose(f); }

static std::string mutate(std::string s, std::mt19937 &rng) {
    if (s.empty()) s = "x";
    int k = rng() % 6;
    for (int t = 0; t < 1 + (int)(rng() % 4); t++) {
        size_t pos = rng() % (s.size() + 1);
        switch (k) {
        case 0: if (pos < s.size()) s.erase(pos, 1 + rng() % 8); break;
        case 1: {
            static const char kDelims[] = "(),:;'\"\t\n\r 0123456789eE.-+_[]";
            s.insert(pos, 1, kDelims[rng() % (sizeof kDelims - 1)]);
            break;
        }
        case 2: if (pos < s.size()) s[pos] = (char)(rng() % 256); break;
        case 3: s = s.substr(0, pos); break;
        case 4: s.insert(pos, std::string(1 + rng() % 3000, 'a' + rng() % 26)); break;
        default: if (pos < s.size()) s.insert(pos, s.substr(pos, 1 + rng() % 50)); break;
        }
    }
    return s;
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const char *gdir = argv[1];  // unpacked golden cases
    const std::string tmp = argv[3];
    std::mt19937 rng(12345);
    const char *cases[] = {"toy", "quirks", "config0", "core", "files"};
    for (const char *c : cases) {
        std::string root = std::string(gdir) + "/" + c + "/input";
        cms_db *db = nullptr;
        if (cms_db_open(root.c_str(), 'G', &db) != CMS_OK) { printf("%s: db %s\n", c, cms_last_error(nullptr)); continue; }
        int L = cms_db_leaf_n(db);
        std::string tp = root + "/table.txt";
        cms_table *t = nullptr;
        if (cms_table_load(tp.c_str(), &t) == CMS_OK) {
            int n = cms_table_n(t);
            std::vector<float> abd((size_t)n * L);
            std::vector<int> m(n);
            cms_table_abundance(db, t, 1, abd.data(), m.data());
            int64_t *rp; int32_t *col; float *val;
            if (cms_table_abundance_csr(db, t, 1, &rp, &col, &val) == CMS_OK) { cms_free(rp); cms_free(col); cms_free(val); }
            std::vector<const char *> names(n);
            for (int i = 0; i < n; i++) names[i] = cms_table_sample_name(t, i);
            std::vector<float> sim((size_t)n * (n > 0 ? n - 1 : 0) / 2 + 1, 0.5f);
            cms_write_dist((tmp + "/out.dist").c_str(), n, names.data(), sim.data(), 0);
            // mutated tables
            std::string txt = slurp(tp.c_str());
            if (txt.size() < 3000000)
                for (int it = 0; it < 60; it++) {
                    spit(tmp + "/tab.txt", mutate(txt, rng));
                    cms_table *u = nullptr;
                    if (cms_table_load((tmp + "/tab.txt").c_str(), &u) == CMS_OK) {
                        int un = cms_table_n(u);
                        std::vector<float> ua((size_t)un * L + 1);
                        cms_table_abundance(db, u, 1, ua.data(), nullptr);
                        if (cms_table_abundance_csr(db, u, 0, &rp, &col, &val) == CMS_OK) { cms_free(rp); cms_free(col); cms_free(val); }
                        cms_table_free(u);
                    }
                }
            cms_table_free(t);
        }
        // sample files + lists (files case)
        std::string fl = root + "/files_all.txt";
        std::string lst = slurp(fl.c_str());
        if (!lst.empty()) {
            std::vector<std::string> paths;
            size_t b = 0;
            while (b < lst.size()) { size_t e = lst.find('\n', b); if (e == std::string::npos) e = lst.size(); if (e > b) paths.push_back(root + "/" + lst.substr(b, e - b)); b = e + 1; }
            std::vector<float> a(L + 1);
            for (auto &p : paths) {
                int m; cms_sample_file_abundance(db, p.c_str(), 1, a.data(), &m);
                std::string txt = slurp(p.c_str());
                for (int it = 0; it < 40; it++) {
                    spit(tmp + "/f.txt", mutate(txt, rng));
                    cms_sample_file_abundance(db, (tmp + "/f.txt").c_str(), 1, a.data(), &m);
                }
            }
            std::vector<const char *> pp; for (auto &p : paths) pp.push_back(p.c_str());
            int64_t *rp; int32_t *col; float *val;
            if (cms_sample_files_csr(db, (int)pp.size(), pp.data(), 1, &rp, &col, &val, nullptr) == CMS_OK) { cms_free(rp); cms_free(col); cms_free(val); }
            for (const char *ln : {"list_ids.txt", "list_paths.txt", "list_prefix.txt"}) {
                std::string lp = root + "/" + ln;
                std::string txt = slurp(lp.c_str());
                for (int it = 0; it < 200; it++) {
                    spit(tmp + "/l.txt", it ? mutate(txt, rng) : txt);
                    cms_list *l = nullptr;
                    if (cms_list_load((tmp + "/l.txt").c_str(), it % 2 ? "pre/" : nullptr, &l) == CMS_OK) {
                        for (int i = 0; i < cms_list_n(l); i++) { (void)strlen(cms_list_path(l, i)); (void)strlen(cms_list_name(l, i)); }
                        cms_list_free(l);
                    }
                }
            }
            for (const char *odd : {"", "/", "//", "a/", "/a", "a//b", "  \n\t\n", "x y\n/p/q/r/s\n", "a/b\r\n"}) {
                spit(tmp + "/l.txt", odd);
                cms_list *l = nullptr;
                if (cms_list_load((tmp + "/l.txt").c_str(), "", &l) == CMS_OK) {
                    for (int i = 0; i < cms_list_n(l); i++) printf("list [%s] -> [%s] [%s]\n", odd, cms_list_path(l, i), cms_list_name(l, i));
                    cms_list_free(l);
                }
            }
        }
        cms_db_close(db);
    }
    // Newick
    std::string tre = slurp(argv[2]);
    for (int it = 0; it < 3000; it++) {
        std::string s = it ? mutate(tre.substr(0, 20000 + rng() % 20000), rng) : tre;
        cms_newick *nw = nullptr;
        if (cms_newick_parse(s.data(), (int64_t)s.size(), &nw) == CMS_OK) {
            int ns, mf, un; cms_newick_info(nw, &ns, &mf, &un);
            const int32_t *o1, *o2, *od; const float *d1, *d2;
            cms_newick_order(nw, &o1, &d1, &o2, &d2, &od);
            cms_newick_free(nw);
        }
    }
    for (const char *odd : {"", ";", "(", ")", "();", "(,);", "((a,b),c", "a;", "(a:1e400,b:-1);", "('q a':1,\"b\":2);", "(a,b)c:1:2;", "(((((a)))));", "(a,(b,(c,(d,(e,f)))));"}) {
        cms_newick *nw = nullptr;
        int rc = cms_newick_parse(odd, (int64_t)strlen(odd), &nw);
        printf("newick [%s] rc=%d %s\n", odd, rc, rc ? cms_last_error(nullptr) : "");
        if (rc == CMS_OK) cms_newick_free(nw);
    }
    // schedules
    for (int it = 0; it < 2000; it++) {
        int on = 1 + rng() % 20, L = 1 + rng() % 30;
        std::vector<int32_t> o1(on), o2(on), od(on); std::vector<float> d1(on), d2(on);
        for (int i = 0; i < on; i++) { o1[i] = (int)(rng() % 200) - 160; o2[i] = (int)(rng() % 200) - 160; od[i] = -(int)(rng() % 160); d1[i] = d2[i] = 0.1f; }
        int a_, b_, c_; cms_validate_tree(L, on, o1.data(), d1.data(), o2.data(), d2.data(), od.data(), &a_, &b_, &c_);
    }
    printf("host_fuzz done\n");
    return 0;
}
