// cms-gen — write a seeded synthetic tree database and OTU count table in the
// reference's file formats (db.config, tree/id.txt, tree/order.txt,
// copy_number.txt, dense -T table).
//
//   cms-gen --root DIR [--table PATH] [--csr PATH] [--leaves L] [--samples N]
//           [--seed S] [--bl-mean m] [--bl-max x] [--nnz LO HI] [--log-uniform]
//           [--zipf s] [--nb MEAN DISP] [--cp-max k] [--no-cp] [--name db]
//
// --csr writes the same table as a sparse table file (cms_table_save_csr),
// which cuda-comp -T reads like the text table.
//
// Seeds follow one convention: tree = seed, samples = seed + 1000,
// copy numbers = seed + 2000.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cms.h"

int main(int argc, char **argv) {
    std::string root, table, csr, name = "synthetic";
    int L = 5000, N = 64, nnz_lo = 100, nnz_hi = 800, log_uniform = 0, cp_max = 15, use_cp = 1;
    double bl_mean = 0.03, bl_max = 0.99, zipf = 1.1, nb_mean = 50, nb_disp = 0.5;
    unsigned long long seed = 1000;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto v = [&](int k) -> const char * {
            if (i + k >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
            return argv[i + k];
        };
        if (a == "--root") root = v(1), i++;
        else if (a == "--table") table = v(1), i++;
        else if (a == "--csr") csr = v(1), i++;
        else if (a == "--name") name = v(1), i++;
        else if (a == "--leaves") L = atoi(v(1)), i++;
        else if (a == "--samples") N = atoi(v(1)), i++;
        else if (a == "--seed") seed = strtoull(v(1), nullptr, 10), i++;
        else if (a == "--bl-mean") bl_mean = atof(v(1)), i++;
        else if (a == "--bl-max") bl_max = atof(v(1)), i++;
        else if (a == "--nnz") { nnz_lo = atoi(v(1)); nnz_hi = atoi(v(2)); i += 2; }
        else if (a == "--log-uniform") log_uniform = 1;
        else if (a == "--zipf") zipf = atof(v(1)), i++;
        else if (a == "--nb") { nb_mean = atof(v(1)); nb_disp = atof(v(2)); i += 2; }
        else if (a == "--cp-max") cp_max = atoi(v(1)), i++;
        else if (a == "--no-cp") use_cp = 0;
        else { fprintf(stderr, "unknown argument %s\n", a.c_str()); return 2; }
    }
    if (root.empty()) { fprintf(stderr, "--root is required\n"); return 2; }
    const int on = L - 1;
    std::vector<int32_t> o1(on), o2(on), od(on);
    std::vector<float> d1(on), d2(on), cp(L);
    int slots = 0;
    if (cms_synth_tree(seed, L, bl_mean, bl_max, o1.data(), d1.data(), o2.data(), d2.data(), od.data(), &slots)) {
        fprintf(stderr, "%s\n", cms_last_error(nullptr));
        return 1;
    }
    if (cms_synth_copy_numbers(seed + 2000, L, cp_max, cp.data())) return 1;
    if (cms_write_db(root.c_str(), name.c_str(), L, on, o1.data(), d1.data(), o2.data(), d2.data(), od.data(),
                     use_cp ? cp.data() : nullptr)) {
        fprintf(stderr, "%s\n", cms_last_error(nullptr));
        return 1;
    }
    if (!table.empty() || !csr.empty()) {
        cms_synth_samples_opts so;
        memset(&so, 0, sizeof so);
        so.seed = seed + 1000;
        so.n_samples = N;
        so.leaf_n = L;
        so.nnz_lo = nnz_lo;
        so.nnz_hi = nnz_hi;
        so.nnz_log_uniform = log_uniform;
        so.zipf_s = zipf;
        so.nb_mean = nb_mean;
        so.nb_dispersion = nb_disp;
        int64_t *rowptr;
        int32_t *col;
        float *cnt;
        if (cms_synth_samples(&so, &rowptr, &col, &cnt)) { fprintf(stderr, "%s\n", cms_last_error(nullptr)); return 1; }
        if (!table.empty() && cms_write_table(table.c_str(), N, L, rowptr, col, cnt)) {
            fprintf(stderr, "%s\n", cms_last_error(nullptr));
            return 1;
        }
        if (!csr.empty() && cms_write_table_csr(csr.c_str(), N, L, rowptr, col, cnt)) {
            fprintf(stderr, "%s\n", cms_last_error(nullptr));
            return 1;
        }
        cms_free(rowptr);
        cms_free(col);
        cms_free(cnt);
    }
    printf("database %s/databases/%s: %d leaves, %d internal nodes, %d register slots\n", root.c_str(), name.c_str(), L, on,
           slots);
    return 0;
}
