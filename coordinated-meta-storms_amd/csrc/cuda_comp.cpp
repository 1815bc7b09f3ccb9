// cuda-comp — command-line driver with the reference's flags and .dist output
// (Parse_Para / Output_Matrix / main of cms_cuda.cu:1222-1342, 2237-2261),
// computing every -M metric (0 Meta-Storms, 1 unweighted, 2..5 cosine,
// Euclidean, JSD, Bray-Curtis) on the GPU through libcms.
//
//   cuda-comp -T table -D G -o out.dist [-r T|F] [-d T|F] [-M 0..5] [-t n]
//             [--kernel walk|sharedmin] [--storage f32|bf16|f16] [--device k]
//             [--format dist|binary] [--stream-bands P]
//   cuda-comp -l list [-p prefix] -D G -o out.dist [...]   one sample per file
//   cuda-comp -i file1 file2 -D G [-r T|F] [-d T|F] [-M 0..5]   one pair, printed
//
// -l and -i read classification files (Multi_Comp / Single_Comp,
// cms_cuda.cu:1344-1403); the last of -T / -l / -i on the command line
// selects the mode, as in the reference.
//
// --format binary writes the packed upper triangle (pair (i, j), i < j, at
// i*n + j - (i+1)(i+2)/2; the values .dist would print: 1 - sim, or sim with
// -d F) as little-endian float32 to the -o file and the sample names, one per
// line, to <file>.names. With the shared-minimum kernel a triangle of more
// than 2^28 pairs is streamed to the file in bands as the GPU computes them
// (no n x n buffer in host memory).
//
// The database is read from $ParallelMETA/databases (db.config, tree/id.txt,
// tree/order.txt, copy_number.txt). Errors exit with status 1 (the reference
// prints and exits 0 or continues with empty data).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cms.h"

namespace {

[[noreturn]] void printhelp() {
    printf(" Welcome to Coordinated Meta Storm (MI355X)\n");
    printf("\tCompute the Meta-Storms distance among samples, default use rRNA copy number correction\n");
    printf("Usage: \ncuda-comp [Option] Value\nOptions: \n");
    printf("\t-D (upper) ref database, default is G\n");
    printf("\t[Input options, required]\n");
    printf("\t  -i Two samples path for single sample comparison\n");
    printf("\tor\n");
    printf("\t  -l Input files list for multi-sample comparison\n");
    printf("\t  -p List files path prefix [Optional for -l]\n");
    printf("\tor\n");
    printf("\t  -T (upper) Input OTU count table (*.OTU.Count) for multi-sample comparison\n");
    printf("\t[Output options]\n");
    printf("\t  -o Output file\n");
    printf("\t  -d Output format, distance (T) or similarity (F), default is T\n");
    printf("\t[Other options]\n");
    printf("\t  -M (upper) Distance Metric, 0: Meta-Storms; 1: Meta-Storms-unweighted; 2: Cosine; 3: Euclidean;\n");
    printf("\t     4: Jensen-Shannon; 5: Bray-Curtis; default is 0\n");
    printf("\t  -r rRNA copy number correction, T(rue) or F(alse), default is T\n");
    printf("\t  -t Number of host threads for loading, default is auto\n");
    printf("\t  --kernel walk|sharedmin  walk (default): the reference's FP32 arithmetic per pair, bit-identical\n");
    printf("\t                           to comp; sharedmin: the fast kernels\n");
    printf("\t  --storage f32|bf16|f16   walk abundance storage, default f32\n");
    printf("\t  --device k               GPU index, default 0\n");
    printf("\t  --format dist|binary     text .dist (default) or packed float32 triangle + .names\n");
    printf("\t  --stream-bands P         binary output with sharedmin: stream bands of P pairs (default: only\n");
    printf("\t                           triangles above 2^28 pairs, in bands of 2^26)\n");
    printf("\t  -h Help\n");
    exit(0);
}

[[noreturn]] void die(const std::string &msg) {
    fprintf(stderr, "Error: %s\n", msg.c_str());
    exit(1);
}

struct BinaryOut {
    FILE *f = nullptr;
    bool is_sim = false;
    int64_t written = 0;
    bool ok = true;
    std::vector<float> buf;
};

// Band callback of cms_all_pairs_stream: values as .dist prints them.
void write_band(void *user, int64_t first, int64_t count, const float *sims) {
    BinaryOut *o = static_cast<BinaryOut *>(user);
    if (!o->ok) return;
    if (first != o->written) {  // bands arrive in packed order when samples are not binned
        o->ok = false;
        return;
    }
    o->buf.resize((size_t)count);
    for (int64_t k = 0; k < count; k++) o->buf[(size_t)k] = o->is_sim ? sims[k] : (float)(1.0 - (double)sims[k]);
    if (fwrite(o->buf.data(), sizeof(float), (size_t)count, o->f) != (size_t)count) o->ok = false;
    o->written += count;
}

}  // namespace

int main(int argc, char **argv) {
    char ref_db = 'G';
    std::string table_path, out_path, list_path, list_prefix, query1, query2;
    int mode = 0;  // 0: -i (the reference's default), 1: -l, 2: -T
    bool cp_correct = true, is_sim = false;
    int metric = 0, device = 0;
    bool binary = false;
    int64_t band_pairs = 0;  // --stream-bands: stream binary output in bands of this many pairs
    cms_opts opts;
    cms_opts_default(&opts);
    if (argc == 1) printhelp();
    int i = 1;
    while (i < argc) {
        std::string a = argv[i];
        if (a.size() < 2 || a[0] != '-') {
            fprintf(stderr, "Argument # %d Error : Arguments must start with -\n", i);
            return 1;
        }
        auto need = [&](int k) -> const char * {
            if (i + k >= argc) die("missing value for " + a);
            return argv[i + k];
        };
        if (a == "--kernel") {
            std::string v = need(1);
            if (v == "walk") opts.kernel = CMS_KERNEL_WALK;
            else if (v == "sharedmin") opts.kernel = CMS_KERNEL_SHAREDMIN;
            else die("unknown kernel " + v);
        } else if (a == "--storage") {
            std::string v = need(1);
            if (v == "f32") opts.storage = CMS_STORE_F32;
            else if (v == "bf16") opts.storage = CMS_STORE_BF16;
            else if (v == "f16") opts.storage = CMS_STORE_F16;
            else die("unknown storage " + v);
        } else if (a == "--device") {
            device = atoi(need(1));
        } else if (a == "--stream-bands") {
            band_pairs = atoll(need(1));
            if (band_pairs <= 0) die("--stream-bands needs a positive pair count");
        } else if (a == "--format") {
            std::string v = need(1);
            if (v == "dist") binary = false;
            else if (v == "binary") binary = true;
            else die("unknown format " + v);
        } else {
            switch (a[1]) {
            case 'D': ref_db = need(1)[0]; break;
            case 'T': table_path = need(1); mode = 2; break;
            case 'l': list_path = need(1); mode = 1; break;
            case 'p': list_prefix = need(1); break;
            case 'i':
                query1 = need(1);
                query2 = need(2);
                mode = 0;
                i++;
                break;
            case 'o': out_path = need(1); break;
            case 'M': metric = atoi(need(1)); break;
            case 'r': { const char *v = need(1); if (v[0] == 'f' || v[0] == 'F') cp_correct = false; } break;
            case 'd': { const char *v = need(1); if (v[0] == 'f' || v[0] == 'F') is_sim = true; } break;
            case 't': (void)need(1); break;
            case 'P': case 'c': (void)need(1); break;
            case 'h': printhelp();
            default:
                fprintf(stderr, "Error: Unrec argument %s\n", argv[i]);
                printhelp();
            }
        }
        i += 2;
    }
    if (mode == 0 && (query1.empty() || query2.empty())) die("an input is required: -i file1 file2, -l list or -T table");
    if (metric < 0 || metric > 5) metric = 4;  // Calc_sim's default case (comp.h:322-325)

    auto t0 = std::chrono::steady_clock::now();
    cms_db *db = nullptr;
    if (cms_db_open(nullptr, ref_db, &db) != CMS_OK) die(cms_last_error(nullptr));
    // without a tree, -M 0 and -M 1 fall back to JSD as in Calc_sim (comp.h:299-302)
    const bool has_tree = cms_db_has_tree(db) && cms_db_order_n(db) > 0;
    if (!has_tree && metric < 2) metric = 4;
    opts.metric = metric;  // walk: the reference's arithmetic (walk or leaf merge); sharedmin: the fast kernels
    const int L = cms_db_leaf_n(db), on = cms_db_order_n(db);
    cms_table *table = nullptr;
    cms_list *list = nullptr;
    int n = 0;
    int64_t *rowptr = nullptr;
    int32_t *leaf = nullptr;
    float *val = nullptr;
    std::vector<int> mapped;
    if (mode == 2) {
        if (cms_table_load(table_path.c_str(), &table) != CMS_OK) die(cms_last_error(nullptr));
        n = cms_table_n(table);
        if (cms_table_abundance_csr(db, table, cp_correct ? 1 : 0, &rowptr, &leaf, &val) != CMS_OK)
            die(cms_last_error(nullptr));
    } else {
        std::vector<const char *> files;
        if (mode == 1) {
            if (cms_list_load(list_path.c_str(), list_prefix.c_str(), &list) != CMS_OK) die(cms_last_error(nullptr));
            n = cms_list_n(list);
            for (int k = 0; k < n; k++) files.push_back(cms_list_path(list, k));
        } else {
            n = 2;
            files = {query1.c_str(), query2.c_str()};
        }
        mapped.assign((size_t)std::max(n, 1), 0);
        if (cms_sample_files_csr(db, n, files.data(), cp_correct ? 1 : 0, &rowptr, &leaf, &val, mapped.data()) !=
            CMS_OK)
            die(cms_last_error(nullptr));
    }
    if (mode == 0) {
        printf("%d OTUs in file 1\n%d OTUs in file 2\n", mapped[0], mapped[1]);
    } else {
        printf("%d files loaded\n", n);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (mode != 0 && out_path.empty()) die("Cannot open output file : ");

    const int32_t *o1 = nullptr, *o2 = nullptr, *od = nullptr;
    const float *d1 = nullptr, *d2 = nullptr;
    const int32_t none_i = 0;
    const float none_f = 0.0f;
    if (on > 0) cms_db_order(db, &o1, &d1, &o2, &d2, &od);
    if (!o1) { o1 = o2 = od = &none_i; d1 = d2 = &none_f; }
    // binary output of a triangle above 1 GiB floats is streamed band by band
    // (unbinned, so bands arrive in the packed order of the table's samples);
    // a smaller one is computed whole with samples binned by sparsity
    const int64_t npairs = (int64_t)n * (n > 0 ? n - 1 : 0) / 2;
    const bool stream = binary && metric < 2 && opts.kernel == CMS_KERNEL_SHAREDMIN &&
                        (band_pairs > 0 || npairs > ((int64_t)1 << 28));
    if (opts.kernel == CMS_KERNEL_SHAREDMIN) opts.bin = stream ? 0 : 1;
    cms_ctx *ctx = nullptr;
    if (cms_create(device, &opts, &ctx) != CMS_OK) die(cms_last_error(nullptr));
    if (cms_set_tree(ctx, L, has_tree ? on : 0, o1, d1, o2, d2, od) != CMS_OK) die(cms_last_error(ctx));
    if (cms_set_samples_csr(ctx, n, rowptr, leaf, val) != CMS_OK) die(cms_last_error(ctx));
    cms_free(rowptr);
    cms_free(leaf);
    cms_free(val);
    std::vector<const char *> names(n);
    for (int k = 0; k < n; k++) names[k] = mode == 2 ? cms_table_sample_name(table, k) : mode == 1 ? cms_list_name(list, k) : "";
    auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
    if (mode == 0) {
        // Single_Comp (cms_cuda.cu:1344-1360): the value on stdout, printed
        // as std::ostream prints a float (sim) or a double (1.0 - sim)
        float sim = 0.0f;
        if (cms_all_pairs(ctx, &sim) != CMS_OK) die(cms_last_error(ctx));
        printf("%.6g\n", is_sim ? (double)sim : 1.0 - (double)sim);
    } else if (binary) {
        const std::string names_path = out_path + ".names";
        FILE *nf = fopen(names_path.c_str(), "w");
        if (!nf) die("Cannot open output file : " + names_path);
        for (int k = 0; k < n; k++) fprintf(nf, "%s\n", names[k]);
        fclose(nf);
        BinaryOut bo;
        bo.f = fopen(out_path.c_str(), "wb");
        bo.is_sim = is_sim;
        if (!bo.f) die("Cannot open output file : " + out_path);
        if (stream) {
            if (cms_all_pairs_stream(ctx, band_pairs > 0 ? band_pairs : (int64_t)1 << 26, write_band, &bo) != CMS_OK)
                die(cms_last_error(ctx));
        } else {
            std::vector<float> sim((size_t)n * (n > 0 ? n - 1 : 0) / 2);
            if (cms_all_pairs(ctx, sim.data()) != CMS_OK) die(cms_last_error(ctx));
            write_band(&bo, 0, (int64_t)sim.size(), sim.data());
        }
        if (fclose(bo.f) != 0) bo.ok = false;
        if (!bo.ok || bo.written != (int64_t)n * (n > 0 ? n - 1 : 0) / 2) die("writing " + out_path + " failed");
        cms_stats st;
        cms_get_stats(ctx, &st);
        auto t2 = std::chrono::steady_clock::now();
        printf("*** load %.3f s, gpu + write %.3f s (pair kernel %.3f ms)\n", sec(t0, t1), sec(t1, t2), st.kernel_ms);
    } else {
        std::vector<float> sim((size_t)n * (n > 0 ? n - 1 : 0) / 2);
        if (cms_all_pairs(ctx, sim.data()) != CMS_OK) die(cms_last_error(ctx));
        cms_stats st;
        cms_get_stats(ctx, &st);
        auto t2 = std::chrono::steady_clock::now();
        if (cms_write_dist(out_path.c_str(), n, names.data(), sim.data(), is_sim ? 1 : 0) != CMS_OK)
            die(cms_last_error(nullptr));
        auto t3 = std::chrono::steady_clock::now();
        printf("*** load %.3f s, gpu %.3f s (pair kernel %.3f ms), write %.3f s\n", sec(t0, t1), sec(t1, t2),
               st.kernel_ms, sec(t2, t3));
    }
    cms_destroy(ctx);
    if (list) cms_list_free(list);
    if (table) cms_table_free(table);
    cms_db_close(db);
    return 0;
}
