// ref_driver.cpp — drives the reference's own C++ Meta-Storms code
// (_Comp_Tree / _Table_Format from src/comp.h and src/table_format.h, compiled
// from /root/reference/src by oracle/Makefile) for two jobs:
//
//   --hex            print every pair's Calc_sim (comp.h:332-375) as a C99
//                    hex float, for bit-exact golden fixtures;
//   --time P T       time Calc_sim over the first P pairs of the upper-triangle
//                    pair list (comp_sam.cpp:321-330) on T OpenMP threads,
//                    dynamic,1 (comp_sam.cpp:341), with a wall clock; prints
//                    one JSON line;
//   --files LIST     for every path in LIST (one per line), the abundances of
//                    _Comp_Tree::Load_abd(const char *, float *, bool)
//                    (comp.h:194-236): "# <k> <mapped>" then "<leaf> <hex>"
//                    for every nonzero leaf.
//
// TEST INFRASTRUCTURE ONLY (the checker / CPU baseline), never part of the
// product. Needs ParallelMETA set, as the reference does at static-init time
// (db.h:285, utility.h:27-41).
//
// Usage: ref_driver -D <letter> -T <table> [-r T|F] [-M metric] (--hex | --time P T)
//        ref_driver -D <letter> [-r T|F] --files LIST
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "comp.h"

int main(int argc, char **argv) {
    char db = 'G';
    std::string table_path, files_list;
    bool cp = true, hex = false;
    int metric = 0;
    long time_pairs = 0;
    int threads = 1;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "-D" && i + 1 < argc) db = argv[++i][0];
        else if (a == "-T" && i + 1 < argc) table_path = argv[++i];
        else if (a == "-r" && i + 1 < argc) { const char *v = argv[++i]; cp = !(v[0] == 'f' || v[0] == 'F'); }
        else if (a == "--hex") hex = true;
        else if (a == "-M" && i + 1 < argc) metric = atoi(argv[++i]);
        else if (a == "--time" && i + 2 < argc) { time_pairs = atol(argv[++i]); threads = atoi(argv[++i]); }
        else if (a == "--files" && i + 1 < argc) files_list = argv[++i];
        else { fprintf(stderr, "unknown argument %s\n", a.c_str()); return 2; }
    }
    if (!files_list.empty()) {
        _Comp_Tree tree(db);
        const int L = tree.Get_LeafN();
        std::vector<float> abd(L);
        std::ifstream in(files_list.c_str());
        std::string path;
        int k = 0;
        while (std::getline(in, path)) {
            if (path.empty()) continue;
            const int m = tree.Load_abd(path.c_str(), abd.data(), cp);
            printf("# %d %d\n", k++, m);
            for (int l = 0; l < L; l++) {
                uint32_t bits;
                memcpy(&bits, &abd[l], 4);
                if (bits) printf("%d %a\n", l, (double)abd[l]);
            }
        }
        return 0;
    }
    if (table_path.empty()) { fprintf(stderr, "need -T table\n"); return 2; }
    _Comp_Tree tree(db);
    _Table_Format table(table_path.c_str());
    const int n = table.Get_Sample_Size();
    const int L = tree.Get_LeafN();
    std::vector<float *> abd(n);
    for (int i = 0; i < n; i++) {
        abd[i] = new float[L];
        tree.Load_abd(&table, abd[i], i, cp);
    }
    std::vector<int> pm, pn;
    for (int i = 0; i < n - 1; i++)
        for (int j = i + 1; j < n; j++) { pm.push_back(i); pn.push_back(j); }
    if (hex) {
        for (size_t p = 0; p < pm.size(); p++) {
            float s = tree.Calc_sim(abd[pm[p]], abd[pn[p]], metric);
            printf("%d %d %a\n", pm[p], pn[p], (double)s);
        }
    }
    if (time_pairs > 0) {
        long P = time_pairs < (long)pm.size() ? time_pairs : (long)pm.size();
        std::vector<float> out(P);
        omp_set_num_threads(threads);
        auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1)
        for (long p = 0; p < P; p++) out[p] = tree.Calc_sim(abd[pm[p]], abd[pn[p]], metric);
        auto t1 = std::chrono::steady_clock::now();
        double sec = std::chrono::duration<double>(t1 - t0).count();
        double chk = 0;
        for (long p = 0; p < P; p++) chk += out[p];
        printf("{\"pairs\": %ld, \"threads\": %d, \"seconds\": %.6f, \"pairs_per_sec\": %.3f, "
               "\"order_n\": %d, \"checksum\": %.9g}\n",
               P, threads, sec, P / sec, tree.getOrderN(), chk);
    }
    for (int i = 0; i < n; i++) delete[] abd[i];
    return 0;
}
