// cms-make-db — build a database directory in the reference's formats from a
// Newick tree: ROOT/databases/db.config (entry 'G'), ROOT/databases/NAME/
// tree/id.txt and tree/order.txt, and copy_number.txt copied from -c.
//
//   cms-make-db -i TREE.tre -o ROOT [-n NAME] [-c copy_number.txt]
//
// Prints leaves, internal nodes and registers used; exit status 1 on errors.
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/cms.h"

int main(int argc, char **argv) {
    const char *in = nullptr, *root = nullptr, *cp = nullptr;
    std::string name = "newick";
    for (int i = 1; i < argc; i++) {
        const bool has = i + 1 < argc;
        if (!strcmp(argv[i], "-i") && has) in = argv[++i];
        else if (!strcmp(argv[i], "-o") && has) root = argv[++i];
        else if (!strcmp(argv[i], "-n") && has) name = argv[++i];
        else if (!strcmp(argv[i], "-c") && has) cp = argv[++i];
        else {
            fprintf(stderr, "usage: cms-make-db -i TREE.tre -o ROOT [-n NAME] [-c copy_number.txt]\n");
            return 1;
        }
    }
    if (!in || !root) {
        fprintf(stderr, "usage: cms-make-db -i TREE.tre -o ROOT [-n NAME] [-c copy_number.txt]\n");
        return 1;
    }
    cms_newick *t = nullptr;
    if (cms_newick_load(in, &t) != CMS_OK) {
        fprintf(stderr, "Error: %s\n", cms_last_error(nullptr));
        return 1;
    }
    if (cms_newick_write_db(t, root, name.c_str(), cp) != CMS_OK) {
        fprintf(stderr, "Error: %s\n", cms_last_error(nullptr));
        cms_newick_free(t);
        return 1;
    }
    int nslots = 0, multi = 0, unary = 0;
    cms_newick_info(t, &nslots, &multi, &unary);
    printf("%d leaves, %d internal nodes, %d registers (%d multifurcations binarised, %d unary nodes merged)\n",
           cms_newick_leaf_n(t), cms_newick_order_n(t), nslots, multi, unary);
    cms_newick_free(t);
    return 0;
}
