// Database builder: Newick tree -> tree/id.txt + tree/order.txt (+ db.config).
//
// The reference only reads these files (_Comp_Tree::Load_id / Load_order,
// comp.h:139-192): id.txt holds one leaf id per line (line = leaf index),
// order.txt one internal node per line in evaluation order,
// "order_1 dist_1 order_2 dist_2 order_d", where order_k >= 0 is a leaf index,
// order_k < 0 is register order_k + 150 (REG_SIZE, comp.h:26), dist_k is child
// k's branch length, and order_d is the register the node's value goes to; the
// last line is the root (its register is never read).
//
// Builder:
//   - leaves are numbered in order of appearance in the Newick text;
//   - a node with one child is merged into the child's edge (lengths add);
//     a node with k > 2 children becomes a left-leaning chain of binary nodes
//     joined by zero-length edges, children kept in text order;
//   - nodes are evaluated in Sethi-Ullman order (the child needing more
//     registers first) and registers are allocated lowest-free after the
//     children's registers are released, so the destination may reuse a
//     child's register (the walk reads both children before writing);
//     register s is written as -(s + 1);
//   - lengths are written as they appear in the text (the reference parses
//     them with operator>> into float); merged lengths as %.9g.
// Parsing and scheduling use explicit stacks: a caterpillar tree of 10^5
// leaves is 10^5 levels deep.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "cms_internal.h"

using cms::set_global_error;

struct cms_newick {
    std::vector<std::string> names;  // leaf index -> id
    std::vector<int32_t> o1, o2, od;
    std::vector<float> d1, d2;
    std::vector<std::string> d1s, d2s;  // lengths as text
    int nslots = 0, multifurcations = 0, unary = 0;
};

namespace {

struct PNode {
    std::vector<int32_t> ch;
    std::string label;
    std::string len_text;  // empty: no length given
    double len = 0.0;
};

struct Parser {
    const char *s;
    size_t n, i = 0;
    std::string err;

    bool fail(const std::string &m) {
        if (err.empty()) err = "Newick: " + m + " at offset " + std::to_string(i);
        return false;
    }
    // whitespace and [comments]
    bool skip() {
        for (;;) {
            while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n')) i++;
            if (i < n && s[i] == '[') {
                const size_t start = i;
                while (i < n && s[i] != ']') i++;
                if (i >= n) {
                    i = start;
                    return fail("unterminated comment");
                }
                i++;
                continue;
            }
            return true;
        }
    }
    bool label(std::string &out) {
        out.clear();
        if (!skip()) return false;
        if (i < n && s[i] == '\'') {
            i++;
            for (;;) {
                if (i >= n) return fail("unterminated quoted label");
                if (s[i] == '\'') {
                    if (i + 1 < n && s[i + 1] == '\'') {
                        out += '\'';
                        i += 2;
                        continue;
                    }
                    i++;
                    break;
                }
                out += s[i++];
            }
            return true;
        }
        while (i < n && !strchr("()[]':;, \t\r\n", s[i])) out += s[i++];
        return true;
    }
    bool length(PNode &nd) {
        if (!skip()) return false;
        if (i >= n || s[i] != ':') return true;
        i++;
        if (!skip()) return false;
        const size_t b = i;
        while (i < n && s[i] != '\0' && (isdigit((unsigned char)s[i]) || strchr("+-.eE", s[i]))) i++;
        nd.len_text.assign(s + b, i - b);
        if (nd.len_text.empty()) return fail("missing branch length after ':'");
        char *end = nullptr;
        errno = 0;
        nd.len = strtod(nd.len_text.c_str(), &end);
        if (*end != '\0' || errno == ERANGE || !std::isfinite(nd.len)) return fail("bad branch length '" + nd.len_text + "'");
        return true;
    }

    // Tree into `nodes` (nodes[root] is the root); false with err on bad input.
    bool parse(std::vector<PNode> &nodes, int32_t &root) {
        std::vector<int32_t> open;  // internal nodes whose ')' has not been read
        root = -1;
        if (!skip()) return false;
        if (i >= n) return fail("empty input");
        for (;;) {
            // a subtree starts here
            if (!skip()) return false;
            if (i < n && s[i] == '(') {
                i++;
                const int32_t id = (int32_t)nodes.size();
                nodes.emplace_back();
                if (!open.empty()) nodes[(size_t)open.back()].ch.push_back(id);
                else if (root >= 0) return fail("text after the tree");
                else root = id;
                open.push_back(id);
                continue;
            }
            PNode leaf;
            if (!label(leaf.label)) return false;
            if (leaf.label.empty()) return fail("unnamed leaf");
            if (!length(leaf)) return false;
            const int32_t id = (int32_t)nodes.size();
            nodes.push_back(std::move(leaf));
            if (!open.empty()) nodes[(size_t)open.back()].ch.push_back(id);
            else if (root >= 0) return fail("text after the tree");
            else root = id;
            // close as many subtrees as end here
            for (;;) {
                if (!skip()) return false;
                if (i >= n) return open.empty() ? fail("missing ';'") : fail("unbalanced '('");
                if (s[i] == ',') {
                    if (open.empty()) return fail("',' outside parentheses");
                    i++;
                    break;  // next sibling
                }
                if (s[i] == ')') {
                    if (open.empty()) return fail("unbalanced ')'");
                    i++;
                    PNode &nd = nodes[(size_t)open.back()];
                    std::string lab;
                    if (!label(lab)) return false;  // internal label (ignored)
                    if (!length(nd)) return false;
                    open.pop_back();
                    continue;
                }
                if (s[i] == ';') {
                    if (!open.empty()) return fail("unbalanced '('");
                    i++;
                    return true;
                }
                return fail(std::string("unexpected '") + s[i] + "'");
            }
        }
    }
};

// Binary schedule from the parsed tree.
int build(const std::vector<PNode> &in, int32_t root, cms_newick &t, std::string &err) {
    // binary nodes: leaves keep a leaf index; internal nodes have two children
    struct BNode {
        int32_t a = -1, b = -1;  // children (binary node ids), -1 for a leaf
        int32_t leaf = -1;
        double len = 0.0;
        std::string len_text;
    };
    std::vector<BNode> bn;
    bn.reserve(in.size() * 2);
    // map parsed node -> binary node, children first (iterative post-order)
    std::vector<int32_t> map(in.size(), -1);
    std::vector<std::pair<int32_t, size_t>> st;
    st.push_back({root, 0});
    // leaf numbering must follow the text: number leaves in a pre-order pass
    {
        std::vector<int32_t> pre{root};
        while (!pre.empty()) {
            const int32_t v = pre.back();
            pre.pop_back();
            const PNode &p = in[(size_t)v];
            if (p.ch.empty()) {
                map[(size_t)v] = -2 - (int32_t)t.names.size();  // provisional: leaf index
                t.names.push_back(p.label);
            }
            for (size_t k = p.ch.size(); k-- > 0;) pre.push_back(p.ch[k]);
        }
    }
    while (!st.empty()) {
        const int32_t v = st.back().first;
        size_t &k = st.back().second;
        const PNode &p = in[(size_t)v];
        if (k < p.ch.size()) {
            st.push_back({p.ch[k++], 0});
            continue;
        }
        st.pop_back();
        int32_t id;
        if (p.ch.empty()) {
            BNode b;
            b.leaf = -2 - map[(size_t)v];
            id = (int32_t)bn.size();
            bn.push_back(b);
        } else if (p.ch.size() == 1) {
            id = map[(size_t)p.ch[0]];  // merged into the child's edge below
            t.unary++;
        } else {
            if (p.ch.size() > 2) t.multifurcations++;
            int32_t cur = map[(size_t)p.ch[0]];
            for (size_t j = 1; j < p.ch.size(); j++) {
                BNode b;
                b.a = cur;
                b.b = map[(size_t)p.ch[j]];
                bn.push_back(b);
                cur = (int32_t)bn.size() - 1;
                if (j + 1 < p.ch.size()) {
                    bn[(size_t)cur].len = 0.0;  // edge into the next chain node
                    bn[(size_t)cur].len_text = "0";
                }
            }
            id = cur;
        }
        // this node's own edge
        BNode &me = bn[(size_t)id];
        if (p.ch.size() == 1) {
            if (!p.len_text.empty()) {
                me.len += p.len;
                me.len_text.clear();  // a sum: printed from the value
            }
        } else {
            me.len = p.len;
            me.len_text = p.len_text.empty() ? "0" : p.len_text;
        }
        map[(size_t)v] = id;
    }
    const int32_t broot = map[(size_t)root];
    if (t.names.empty()) {
        err = "Newick: no leaves";
        return CMS_E_INVAL;
    }
    // Sethi-Ullman register need (leaves are read directly: need 0)
    std::vector<int32_t> need(bn.size(), 0);
    for (size_t v = 0; v < bn.size(); v++) {  // children precede parents in bn
        const BNode &b = bn[v];
        if (b.leaf >= 0) continue;
        const int32_t x = need[(size_t)b.a], y = need[(size_t)b.b];
        need[v] = x == y ? x + 1 : std::max(x, y);
    }
    // evaluation with lowest-free registers
    std::vector<char> busy;
    auto alloc = [&]() {
        size_t s = 0;
        while (s < busy.size() && busy[s]) s++;
        if (s == busy.size()) busy.push_back(0);
        busy[s] = 1;
        t.nslots = std::max<int>(t.nslots, (int)s + 1);
        return (int32_t)s;
    };
    std::vector<int32_t> val(bn.size(), 0);  // >= 0 leaf index, < 0 register -(s+1)
    auto fmt_len = [](const BNode &b) {
        if (!b.len_text.empty()) return b.len_text;
        char buf[64];
        snprintf(buf, sizeof buf, "%.9g", (double)(float)b.len);
        return std::string(buf);
    };
    std::vector<std::pair<int32_t, int>> ev;
    ev.push_back({broot, 0});
    while (!ev.empty()) {
        const int32_t v = ev.back().first;
        int &stage = ev.back().second;
        const BNode &b = bn[(size_t)v];
        if (b.leaf >= 0) {
            val[(size_t)v] = b.leaf;
            ev.pop_back();
            continue;
        }
        const bool b_first = need[(size_t)b.b] > need[(size_t)b.a];
        if (stage < 2) {
            const int32_t next = (stage == 0) == b_first ? b.b : b.a;
            stage++;
            ev.push_back({next, 0});
            continue;
        }
        ev.pop_back();
        for (int32_t c : {b.a, b.b})
            if (bn[(size_t)c].leaf < 0) busy[(size_t)(-val[(size_t)c] - 1)] = 0;
        const int32_t d = alloc();
        if (d >= CMS_REG_SIZE) {
            err = "Newick: schedule needs more than " + std::to_string(CMS_REG_SIZE) + " registers";
            return CMS_E_UNSUPPORTED;
        }
        val[(size_t)v] = -(d + 1);
        const BNode &A = bn[(size_t)b.a], &B = bn[(size_t)b.b];
        t.o1.push_back(val[(size_t)b.a]);
        t.o2.push_back(val[(size_t)b.b]);
        t.od.push_back(val[(size_t)v]);
        t.d1.push_back((float)A.len);
        t.d2.push_back((float)B.len);
        t.d1s.push_back(fmt_len(A));
        t.d2s.push_back(fmt_len(B));
    }
    return CMS_OK;
}

bool mkdirs(const std::string &path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); i++) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty()) {
                struct stat sb;
                if (stat(cur.c_str(), &sb) != 0 && mkdir(cur.c_str(), 0755) != 0) return false;
            }
        }
        if (i < path.size()) cur += path[i];
    }
    return true;
}

}  // namespace

extern "C" {

int cms_newick_parse(const char *text, int64_t len, cms_newick **out) {
    if (!text || !out || len < 0) {
        set_global_error("cms_newick_parse: null argument");
        return CMS_E_INVAL;
    }
    *out = nullptr;
    Parser p{text, (size_t)len};
    std::vector<PNode> nodes;
    int32_t root = -1;
    if (!p.parse(nodes, root)) {
        set_global_error(p.err);
        return CMS_E_INVAL;
    }
    cms_newick *t = new cms_newick();
    std::string err;
    const int rc = build(nodes, root, *t, err);
    if (rc) {
        delete t;
        set_global_error(err);
        return rc;
    }
    *out = t;
    return CMS_OK;
}

int cms_newick_load(const char *path, cms_newick **out) {
    if (!path || !out) {
        set_global_error("cms_newick_load: null argument");
        return CMS_E_INVAL;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        set_global_error(std::string("Cannot open file : ") + path);
        return CMS_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    return cms_newick_parse(s.data(), (int64_t)s.size(), out);
}

void cms_newick_free(cms_newick *t) { delete t; }

int cms_newick_leaf_n(const cms_newick *t) { return t ? (int)t->names.size() : 0; }

int cms_newick_order_n(const cms_newick *t) { return t ? (int)t->o1.size() : 0; }

const char *cms_newick_leaf_name(const cms_newick *t, int i) {
    if (!t || i < 0 || i >= (int)t->names.size()) return nullptr;
    return t->names[(size_t)i].c_str();
}

int cms_newick_order(const cms_newick *t, const int32_t **o1, const float **d1, const int32_t **o2, const float **d2,
                     const int32_t **od) {
    if (!t || !o1 || !d1 || !o2 || !d2 || !od) {
        set_global_error("cms_newick_order: null argument");
        return CMS_E_INVAL;
    }
    *o1 = t->o1.data();
    *d1 = t->d1.data();
    *o2 = t->o2.data();
    *d2 = t->d2.data();
    *od = t->od.data();
    return CMS_OK;
}

int cms_newick_info(const cms_newick *t, int *nslots, int *multifurcations, int *unary) {
    if (!t) {
        set_global_error("cms_newick_info: null tree");
        return CMS_E_INVAL;
    }
    if (nslots) *nslots = t->nslots;
    if (multifurcations) *multifurcations = t->multifurcations;
    if (unary) *unary = t->unary;
    return CMS_OK;
}

int cms_newick_write_db(const cms_newick *t, const char *root, const char *name, const char *cp_path) {
    if (!t || !root || !name || !*name) {
        set_global_error("cms_newick_write_db: bad argument");
        return CMS_E_INVAL;
    }
    std::string cp_text;
    if (cp_path) {
        std::ifstream f(cp_path, std::ios::binary);
        if (!f) {
            set_global_error(std::string("Cannot open file : ") + cp_path);
            return CMS_E_IO;
        }
        std::stringstream ss;
        ss << f.rdbuf();
        cp_text = ss.str();
    }
    const std::string base = std::string(root) + "/databases/";
    const std::string dir = base + name + "/";
    if (!mkdirs(dir + "tree")) {
        set_global_error("cms_newick_write_db: cannot create " + dir + "tree");
        return CMS_E_IO;
    }
    FILE *f = fopen((base + "db.config").c_str(), "w");
    if (!f) {
        set_global_error("cms_newick_write_db: cannot write db.config");
        return CMS_E_IO;
    }
    fprintf(f, "#Database id\nG\n#domain\n0\n#name\n%s\n#description\nTree of %zu leaves (Newick)\n"
               "#taxonomy\nN\n#tree\nY\n#copy number\n%s\n#function\nN\n",
            name, t->names.size(), cp_path ? "Y" : "N");
    fclose(f);
    f = fopen((dir + "tree/id.txt").c_str(), "w");
    if (!f) {
        set_global_error("cms_newick_write_db: cannot write id.txt");
        return CMS_E_IO;
    }
    for (const std::string &s : t->names) fprintf(f, "%s\n", s.c_str());
    fclose(f);
    f = fopen((dir + "tree/order.txt").c_str(), "w");
    if (!f) {
        set_global_error("cms_newick_write_db: cannot write order.txt");
        return CMS_E_IO;
    }
    for (size_t i = 0; i < t->o1.size(); i++)
        fprintf(f, "%d\t%s\t%d\t%s\t%d\n", t->o1[i], t->d1s[i].c_str(), t->o2[i], t->d2s[i].c_str(), t->od[i]);
    fclose(f);
    if (cp_path) {
        f = fopen((dir + "copy_number.txt").c_str(), "wb");
        if (!f) {
            set_global_error("cms_newick_write_db: cannot write copy_number.txt");
            return CMS_E_IO;
        }
        fwrite(cp_text.data(), 1, cp_text.size(), f);
        fclose(f);
    }
    return CMS_OK;
}

}  // extern "C"
