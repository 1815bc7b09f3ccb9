// Internal host-side structures shared by the loader, the schedule compiler and
// the HIP backend of libcms. Nothing here is part of the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cms.h"

namespace cms {

// Error text of the last failure without a context (thread-local).
void set_global_error(const std::string &msg);
const char *global_error();

// Compiled walk schedule node, 16 bytes, read with one scalar load per node.
//   code bits  0..7  : slot of child 1 (valid when not a leaf)
//             8..15  : slot of child 2
//            16..23  : destination slot
//            24      : child 1 is a leaf
//            25      : child 2 is a leaf
//            26      : destination is written (value is read later)
//   leafpos           : leaf-stream position of this node's first leaf child
//   w1, w2            : 1 - dist clamped to MIN_DIST when negative (comp.h:341-344)
struct WalkNode {
    uint32_t code;
    int32_t leafpos;
    float w1;
    float w2;
};
static_assert(sizeof(WalkNode) == 16, "WalkNode must be 16 bytes");

constexpr uint32_t kC1Leaf = 1u << 24;
constexpr uint32_t kC2Leaf = 1u << 25;
constexpr uint32_t kDstLive = 1u << 26;

// A chunk of consecutive schedule nodes whose leaf children occupy at most
// kLeafChunk consecutive leaf-stream positions. The walk kernel stages one
// chunk of leaf values in LDS at a time.
struct WalkChunk {
    int32_t node_begin;
    int32_t node_end;
    int32_t leaf_begin;
    int32_t leaf_end;
};

struct CompiledTree {
    int leaf_n = 0;
    int order_n = 0;
    int nslots = 0;          // slots after liveness re-allocation
    bool is_tree = false;    // every leaf and every value consumed at most once
    bool nonneg_beta = true; // all 1 - w >= 0 (required by the embedding form)
    std::vector<WalkNode> nodes;
    std::vector<int32_t> leaf_stream;  // original leaf index per stream position
    std::vector<WalkChunk> chunks;
    int leaf_chunk = 0;

    // Embedding form: one entry per child edge (2 per node).
    //   entry 2i+j is child j of node i; value y = beta * x(child)
    //   child_leafpos[e] >= 0: leaf-stream position, else -1 - (producer node)
    std::vector<int32_t> ent_child;
    std::vector<double> ent_beta;

    // Subtree segments for the parallel embedding: node ranges
    // [seg_first, seg_last] that are complete subtrees, contiguous in the
    // schedule, of at most seg_max nodes. seg_out[k] = 1 when the segment
    // root's value is read by a later node. Nodes outside every segment are
    // "top" nodes, evaluated after all segments; top_src[2t+j] says where
    // child j of top node t (node index top_nodes[t]) comes from:
    // -1 = its leaf or register as in the walk schedule, k >= 0 = the
    // output of segment k.
    std::vector<int32_t> seg_first, seg_last, seg_out;
    std::vector<int32_t> top_nodes, top_src;
};

// The host C library's logf on (0, 4] (the JSD of the leaf merge kernels):
// inputs whose logf is not log((double)x) rounded to float, or whose double
// evaluation lies within a small margin of a rounding midpoint, with the
// library's result. bitmap bit u is set for each such input bit pattern u;
// table holds (u << 32 | result bits), ascending. Built once per process.
constexpr uint32_t kLogfMaxBits = 0x40800000u;  // 4.0f
// The device's log in double (table_log): x = 2^k z, z in [0.699, 1.398) cut
// into 64 pieces at the float bits u - kLogOff; log x = k ln2 + log c_i +
// log1p(z / c_i - 1), the last by a degree-8 polynomial (|z / c_i - 1| <
// 2^-7; c = 1 for the piece holding 1, so the result stays relatively exact
// near x = 1). invc = 1 / c_i and logc = -log(invc) in double.
constexpr uint32_t kLogOff = 0x3f330000u;
constexpr int kLogPieces = 64;
struct LogfTable {
    std::vector<uint32_t> bitmap;  // (kLogfMaxBits >> 5) + 1 words
    std::vector<uint32_t> blocks;  // bit b: some input in [256 b, 256 b + 255] is listed
    std::vector<uint64_t> table;
    uint32_t ln2_bits = 0;         // logf(2.0f)
    double invc[kLogPieces], logc[kLogPieces];
};
// The device's table_log, evaluated on the host with the same IEEE operations.
double table_log(float x, const double *invc, const double *logc);
const LogfTable &host_logf_table();

// Split a tree schedule into subtree segments (fills the seg_* and top_*
// fields). Schedules whose subtrees are not contiguous get no segments: every
// node is then a top node.
void plan_segments(CompiledTree &t, int seg_max);

// Validate and compile the reference schedule arrays (comp.h:161-192 layout).
// Returns CMS_OK or CMS_E_SCHEDULE with a message in `err`.
int compile_tree(int leaf_n, int order_n, const int32_t *o1, const float *d1,
                 const int32_t *o2, const float *d2, const int32_t *od,
                 int leaf_chunk, CompiledTree &out, std::string &err);

// Float formatting exactly like std::ostream's default for a double
// (precision 6, %g conversion). Appends to `s`. Values in [1e-15, 1e5) take
// an integer path (exact six-digit rounding of the binary value), the rest
// snprintf.
void append_g6(std::string &s, double v);
// The same text written at out (room for 32 bytes); returns its length.
int write_g6(char *out, double v);
// 1 when the integer path and snprintf("%.6g") disagree on v (tests).
int g6_fast_differs(double v);

} // namespace cms
