// Context state shared by the C ABI implementation (cms_api.hip) and the
// serving loop (cms_serve.hip). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/cms.h"
#include "cms_internal.h"
#include "kernels.h"

using cms::CompiledTree;
using cms::EmbedNode;
using cms::WalkChunk;
using cms::WalkNode;

struct cms_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    cms_opts opts{};
    std::string err;

    bool has_tree = false;
    CompiledTree tree;
    WalkNode *d_nodes = nullptr;
    WalkChunk *d_chunks = nullptr;
    int32_t *d_leaf_lp_ptr = nullptr, *d_leaf_lp = nullptr, *d_leaf_uses = nullptr;
    EmbedNode *d_embnodes = nullptr, *d_embnodes_u = nullptr;
    int64_t nrows_u = 0;
    int32_t *d_seg_first = nullptr, *d_seg_last = nullptr;
    EmbedNode *d_toprecs = nullptr, *d_toprecs_u = nullptr;  // records of the nodes above the segments
    int32_t *d_seg_lp = nullptr;   // each segment's first leaf-stream position
    int32_t *d_topidx = nullptr;   // per leaf-stream position: top leaf index, or -1
    int ntop_leaves = 0;           // leaves read by the nodes above the segments
    int64_t nrows = 0;

    int n = -1;
    int64_t nnz = 0;
    bool has_negative = false;
    int64_t *d_rowptr = nullptr;
    int32_t *d_col = nullptr;
    float *d_val = nullptr;

    uint8_t *d_ab = nullptr;
    int ab_storage = -1;  // storage of the valid walk layout, -1 = none
    float *d_scale = nullptr, *d_inv = nullptr;
    size_t cap_ab = 0, cap_scale = 0, cap_inv = 0;

    bool sm_ready = false;
    int sm_metric = -1;
    int64_t sm_rows = 0;
    bool ab_binned = false;          // walk layout holds samples at stored (binned) positions
    bool sm_binned = false;          // the shared-minimum layout is binned
    std::vector<int32_t> h_perm;     // stored position -> sample (binned layout)
    int32_t *d_perm = nullptr, *d_pos = nullptr;
    unsigned long long *d_work = nullptr;  // persistent-kernel tile counter
    int variant = -1;                // tuned persistent-kernel variant
    int grid_waves = 0;              // persistent grid size in waves
    hipStream_t copy_stream = nullptr;
    hipGraphExec_t band_exec = nullptr;
    unsigned long long *d_colsum = nullptr;
    size_t cap_colsum = 0;
    uint32_t *d_Y = nullptr, *d_bm = nullptr;
    uint32_t *d_lbm = nullptr, *d_pl = nullptr;  // light-entry bitmaps and their per-word prefix
    uint4 *d_crl = nullptr;                     // records of the light entries, block after block in rank order
    int64_t *d_croff = nullptr;                 // each block's first record in d_crl (nblocks + 1)
    size_t cap_lbm = 0, cap_pl = 0, cap_crl = 0, cap_croff = 0;
    // light records of the prepared samples, counted once per sample set,
    // tree, metric and binning (the count sizes d_crl; later preparations of
    // the same inputs reuse it without reading it back)
    int64_t nrec = 0;
    bool nrec_known = false;
    int nrec_metric = -1;
    bool nrec_binned = false;
    double *d_mass = nullptr;
    double *d_xroot = nullptr;
    size_t cap_xroot = 0;
    // leaf lists of the stored samples (embedding input)
    int32_t *d_ll_pos = nullptr;
    float *d_ll_val = nullptr, *d_ll_top = nullptr;
    int64_t *d_ll_start = nullptr, *d_ll_end = nullptr;
    size_t cap_ll_pos = 0, cap_ll_val = 0, cap_ll_top = 0, cap_ll_start = 0, cap_ll_end = 0;
    size_t cap_Y = 0, cap_bm = 0, cap_mass = 0;
    int64_t npad = 0, nwords = 0;
    int scale_bits = 0;
    double max_mass = 0.0;  // largest sample column sum of y (FP64) that chose scale_bits
    int servers = 0;        // serving loops reading this context's layout
    // leaf metrics (-M 2..5)
    bool leaf_ready = false;
    float *d_leafP = nullptr;
    uint32_t *d_leafbm = nullptr;
    double *d_leafstats = nullptr;
    size_t cap_leafP = 0, cap_leafbm = 0, cap_leafstats = 0;
    // leaf metrics in the reference's arithmetic (merge kernels)
    int lm_layout = -1;  // -1 none, 0 raw values (cosine), 1 normalised
    uint2 *d_lment = nullptr;
    float2 *d_lmstat = nullptr;
    int32_t *d_lmperm = nullptr;
    uint32_t *d_logf_bm = nullptr, *d_logf_blocks = nullptr;  // host logf exceptions (JSD), uploaded once
    float logf_ln2 = 0.0f;
    double *d_logf_pieces = nullptr;
    unsigned long long *d_logf_tab = nullptr;
    int logf_ntab = -1;                       // -1 = not uploaded
    std::vector<int32_t> h_leaf_uses;  // per leaf: leaf-stream positions consuming it
    unsigned long long *d_counter = nullptr;  // entries visited by shared-minimum tiles

    float *d_out = nullptr;
    size_t out_cap = 0;
    int32_t *d_prow = nullptr, *d_pcol = nullptr;  // cms_sim_pairs pair lists (grown, reused)
    size_t cap_prow = 0, cap_pcol = 0;
    void *scratch = nullptr;  // this device's static scratch (cms_sim_pairs_noalloc), from hipGetSymbolAddress
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evp0 = nullptr, evp1 = nullptr;
    bool kern_pending = false, prep_pending = false;
    cms_stats stats{};
};

namespace cms {
// Shared-minimum layout of the context's samples for its metric (built if
// needed); the context must hold a tree and samples.
int ctx_prepare(cms_ctx *c);
SmData ctx_sm_data(const cms_ctx *c);
int ctx_fail(cms_ctx *c, int code, const std::string &msg);
// Free the walk layout (rebuilt on demand by later calls).
void ctx_release_walk_layout(cms_ctx *c);
}  // namespace cms
