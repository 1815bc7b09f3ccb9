// C ABI of libcms: contexts, device layouts and dispatch to the walk and
// shared-minimum kernels. Replaces the per-tile, per-GPU host loop of
// Multi_GPU_split_MetaStorm (cms_cuda.cu:1470-2235): inputs are uploaded once,
// laid out for the kernels on the device, and every pair is computed by one
// launch per call on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "cms_ctx.h"

using namespace cms;

// Device scratch of cms_sim_pairs_noalloc: the staged row list, column list
// and results of up to CMS_NOALLOC_MAX_PAIRS pairs (12 MiB). Every device the
// library runs on holds its own copy until the process ends, shared by all
// contexts on that device (calls that overlap on one device race). In host
// code the array's name is only a stand-in: the device address comes from
// hipGetSymbolAddress, taken once per device by cms_create with that device
// current.
static __device__ __align__(256) uint8_t scratch[(size_t)CMS_NOALLOC_MAX_PAIRS * 12];

namespace {

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device, null until taken

int fail(cms_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    set_global_error(msg);
    return code;
}

#define HIPCALL(ctx, expr)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, e_ == hipErrorOutOfMemory ? CMS_E_NOMEM : CMS_E_HIP,                  \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <typename T>
hipError_t dalloc(T *&p, size_t count) {
    dfree(p);
    if (count == 0) count = 1;
    return hipMalloc((void **)&p, count * sizeof(T));
}

// Reuse an allocation when it is large enough (layouts are rebuilt in place
// by cms_prepare without returning memory to the driver).
template <typename T>
hipError_t dreserve(T *&p, size_t &cap_bytes, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (p && cap_bytes >= bytes) return hipSuccess;
    dfree(p);
    cap_bytes = 0;
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e == hipSuccess) cap_bytes = bytes;
    return e;
}

void drop_layouts(cms_ctx *c) {
    dfree(c->d_ab);
    c->cap_ab = 0;
    c->ab_storage = -1;
    dfree(c->d_scale);
    dfree(c->d_inv);
    c->cap_scale = c->cap_inv = 0;
    dfree(c->d_Y);
    dfree(c->d_bm);
    dfree(c->d_mass);
    dfree(c->d_xroot);
    dfree(c->d_colsum);
    c->cap_xroot = c->cap_colsum = 0;
    dfree(c->d_ll_pos);
    dfree(c->d_ll_val);
    dfree(c->d_ll_top);
    dfree(c->d_ll_start);
    dfree(c->d_ll_end);
    c->cap_ll_pos = c->cap_ll_val = c->cap_ll_top = c->cap_ll_start = c->cap_ll_end = 0;
    c->cap_Y = c->cap_bm = c->cap_mass = 0;
    c->sm_ready = false;
    c->nrec_known = false;
    dfree(c->d_leafP);
    dfree(c->d_leafbm);
    dfree(c->d_leafstats);
    c->cap_leafP = c->cap_leafbm = c->cap_leafstats = 0;
    c->leaf_ready = false;
    dfree(c->d_lment);
    dfree(c->d_lmstat);
    dfree(c->d_lmperm);
    c->lm_layout = -1;
    c->variant = -1;  // chosen again for the new samples or tree
}

// -M 2..5 run on the leaf-metric kernels: with CMS_KERNEL_WALK the merge
// kernels (the reference's FP32 arithmetic), with CMS_KERNEL_SHAREDMIN the
// leaf tiles (FP64 sums).
bool is_leaf_metric(const cms_ctx *c) { return c->opts.metric >= CMS_METRIC_COS; }
bool use_leaf_merge(const cms_ctx *c) { return is_leaf_metric(c) && c->opts.kernel == CMS_KERNEL_WALK; }

// The shared-minimum kernels serve CMS_KERNEL_SHAREDMIN and every
// unweighted request (the walk computes the weighted metric only).
bool use_sm(const cms_ctx *c) {
    return !is_leaf_metric(c) && c->opts.kernel == CMS_KERNEL_SHAREDMIN;
}

int check_ready(cms_ctx *c) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    if (!c->has_tree) return fail(c, CMS_E_STATE, "no tree loaded (cms_set_tree)");
    if (c->n < 0) return fail(c, CMS_E_STATE, "no samples loaded (cms_set_samples)");
    HIPCALL(c, hipSetDevice(c->device));
    return CMS_OK;
}

// Block-leaf-major walk layout of stored positions [p0, p0 + cnt) in d_ab
// (zeros elsewhere): the samples stored there are h_perm's when `binned`.
int build_walk_slice(cms_ctx *c, int storage, bool binned, int p0, int cnt) {
    const int64_t nlp = (int64_t)c->tree.leaf_stream.size();
    const int64_t nblk = ((int64_t)cnt + kWalkBlock - 1) / kWalkBlock;
    const size_t elem = storage == CMS_STORE_F32 ? 4 : 2;
    const size_t bytes = (size_t)std::max<int64_t>(nblk, 1) * (size_t)std::max<int64_t>(nlp, 1) * kWalkBlock * elem;
    HIPCALL(c, dreserve(c->d_ab, c->cap_ab, bytes));
    HIPCALL(c, hipMemsetAsync(c->d_ab, 0, bytes, c->stream));
    HIPCALL(c, launch_walk_scatter(c->d_rowptr, c->d_col, c->d_val, c->d_leaf_lp_ptr, c->d_leaf_lp, nlp, storage,
                                   c->d_scale, binned ? c->d_perm : nullptr, p0, cnt, c->d_ab, c->stream));
    return CMS_OK;
}

// Walk layout of all samples, at their stored positions when `binned`
// (shared-minimum only).
int ensure_walk_layout(cms_ctx *c, int storage, bool binned = false) {
    if (c->ab_storage == storage && c->ab_binned == binned) return CMS_OK;
    c->ab_storage = -1;
    if (storage != CMS_STORE_F32) {
        HIPCALL(c, dreserve(c->d_scale, c->cap_scale, (size_t)std::max(c->n, 1)));
        HIPCALL(c, dreserve(c->d_inv, c->cap_inv, (size_t)std::max(c->n, 1)));
        HIPCALL(c, launch_sample_scale(c->d_rowptr, c->d_val, c->n, storage, c->d_scale, c->d_inv, c->stream));
    }
    int rc = build_walk_slice(c, storage, binned, 0, c->n);
    if (rc) return rc;
    c->ab_storage = storage;
    c->ab_binned = binned;
    return CMS_OK;
}

WalkData walk_data(const cms_ctx *c) {
    WalkData d{};
    d.nodes = c->d_nodes;
    d.chunks = c->d_chunks;
    d.nchunks = (int)c->tree.chunks.size();
    d.order_n = c->tree.order_n;
    d.nslots = std::max(c->tree.nslots, 1);
    d.nlp = (int64_t)c->tree.leaf_stream.size();
    d.ab = c->d_ab;
    d.inv_scale = c->d_inv;
    d.n = c->n;
    d.storage = c->ab_storage;
    d.unweighted = c->opts.metric == CMS_METRIC_UNWEIGHTED ? 1 : 0;
    return d;
}

// Abundance storage of the walk: the unweighted walk tests presence and keeps
// F32 (a 16-bit value can round a small abundance to zero).
int walk_storage(const cms_ctx *c) {
    return c->opts.metric == CMS_METRIC_UNWEIGHTED ? CMS_STORE_F32 : c->opts.storage;
}

int sharedmin_supported(cms_ctx *c) {
    if (!c->tree.is_tree)
        return fail(c, CMS_E_UNSUPPORTED, "shared-minimum kernel needs a tree schedule (a value or leaf is read twice)");
    if (!c->tree.nonneg_beta)
        return fail(c, CMS_E_UNSUPPORTED, "shared-minimum kernel needs non-negative branch lengths");
    if (c->has_negative) return fail(c, CMS_E_UNSUPPORTED, "shared-minimum kernel needs non-negative abundances");
    if (c->tree.nslots > kEmbedMaxSlots)
        return fail(c, CMS_E_UNSUPPORTED, "schedule keeps too many registers live for the embedding kernel");
    return CMS_OK;
}

// Largest fixed-point scale K with max column sum * 2^K + rounding below
// kSmFixedMax (2^31).
int choose_scale_bits(double max_colsum, int64_t nrows) {
    if (!(max_colsum > 0)) return 24;
    const double room = kSmFixedMax - (double)nrows * 0.5 - 64.0;
    int K = (int)std::floor(std::log2(room / max_colsum));
    while (K > -60 && std::ldexp(max_colsum, K) + (double)nrows * 0.5 >= kSmFixedMax) K--;
    return std::min(K, 60);
}

// Fixed-point embedding Y, the per-block entry bitmaps and per-sample column
// sums, for the context's metric.
int ensure_sharedmin(cms_ctx *c) {
    const int metric = c->opts.metric;
    if (c->sm_ready && c->sm_metric == metric && c->sm_binned == (c->opts.bin != 0)) return CMS_OK;
    c->sm_ready = false;
    int rc = sharedmin_supported(c);
    if (rc) return rc;
    const bool binned = c->opts.bin != 0;
    if (binned) {
        // stored order: densest samples first, so 32-sample blocks have
        // similar sparsity and sparse blocks have small root-path unions
        std::vector<int64_t> rp((size_t)c->n + 1);
        if (c->n > 0) HIPCALL(c, hipMemcpy(rp.data(), c->d_rowptr, sizeof(int64_t) * (c->n + 1), hipMemcpyDeviceToHost));
        c->h_perm.resize(c->n);
        for (int s = 0; s < c->n; s++) c->h_perm[s] = s;
        std::stable_sort(c->h_perm.begin(), c->h_perm.end(),
                         [&](int32_t a, int32_t b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
        std::vector<int32_t> pos(c->n);
        for (int p = 0; p < c->n; p++) pos[c->h_perm[p]] = p;
        HIPCALL(c, dalloc(c->d_perm, (size_t)std::max(c->n, 1)));
        HIPCALL(c, dalloc(c->d_pos, (size_t)std::max(c->n, 1)));
        if (c->n > 0) {
            HIPCALL(c, hipMemcpy(c->d_perm, c->h_perm.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice));
            HIPCALL(c, hipMemcpy(c->d_pos, pos.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice));
        }
    }
    c->sm_binned = binned;
    const bool unw = metric == CMS_METRIC_UNWEIGHTED;
    const int64_t nrows = unw ? c->nrows_u : c->nrows;
    c->npad = ((int64_t)c->n + 63) / 64 * 64;
    c->nwords = (nrows + 31) / 32;
    // The embedding reads the samples' leaf lists (their nonzero leaves in
    // leaf-stream order, built here) where the leaf stream fits the list
    // builder, else the F32 walk layout. The walk layout is built and
    // embedded in slices of stored positions (multiples of 64 samples) above
    // kEmbedSliceBytes, so the transient layout stays bounded beside Y;
    // cms_opts.embed_slice forces slices on either input.
    const int64_t nlp = (int64_t)c->tree.leaf_stream.size();
    const int nseg = (int)c->tree.seg_first.size();
    const bool lists = c->opts.embed_input == 0 && nlp >= 1 && nlp <= kLeafListMaxPositions;
    const size_t per_sample = (size_t)std::max<int64_t>(nlp, 1) * sizeof(float);
    int64_t slice = c->npad;
    if (c->opts.embed_slice > 0)
        slice = std::max<int64_t>(64, (int64_t)c->opts.embed_slice / 64 * 64);
    else if (!lists && (size_t)c->npad * per_sample > kEmbedSliceBytes)
        slice = std::max<int64_t>(64, (int64_t)(kEmbedSliceBytes / per_sample) / 64 * 64);
    const int64_t nslices = c->n > 0 ? (c->n + slice - 1) / slice : 1;
    if (lists) {
        const int64_t ntl = c->ntop_leaves;
        HIPCALL(c, dreserve(c->d_ll_pos, c->cap_ll_pos, (size_t)std::max<int64_t>(c->nnz, 1)));
        HIPCALL(c, dreserve(c->d_ll_val, c->cap_ll_val, (size_t)std::max<int64_t>(c->nnz, 1)));
        HIPCALL(c, dreserve(c->d_ll_start, c->cap_ll_start, (size_t)std::max<int64_t>((int64_t)c->n * nseg, 1)));
        HIPCALL(c, dreserve(c->d_ll_end, c->cap_ll_end, (size_t)std::max(c->n, 1)));
        HIPCALL(c, dreserve(c->d_ll_top, c->cap_ll_top, (size_t)std::max<int64_t>(ntl * c->npad, 1)));
        if (ntl > 0) HIPCALL(c, hipMemsetAsync(c->d_ll_top, 0, sizeof(float) * ntl * c->npad, c->stream));
        HIPCALL(c, launch_leaf_lists(c->d_rowptr, c->d_col, c->d_val, binned ? c->d_perm : nullptr, c->n,
                                     c->d_leaf_lp_ptr, c->d_leaf_lp, nlp, c->d_seg_lp, nseg, c->d_topidx,
                                     c->d_ll_pos, c->d_ll_val, c->d_ll_start, c->d_ll_end, c->d_ll_top, c->npad,
                                     c->stream));
    } else if (nslices == 1) {
        rc = ensure_walk_layout(c, CMS_STORE_F32, binned);
        if (rc) return rc;
    } else {
        c->ab_storage = -1;  // the layout will hold the last slice only
    }
    const int64_t spad = std::min<int64_t>(slice, c->npad);
    HIPCALL(c, dreserve(c->d_xroot, c->cap_xroot, (size_t)std::max<size_t>(c->tree.seg_first.size(), 1) * spad));
    HIPCALL(c, dreserve(c->d_colsum, c->cap_colsum, (size_t)std::max<int64_t>(c->npad, 1)));
    HIPCALL(c, dreserve(c->d_mass, c->cap_mass, (size_t)std::max<int64_t>(c->npad, 1)));
    EmbedData ed{};
    ed.nodes = unw ? c->d_embnodes_u : c->d_embnodes;
    ed.order_n = c->tree.order_n;
    ed.nslots = std::max(c->tree.nslots, 1);
    ed.seg_first = c->d_seg_first;
    ed.seg_last = c->d_seg_last;
    ed.nseg = (int)c->tree.seg_first.size();
    ed.top = unw ? c->d_toprecs_u : c->d_toprecs;
    ed.ntop = (int)c->tree.top_nodes.size();
    ed.ab = (const float *)c->d_ab;
    ed.sparse = lists ? 1 : 0;
    ed.nlp = nlp;
    ed.nwords = c->nwords;
    ed.ystride = c->npad;
    ed.xroot = c->d_xroot;
    ed.unweighted = unw ? 1 : 0;
    // one embedding pass over all slices (first pass: column sums only)
    auto embed_all = [&](bool sums_only) -> int {
        for (int64_t k = 0; k < nslices; k++) {
            const int64_t p0 = k * slice, cnt = std::min<int64_t>(slice, c->n - p0);
            if (nslices > 1 && !lists) {
                int r = build_walk_slice(c, CMS_STORE_F32, binned, (int)p0, (int)cnt);
                if (r) return r;
            }
            EmbedData e = ed;
            e.ab = (const float *)c->d_ab;
            if (lists)
                e.ll = LeafLists{c->d_ll_pos, c->d_ll_val, c->d_ll_start + p0 * nseg, c->d_ll_end + p0,
                                 c->d_ll_top + p0, c->npad};
            e.n = (int)cnt;
            e.npad = (cnt + 63) / 64 * 64;
            e.sums_only = sums_only ? 1 : 0;
            e.Y = c->d_Y ? c->d_Y + p0 : nullptr;
            e.bm = c->d_bm ? c->d_bm + (p0 / kSmBlock) * c->nwords : nullptr;
            e.lbm = (c->d_lbm && !sums_only) ? c->d_lbm + (p0 / kSmBlock) * c->nwords : nullptr;
            e.colsum = c->d_colsum + p0;
            e.colsum_f64 = c->d_mass + p0;
            HIPCALL(c, launch_embed(e, c->stream));
        }
        return CMS_OK;
    };
    if (unw) {
        // column sums of y (shared weight bound) from a first, write-free pass
        HIPCALL(c, hipMemsetAsync(c->d_mass, 0, sizeof(double) * std::max<int64_t>(c->npad, 1), c->stream));
        rc = embed_all(true);
        if (rc) return rc;
    } else {
        // a column sum of y equals the sample's consumed leaf mass
        HIPCALL(c, launch_sample_mass(c->d_rowptr, c->d_col, c->d_val, c->n, c->d_leaf_uses, c->d_mass, c->stream));
    }
    std::vector<double> mass(std::max(c->n, 1), 0.0);
    if (c->n > 0)
        HIPCALL(c, hipMemcpyAsync(mass.data(), c->d_mass, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    double maxm = 0.0;
    for (int s = 0; s < c->n; s++) maxm = std::max(maxm, mass[s]);
    c->max_mass = maxm;
    const int K = choose_scale_bits(maxm, nrows);
    c->scale_bits = K;
    const int64_t nblk = c->npad / kSmBlock;
    const size_t nbm = (size_t)std::max<int64_t>(nblk, 1) * std::max<int64_t>(c->nwords, 1);
    // Y holds one more row than the entries: an all-zero row the split tiles
    // stage in place of missing rows (cleared here; the embedding writes
    // only the entries' rows)
    HIPCALL(c, dreserve(c->d_Y, c->cap_Y, (size_t)(std::max<int64_t>(nrows, 0) + 1) * std::max<int64_t>(c->npad, 1)));
    HIPCALL(c, hipMemsetAsync(c->d_Y + (size_t)nrows * c->npad, 0, sizeof(uint32_t) * c->npad, c->stream));
    HIPCALL(c, dreserve(c->d_bm, c->cap_bm, nbm));
    HIPCALL(c, hipMemsetAsync(c->d_bm, 0, sizeof(uint32_t) * nbm, c->stream));
    // light-entry bitmaps of the split tiles (filled by the embedding)
    HIPCALL(c, dreserve(c->d_lbm, c->cap_lbm, nbm));
    HIPCALL(c, hipMemsetAsync(c->d_lbm, 0, sizeof(uint32_t) * nbm, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_colsum, 0, sizeof(unsigned long long) * std::max<int64_t>(c->npad, 1), c->stream));
    ed.scale = std::ldexp(1.0, K);
    rc = embed_all(false);
    if (rc) return rc;
    // records of the light entries in rank order, packed block after block:
    // the prefixes give each block's offset and the count, which sizes the
    // allocation (the bound nwords * 32 per block is about 10x the count)
    HIPCALL(c, dreserve(c->d_pl, c->cap_pl, nbm));
    HIPCALL(c, dreserve(c->d_croff, c->cap_croff, (size_t)std::max<int64_t>(nblk, 1) + 1));
    HIPCALL(c, launch_light_prefix(c->d_lbm, nblk, c->nwords, c->d_pl, c->d_croff, c->stream));
    if (!c->nrec_known || c->nrec_metric != metric || c->nrec_binned != binned) {
        int64_t nrec = 0;
        if (nblk > 0) {
            HIPCALL(c, hipMemcpyAsync(&nrec, c->d_croff + nblk, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
            HIPCALL(c, hipStreamSynchronize(c->stream));
        }
        c->nrec = nrec;
        c->nrec_known = true;
        c->nrec_metric = metric;
        c->nrec_binned = binned;
    }
    HIPCALL(c, dreserve(c->d_crl, c->cap_crl, (size_t)std::max<int64_t>(c->nrec, 1)));
    HIPCALL(c, launch_light_records(c->d_Y, c->npad, c->d_lbm, nblk, c->nwords, c->d_pl, c->d_croff, c->d_crl,
                                    c->nrec, c->stream));
    c->sm_rows = nrows;
    c->sm_metric = metric;
    c->sm_ready = true;
    return CMS_OK;
}

SmData sm_data(const cms_ctx *c) {
    SmData d{};
    d.Y = c->d_Y;
    d.bm = c->d_bm;
    d.nrows = c->sm_rows;
    d.npad = c->npad;
    d.nwords = c->nwords;
    d.wide_rows = c->opts.wide_rows;
    d.n = c->n;
    d.inv_scale = 1.0 / (100.0 * std::ldexp(1.0, c->scale_bits));
    d.counter = c->d_counter;
    d.unweighted = c->sm_metric == CMS_METRIC_UNWEIGHTED ? 1 : 0;
    d.colsum = c->d_colsum;
    d.perm = c->sm_binned ? c->d_perm : nullptr;
    d.pos = c->sm_binned ? c->d_pos : nullptr;
    d.lbm = c->d_lbm;
    d.crl = c->d_crl;
    d.croff = c->d_croff;
    d.pl = c->d_pl;
    d.zrow = c->sm_rows;  // the zero row after the entries' rows
    return d;
}

// Normalised leaf abundances, block bitmaps over leaves and per-sample
// totals for the leaf metrics.
int ensure_leaf_layout(cms_ctx *c) {
    if (c->leaf_ready) return CMS_OK;
    const int64_t npad = ((int64_t)c->n + 63) / 64 * 64;
    const int L = c->tree.leaf_n;
    const int64_t nwords = ((int64_t)L + 31) / 32;
    const size_t nP = (size_t)std::max<int64_t>((int64_t)L * npad, 1);
    const size_t nbm = (size_t)std::max<int64_t>(npad / kSmBlock * nwords, 1);
    HIPCALL(c, dreserve(c->d_leafP, c->cap_leafP, nP));
    HIPCALL(c, dreserve(c->d_leafbm, c->cap_leafbm, nbm));
    HIPCALL(c, dreserve(c->d_leafstats, c->cap_leafstats, (size_t)std::max<int64_t>(3 * npad, 1)));
    HIPCALL(c, hipMemsetAsync(c->d_leafP, 0, sizeof(float) * nP, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_leafbm, 0, sizeof(uint32_t) * nbm, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_leafstats, 0, sizeof(double) * std::max<int64_t>(3 * npad, 1), c->stream));
    LeafData d{};
    d.P = c->d_leafP;
    d.bm = c->d_leafbm;
    d.stats = c->d_leafstats;
    d.npad = npad;
    d.nwords = nwords;
    d.n = c->n;
    d.leaf_n = L;
    HIPCALL(c, launch_leaf_prepare(c->d_rowptr, c->d_col, c->d_val, c->n, d, c->stream));
    c->leaf_ready = true;
    return CMS_OK;
}

// Entry lists, per-sample FP32 totals and the visiting order of the merge
// kernels; cosine merges the values, the others a / sum a.
int ensure_leaf_merge(cms_ctx *c) {
    const int layout = c->opts.metric == CMS_METRIC_COS ? 0 : 1;
    if (c->lm_layout == layout) return CMS_OK;
    const int n = c->n;
    std::vector<int64_t> rp((size_t)n + 1, 0);
    if (n > 0) HIPCALL(c, hipMemcpy(rp.data(), c->d_rowptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    // most entries first: the 8 x 8 pairs of a wave merge lists of similar length
    std::vector<int32_t> perm((size_t)n);
    for (int s = 0; s < n; s++) perm[s] = s;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int32_t a, int32_t b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
    HIPCALL(c, dalloc(c->d_lment, (size_t)std::max<int64_t>(c->nnz, 1)));
    HIPCALL(c, dalloc(c->d_lmstat, (size_t)std::max(n, 1)));
    HIPCALL(c, dalloc(c->d_lmperm, (size_t)std::max(n, 1)));
    if (n > 0) HIPCALL(c, hipMemcpy(c->d_lmperm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCALL(c, launch_leaf_merge_prepare(c->d_rowptr, c->d_col, c->d_val, n, layout == 1, c->d_lment, c->d_lmstat,
                                         c->stream));
    c->lm_layout = layout;
    return CMS_OK;
}

// JSD in the reference's arithmetic needs the host logf's exceptions on the
// device (built once per process, uploaded once per context).
int ensure_logf_table(cms_ctx *c) {
    if (c->logf_ntab >= 0) return CMS_OK;
    const LogfTable &t = host_logf_table();
    HIPCALL(c, dalloc(c->d_logf_bm, t.bitmap.size()));
    HIPCALL(c, dalloc(c->d_logf_blocks, t.blocks.size()));
    HIPCALL(c, hipMemcpy(c->d_logf_blocks, t.blocks.data(), sizeof(uint32_t) * t.blocks.size(), hipMemcpyHostToDevice));
    std::memcpy(&c->logf_ln2, &t.ln2_bits, 4);
    HIPCALL(c, dalloc(c->d_logf_pieces, 2 * (size_t)kLogPieces));
    HIPCALL(c, hipMemcpy(c->d_logf_pieces, t.invc, sizeof(double) * kLogPieces, hipMemcpyHostToDevice));
    HIPCALL(c, hipMemcpy(c->d_logf_pieces + kLogPieces, t.logc, sizeof(double) * kLogPieces, hipMemcpyHostToDevice));
    HIPCALL(c, dalloc(c->d_logf_tab, std::max<size_t>(t.table.size(), 1)));
    HIPCALL(c, hipMemcpy(c->d_logf_bm, t.bitmap.data(), sizeof(uint32_t) * t.bitmap.size(), hipMemcpyHostToDevice));
    if (!t.table.empty())
        HIPCALL(c, hipMemcpy(c->d_logf_tab, t.table.data(), sizeof(uint64_t) * t.table.size(), hipMemcpyHostToDevice));
    c->logf_ntab = (int)t.table.size();
    return CMS_OK;
}

LeafData leaf_data(const cms_ctx *c) {
    LeafData d{};
    d.P = c->d_leafP;
    d.bm = c->d_leafbm;
    d.stats = c->d_leafstats;
    d.npad = ((int64_t)c->n + 63) / 64 * 64;
    d.nwords = ((int64_t)c->tree.leaf_n + 31) / 32;
    d.n = c->n;
    d.leaf_n = c->tree.leaf_n;
    return d;
}

int prepare(cms_ctx *c) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (is_leaf_metric(c)) {
        const bool merge_ready = c->lm_layout == (c->opts.metric == CMS_METRIC_COS ? 0 : 1) &&
                                 (c->opts.metric != CMS_METRIC_JSD || c->logf_ntab >= 0);
        if (use_leaf_merge(c) ? merge_ready : c->leaf_ready) return CMS_OK;
        if (c->has_negative) return fail(c, CMS_E_UNSUPPORTED, "-M 2..5 on the GPU need non-negative abundances");
        HIPCALL(c, hipEventRecord(c->evp0, c->stream));
        rc = use_leaf_merge(c) ? ensure_leaf_merge(c) : ensure_leaf_layout(c);
        if (!rc && use_leaf_merge(c) && c->opts.metric == CMS_METRIC_JSD) rc = ensure_logf_table(c);
    } else if (use_sm(c)) {
        if (c->sm_ready && c->sm_metric == c->opts.metric && c->sm_binned == (c->opts.bin != 0)) return CMS_OK;
        HIPCALL(c, hipEventRecord(c->evp0, c->stream));
        rc = ensure_sharedmin(c);
    } else {
        if (c->tree.nslots > walk_max_slots())
            return fail(c, CMS_E_UNSUPPORTED,
                        "schedule keeps " + std::to_string(c->tree.nslots) + " registers live; the walk kernel holds " +
                            std::to_string(walk_max_slots()) + " (re-order the tree, e.g. Sethi-Ullman)");
        if (c->ab_storage == walk_storage(c) && !c->ab_binned) return CMS_OK;
        HIPCALL(c, hipEventRecord(c->evp0, c->stream));
        rc = ensure_walk_layout(c, walk_storage(c), false);
    }
    if (rc) return rc;
    HIPCALL(c, hipEventRecord(c->evp1, c->stream));
    c->prep_pending = true;
    return CMS_OK;
}

int upload_csr(cms_ctx *c, int n, const int64_t *rowptr, const int32_t *col, const float *val) {
    drop_layouts(c);
    dfree(c->d_rowptr);
    dfree(c->d_col);
    dfree(c->d_val);
    const int64_t nnz = rowptr[n];
    HIPCALL(c, dalloc(c->d_rowptr, (size_t)n + 1));
    HIPCALL(c, dalloc(c->d_col, (size_t)std::max<int64_t>(nnz, 1)));
    HIPCALL(c, dalloc(c->d_val, (size_t)std::max<int64_t>(nnz, 1)));
    HIPCALL(c, hipMemcpy(c->d_rowptr, rowptr, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyHostToDevice));
    if (nnz > 0) {
        HIPCALL(c, hipMemcpy(c->d_col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
        HIPCALL(c, hipMemcpy(c->d_val, val, sizeof(float) * nnz, hipMemcpyHostToDevice));
    }
    c->n = n;
    c->nnz = nnz;
    return CMS_OK;
}

int ensure_out(cms_ctx *c, size_t count) {
    if (count <= c->out_cap && c->d_out) return CMS_OK;
    HIPCALL(c, dalloc(c->d_out, std::max<size_t>(count, 1)));
    c->out_cap = std::max<size_t>(count, 1);
    return CMS_OK;
}

}  // namespace

extern "C" {

void cms_opts_default(cms_opts *o) {
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->kernel = CMS_KERNEL_WALK;
    o->storage = CMS_STORE_F32;
}

namespace {
const char *opts_error(const cms_opts *o) {
    if (o->kernel != CMS_KERNEL_WALK && o->kernel != CMS_KERNEL_SHAREDMIN) return "unknown kernel";
    if ((o->tile_a != 0 && o->tile_a != kSmBlock) || (o->tile_b != 0 && o->tile_b != kSmBlock))
        return "tile_a / tile_b: the shared-minimum tiles are 32 x 32 samples per wave (0 or 32)";
    if (o->chunk != 0 && o->chunk != 8 && o->chunk != 16 && o->chunk != 32) return "chunk must be 0, 8, 16 or 32";
    if (o->variant < 0 || o->variant > kSmVariants) return "variant must be 0..16";
    if (o->variant > 0 && o->chunk != 0 && kSmVariantTable[o->variant - 1].chunk != o->chunk)
        return "variant and chunk disagree";
    if (o->metric < CMS_METRIC_WEIGHTED || o->metric > CMS_METRIC_BC) return "unknown metric";
    if (o->storage < CMS_STORE_F32 || o->storage > CMS_STORE_F16) return "unknown storage";
    if (o->embed_slice < 0) return "embed_slice must be >= 0";
    if (o->embed_input != 0 && o->embed_input != 1) return "embed_input must be 0 or 1";
    if (o->wide_rows != 0 && o->wide_rows != 1) return "wide_rows must be 0 or 1";
    return nullptr;
}
}  // namespace

int cms_create(int device, const cms_opts *o, cms_ctx **out) {
    if (!out) return fail(nullptr, CMS_E_INVAL, "cms_create: out is NULL");
    *out = nullptr;
    if (o && opts_error(o)) return fail(nullptr, CMS_E_INVAL, std::string("cms_create: ") + opts_error(o));
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(nullptr, CMS_E_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= count) return fail(nullptr, CMS_E_INVAL, "cms_create: device index out of range");
    cms_ctx *c = new cms_ctx();
    c->device = device;
    cms_opts_default(&c->opts);
    if (o) c->opts = *o;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->evp0) != hipSuccess || hipEventCreate(&c->evp1) != hipSuccess ||
        hipMalloc((void **)&c->d_counter, sizeof(unsigned long long) * kCounters) != hipSuccess) {
        delete c;
        return fail(nullptr, CMS_E_HIP, "cms_create: stream/event creation failed");
    }
    {
        // the no-allocation entry point's scratch on this device (device current)
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (g_scratch_ptr.size() < (size_t)count) g_scratch_ptr.resize(count, nullptr);
        if (!g_scratch_ptr[device]) {
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                cms_destroy(c);
                return fail(nullptr, CMS_E_HIP, std::string("cms_create: hipGetSymbolAddress(scratch): ") +
                                                    hipGetErrorString(e));
            }
            g_scratch_ptr[device] = p;
        }
        c->scratch = g_scratch_ptr[device];
    }
    *out = c;
    return CMS_OK;
}

void cms_destroy(cms_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_layouts(c);
    dfree(c->d_nodes);
    dfree(c->d_chunks);
    dfree(c->d_leaf_lp_ptr);
    dfree(c->d_leaf_lp);
    dfree(c->d_leaf_uses);
    dfree(c->d_embnodes);
    dfree(c->d_embnodes_u);
    dfree(c->d_seg_first);
    dfree(c->d_seg_last);
    dfree(c->d_toprecs);
    dfree(c->d_toprecs_u);
    dfree(c->d_seg_lp);
    dfree(c->d_topidx);
    dfree(c->d_rowptr);
    dfree(c->d_col);
    dfree(c->d_val);
    dfree(c->d_lbm);
    dfree(c->d_pl);
    dfree(c->d_crl);
    dfree(c->d_croff);
    dfree(c->d_out);
    dfree(c->d_prow);
    dfree(c->d_pcol);
    dfree(c->d_counter);
    dfree(c->d_work);
    dfree(c->d_perm);
    dfree(c->d_pos);
    dfree(c->d_logf_bm);
    dfree(c->d_logf_blocks);
    dfree(c->d_logf_tab);
    dfree(c->d_logf_pieces);
    if (c->band_exec) (void)hipGraphExecDestroy(c->band_exec);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evp0) (void)hipEventDestroy(c->evp0);
    if (c->evp1) (void)hipEventDestroy(c->evp1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *cms_last_error(const cms_ctx *c) { return c ? c->err.c_str() : global_error(); }

int cms_set_opts(cms_ctx *c, const cms_opts *o) {
    if (!c || !o) return fail(c, CMS_E_INVAL, "cms_set_opts: null argument");
    if (c->servers > 0) return fail(c, CMS_E_STATE, "the context is serving (destroy its server first)");
    if (opts_error(o)) return fail(c, CMS_E_INVAL, opts_error(o));
    c->opts = *o;
    c->variant = -1;  // chosen again under the new options
    return CMS_OK;
}

int cms_set_tree(cms_ctx *c, int leaf_n, int order_n, const int32_t *o1, const float *d1, const int32_t *o2,
                 const float *d2, const int32_t *od) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    if (c->servers > 0) return fail(c, CMS_E_STATE, "the context is serving (destroy its server first)");
    std::string err;
    CompiledTree t;
    int rc = compile_tree(leaf_n, order_n, o1, d1, o2, d2, od, kWalkLeafChunk, t, err);
    if (rc) return fail(c, rc, err);
    HIPCALL(c, hipSetDevice(c->device));
    drop_layouts(c);
    // leaf -> leaf-stream positions (CSR), and consumption counts
    std::vector<int32_t> ptr((size_t)leaf_n + 1, 0), lp(t.leaf_stream.size()), uses((size_t)std::max(leaf_n, 1), 0);
    for (int32_t l : t.leaf_stream) ptr[(size_t)l + 1]++;
    for (int l = 0; l < leaf_n; l++) {
        uses[l] = ptr[(size_t)l + 1];
        ptr[(size_t)l + 1] += ptr[l];
    }
    {
        std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
        for (size_t k = 0; k < t.leaf_stream.size(); k++) lp[fill[t.leaf_stream[k]]++] = (int32_t)k;
    }
    // active entries (beta != 0) get consecutive rows in schedule order
    std::vector<int32_t> entrow(t.ent_beta.size(), -1);
    int64_t nrows = 0;
    for (size_t e = 0; e < t.ent_beta.size(); e++)
        if (t.ent_beta[e] != 0.0) entrow[e] = (int32_t)nrows++;
    HIPCALL(c, dalloc(c->d_nodes, t.nodes.size()));
    HIPCALL(c, dalloc(c->d_chunks, t.chunks.size()));
    HIPCALL(c, dalloc(c->d_leaf_lp_ptr, ptr.size()));
    HIPCALL(c, dalloc(c->d_leaf_lp, std::max<size_t>(lp.size(), 1)));
    HIPCALL(c, dalloc(c->d_leaf_uses, uses.size()));
    plan_segments(t, kEmbedSegMax);
    std::vector<int32_t> export_seg(t.nodes.size(), -1);
    for (size_t k = 0; k < t.seg_first.size(); k++)
        if (t.seg_out[k]) export_seg[t.seg_last[k]] = (int32_t)k;
    // unweighted form: every edge with w != 0 is a row, coefficient w
    std::vector<int32_t> entrow_u(t.ent_beta.size(), -1);
    int64_t nrows_u = 0;
    for (size_t i = 0; i < t.nodes.size(); i++) {
        if (t.nodes[i].w1 != 0.0f) entrow_u[2 * i] = (int32_t)nrows_u++;
        if (t.nodes[i].w2 != 0.0f) entrow_u[2 * i + 1] = (int32_t)nrows_u++;
    }
    std::vector<EmbedNode> emb(t.nodes.size()), emb_u(t.nodes.size());
    for (size_t i = 0; i < t.nodes.size(); i++) {
        EmbedNode &e = emb[i];
        e.code = t.nodes[i].code;
        e.leafpos = t.nodes[i].leafpos;
        e.w1 = t.nodes[i].w1;
        e.w2 = t.nodes[i].w2;
        e.row1 = entrow[2 * i];
        e.row2 = entrow[2 * i + 1];
        e.pad0 = export_seg[i];
        e.pad1 = 0;
        e.beta1 = t.ent_beta[2 * i];
        e.beta2 = t.ent_beta[2 * i + 1];
        EmbedNode &u = emb_u[i];
        u = e;
        u.row1 = entrow_u[2 * i];
        u.row2 = entrow_u[2 * i + 1];
        u.beta1 = (double)t.nodes[i].w1;
        u.beta2 = (double)t.nodes[i].w2;
    }
    HIPCALL(c, dalloc(c->d_embnodes, std::max<size_t>(emb.size(), 1)));
    HIPCALL(c, dalloc(c->d_embnodes_u, std::max<size_t>(emb_u.size(), 1)));
    auto upload_i32 = [&](int32_t *&d, const std::vector<int32_t> &v) -> hipError_t {
        hipError_t e = dalloc(d, std::max<size_t>(v.size(), 1));
        if (e == hipSuccess && !v.empty()) e = hipMemcpy(d, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice);
        return e;
    };
    HIPCALL(c, upload_i32(c->d_seg_first, t.seg_first));
    HIPCALL(c, upload_i32(c->d_seg_last, t.seg_last));
    {
        // the top nodes' records, contiguous, with their children's segments
        // (leaf children: their top leaf index); the segments' first leaf
        // positions and the top leaf index of every leaf-stream position
        std::vector<EmbedNode> top(t.top_nodes.size()), top_u(t.top_nodes.size());
        std::vector<int32_t> topidx(std::max<size_t>(t.leaf_stream.size(), 1), -1), seg_lp(t.seg_first.size());
        int ntl = 0;
        for (size_t k = 0; k < t.top_nodes.size(); k++) {
            const int i = t.top_nodes[k];
            top[k] = emb[i];
            top_u[k] = emb_u[i];
            int32_t src[2] = {t.top_src[2 * k], t.top_src[2 * k + 1]};
            int lp = t.nodes[i].leafpos;
            for (int j = 0; j < 2; j++)
                if (t.nodes[i].code & (j == 0 ? kC1Leaf : kC2Leaf)) {
                    topidx[lp++] = ntl;
                    src[j] = ntl++;
                }
            top[k].pad0 = top_u[k].pad0 = src[0];
            top[k].pad1 = top_u[k].pad1 = src[1];
        }
        for (size_t k = 0; k < t.seg_first.size(); k++) seg_lp[k] = t.nodes[t.seg_first[k]].leafpos;
        c->ntop_leaves = ntl;
        HIPCALL(c, upload_i32(c->d_topidx, topidx));
        HIPCALL(c, upload_i32(c->d_seg_lp, seg_lp));
        HIPCALL(c, dalloc(c->d_toprecs, std::max<size_t>(top.size(), 1)));
        HIPCALL(c, dalloc(c->d_toprecs_u, std::max<size_t>(top.size(), 1)));
        if (!top.empty()) {
            HIPCALL(c, hipMemcpy(c->d_toprecs, top.data(), sizeof(EmbedNode) * top.size(), hipMemcpyHostToDevice));
            HIPCALL(c, hipMemcpy(c->d_toprecs_u, top_u.data(), sizeof(EmbedNode) * top.size(), hipMemcpyHostToDevice));
        }
    }
    if (!t.nodes.empty()) {
        HIPCALL(c, hipMemcpy(c->d_nodes, t.nodes.data(), sizeof(WalkNode) * t.nodes.size(), hipMemcpyHostToDevice));
        HIPCALL(c, hipMemcpy(c->d_chunks, t.chunks.data(), sizeof(WalkChunk) * t.chunks.size(), hipMemcpyHostToDevice));
        HIPCALL(c, hipMemcpy(c->d_embnodes, emb.data(), sizeof(EmbedNode) * emb.size(), hipMemcpyHostToDevice));
        HIPCALL(c, hipMemcpy(c->d_embnodes_u, emb_u.data(), sizeof(EmbedNode) * emb_u.size(), hipMemcpyHostToDevice));
    }
    HIPCALL(c, hipMemcpy(c->d_leaf_lp_ptr, ptr.data(), sizeof(int32_t) * ptr.size(), hipMemcpyHostToDevice));
    if (!lp.empty()) HIPCALL(c, hipMemcpy(c->d_leaf_lp, lp.data(), sizeof(int32_t) * lp.size(), hipMemcpyHostToDevice));
    HIPCALL(c, hipMemcpy(c->d_leaf_uses, uses.data(), sizeof(int32_t) * uses.size(), hipMemcpyHostToDevice));
    c->tree = std::move(t);
    c->h_leaf_uses = std::move(uses);
    c->nrows = nrows;
    c->nrows_u = nrows_u;
    c->has_tree = true;
    c->stats.nslots = c->tree.nslots;
    c->stats.entries = (int)nrows;
    return CMS_OK;
}

int cms_set_samples_csr(cms_ctx *c, int n, const int64_t *rowptr, const int32_t *leaf, const float *val) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    if (c->servers > 0) return fail(c, CMS_E_STATE, "the context is serving (destroy its server first)");
    if (!c->has_tree) return fail(c, CMS_E_STATE, "cms_set_samples before cms_set_tree");
    if (n < 0 || !rowptr || (rowptr[n] > 0 && (!leaf || !val))) return fail(c, CMS_E_INVAL, "cms_set_samples_csr: bad arrays");
    if (rowptr[0] != 0) return fail(c, CMS_E_INVAL, "rowptr[0] must be 0");
    bool neg = false;
    for (int s = 0; s < n; s++) {
        if (rowptr[s + 1] < rowptr[s]) return fail(c, CMS_E_INVAL, "rowptr not non-decreasing");
        for (int64_t k = rowptr[s]; k < rowptr[s + 1]; k++) {
            if (leaf[k] < 0 || leaf[k] >= c->tree.leaf_n) return fail(c, CMS_E_INVAL, "leaf index out of range");
            if (!std::isfinite(val[k])) return fail(c, CMS_E_INVAL, "abundances must be finite");
            if (!(val[k] >= 0.0f)) neg = true;
        }
        // a leaf listed twice in one row would be written twice (last wins) in
        // the dense form; reject to keep both forms identical
        for (int64_t k = rowptr[s] + 1; k < rowptr[s + 1]; k++)
            if (leaf[k] <= leaf[k - 1]) return fail(c, CMS_E_INVAL, "CSR columns must be strictly ascending per row");
    }
    HIPCALL(c, hipSetDevice(c->device));
    int rc = upload_csr(c, n, rowptr, leaf, val);
    if (rc) return rc;
    c->has_negative = neg;
    return CMS_OK;
}

int cms_set_samples(cms_ctx *c, int n, const float *abd) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    if (c->servers > 0) return fail(c, CMS_E_STATE, "the context is serving (destroy its server first)");
    if (!c->has_tree) return fail(c, CMS_E_STATE, "cms_set_samples before cms_set_tree");
    if (n < 0 || (n > 0 && c->tree.leaf_n > 0 && !abd)) return fail(c, CMS_E_INVAL, "cms_set_samples: bad arrays");
    const int L = c->tree.leaf_n;
    std::vector<int64_t> rowptr((size_t)n + 1, 0);
    std::vector<int64_t> cnt(n, 0);
    for (int s = 0; s < n; s++) {
        int64_t k = 0;
        for (int l = 0; l < L; l++) k += abd[(size_t)s * L + l] != 0.0f;
        cnt[s] = k;
    }
    for (int s = 0; s < n; s++) rowptr[s + 1] = rowptr[s] + cnt[s];
    std::vector<int32_t> col((size_t)std::max<int64_t>(rowptr[n], 1));
    std::vector<float> val((size_t)std::max<int64_t>(rowptr[n], 1));
    for (int s = 0; s < n; s++) {
        int64_t k = rowptr[s];
        for (int l = 0; l < L; l++) {
            float v = abd[(size_t)s * L + l];
            if (v != 0.0f) {
                col[k] = l;
                val[k] = v;
                k++;
            }
        }
    }
    return cms_set_samples_csr(c, n, rowptr.data(), col.data(), val.data());
}

}  // extern "C"

namespace {

// Persistent-kernel variant: forced by opts.variant, else the default (split
// tiles shared by four waves when the whole triangle has fewer than
// kSmSharedTileMaxTiles tiles, one wave per tile otherwise), or the fastest
// of all variants when opts.autotune is set.
//
// The tuner times every candidate on the same segments of consecutive tiles
// spread over the whole triangle (the centres of equal strata): binned tiles
// get sparser along the triangle, so segments from its start alone mislead.
// A segment holds about one tile per wave slot, which the variants with one
// wave per tile finish in one round of tiles while those that share tiles
// among waves finish it in several: short segments favour the latter, which
// lose over a long triangle. So the candidates are one family, the one the
// default picks by the triangle's size: tiles shared by waves below
// kSmSharedTileMaxTiles tiles, one wave per tile above.
//
// Real work (out_mode 0, `out` the whole packed triangle, and at least two
// segments per candidate): 2 x candidates segments at evenly spaced
// positions, candidate k timing segments k and 2 nc - 1 - k (one from each
// half of the triangle), its time normalised by the rows it visited. The
// tiles stay computed: *done receives the computed ranges [t0, t1) in
// ascending order, and their rows are counted in c->d_counter, which the
// caller zeroes before. Otherwise the timings are discarded (*done empty):
// three segments per candidate at the centres of three strata, the total
// time compared; with out_mode 1 (a band buffer) each segment's results are
// written from the start of `out` (its first packed index as out_base), and
// a segment whose rows do not fit `out` is shortened.
int pick_variant(cms_ctx *c, int64_t nb, int64_t ntiles, float *out, int64_t out_cap, int out_mode,
                 hipStream_t st, std::vector<std::pair<int64_t, int64_t>> *done) {
    done->clear();
    const int64_t all_tiles = nb * (nb + 1) / 2;  // the whole triangle (the default depends on it)
    if (!c->d_work) HIPCALL(c, dalloc(c->d_work, kWorkWords));
    if (c->grid_waves == 0) {
        int cus = 0;
        HIPCALL(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        c->grid_waves = std::max(cus, 1) * 32;  // 8 waves per SIMD
    }
    if (c->opts.variant >= 1 && c->opts.variant <= kSmVariants) {
        c->variant = c->opts.variant - 1;
        return CMS_OK;
    }
    if (c->variant >= 0) return CMS_OK;
    // opts.chunk (if set) restricts the candidates to that LDS chunk
    auto allowed = [&](int v) { return c->opts.chunk == 0 || kSmVariantTable[v].chunk == c->opts.chunk; };
    if (!c->opts.autotune || ntiles <= 0) {
        // few tiles: four waves share each tile (at 2,048 samples, 2,080
        // tiles: 3.2 ms against 4.5 ms with one wave per tile; at 4,096
        // samples one wave per tile is faster)
        c->variant = all_tiles < kSmSharedTileMaxTiles ? kSmSharedTileVariant : kSmDefaultVariant;
        if (!allowed(c->variant))
            for (int v = 0; v < kSmVariants; v++)
                if (allowed(v)) {
                    c->variant = v;
                    break;
                }
        return CMS_OK;
    }
    const int64_t n = c->n;
    auto row_start = [n](int64_t r) { return r * n - r * (r + 1) / 2; };
    // packed-index range [lo, hi) of the stored-position pairs of tiles [t0, t1)
    auto span = [&](int64_t t0, int64_t t1, int64_t &lo, int64_t &hi) {
        int64_t ra, ca, rb, cbb;
        host_decode_upper(t0, nb, ra, ca);
        host_decode_upper(t1 - 1, nb, rb, cbb);
        lo = row_start(std::min(ra * kSmBlock, n));
        hi = row_start(std::min((rb + 1) * kSmBlock, n));
    };
    auto timed = [&](int v, int64_t t0, int64_t len, float &ms) -> int {
        int64_t base = 0;
        if (out_mode == 1) {
            int64_t lo, hi;
            span(t0, t0 + len, lo, hi);
            base = lo;
        }
        HIPCALL(c, hipEventRecord(c->evp0, st));
        HIPCALL(c, launch_zero_counter(c->d_work, st));
        HIPCALL(c, launch_sm_persistent(sm_data(c), v, nb, t0, len, c->d_work, out, out_cap, out_mode, base,
                                        c->grid_waves, st));
        HIPCALL(c, hipEventRecord(c->evp1, st));
        HIPCALL(c, hipEventSynchronize(c->evp1));
        HIPCALL(c, hipEventElapsedTime(&ms, c->evp0, c->evp1));
        return CMS_OK;
    };
    // tiles per segment: about one per wave slot (grid_waves / 2, 16 per CU)
    const int64_t kSeg = std::max<int64_t>(1024, (int64_t)c->grid_waves / 2);
    float best = 1e30f;
    int best_v = -1;
    std::vector<int> cand;
    const bool shared = all_tiles < kSmSharedTileMaxTiles;
    for (int v = 0; v < kSmVariants; v++)
        if (allowed(v) && (kSmVariantTable[v].parts > 1) == shared) cand.push_back(v);
    if (cand.empty())
        for (int v = 0; v < kSmVariants; v++)
            if (allowed(v)) cand.push_back(v);
    const int nc = (int)cand.size();
    for (int v = 0; v < kSmVariants; v++) c->stats.tune_ms[v] = 0.0;
    if (out_mode == 0 && ntiles == all_tiles && ntiles >= 2 * (int64_t)nc * kSeg) {
        std::vector<double> rows(kSmVariants, 0.0);
        std::vector<float> msum(kSmVariants, 0.0f);
        unsigned long long before = 0, after = 0;
        HIPCALL(c, hipMemcpy(&before, c->d_counter, sizeof before, hipMemcpyDeviceToHost));
        const int nseg = 2 * nc;
        for (int k = 0; k < nseg; k++) {
            const int v = cand[k < nc ? k : nseg - 1 - k];
            const int64_t t0 = (ntiles - kSeg) * k / (nseg - 1);
            float ms = 0;
            int rc = timed(v, t0, kSeg, ms);
            if (rc) return rc;
            HIPCALL(c, hipMemcpy(&after, c->d_counter, sizeof after, hipMemcpyDeviceToHost));
            msum[v] += ms;
            rows[v] += (double)(after - before);
            before = after;
            done->push_back({t0, t0 + kSeg});
        }
        for (int v : cand) {
            c->stats.tune_ms[v] = msum[v];
            const double per_row = rows[v] > 0 ? msum[v] / rows[v] : 1e30;
            if (per_row < best) {
                best = (float)per_row;
                best_v = v;
            }
        }
    } else {
        const int nseg = all_tiles >= 3 * kSeg ? 3 : 1;
        std::vector<std::pair<int64_t, int64_t>> segs;
        for (int k = 0; k < nseg; k++) {
            const int64_t len0 = std::min<int64_t>(all_tiles, kSeg);
            const int64_t t0 = nseg == 1 ? 0 : all_tiles * (2 * k + 1) / (2 * nseg) - len0 / 2;
            int64_t len = len0;
            if (out_mode == 1) {
                // the segment's rows must fit the buffer: shorten it
                int64_t lo, hi;
                span(t0, t0 + len, lo, hi);
                while (len > 1 && hi - lo > out_cap) {
                    len /= 2;
                    span(t0, t0 + len, lo, hi);
                }
                if (hi - lo > out_cap) continue;
            }
            segs.push_back({t0, len});
        }
        for (int v : cand) {
            float total = 0;
            for (const auto &sg : segs) {
                float ms = 0;
                int rc = timed(v, sg.first, sg.second, ms);
                if (rc) return rc;
                total += ms;
            }
            c->stats.tune_ms[v] = total;
            if (total < best) {
                best = total;
                best_v = v;
            }
        }
    }
    c->variant = best_v >= 0 ? best_v : kSmDefaultVariant;
    c->prep_pending = false;  // tuning reused the prep events
    return CMS_OK;
}

}  // namespace

extern "C" {

int cms_all_pairs_device(cms_ctx *c, float *d_sim, void *stream) {
    int rc = prepare(c);
    if (rc) return rc;
    if (!d_sim) {
        rc = ensure_out(c, (size_t)c->n * (c->n > 0 ? c->n - 1 : 0) / 2);
        if (rc) return rc;
        d_sim = c->d_out;
    }
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (st != c->stream) {
        // layouts are built on the context stream: order the caller's stream after them
        HIPCALL(c, hipEventRecord(c->evp1, c->stream));
        HIPCALL(c, hipStreamWaitEvent(st, c->evp1, 0));
    }
    const double pairs = (double)c->n * (c->n - 1) / 2.0;
    c->stats.dense_updates = pairs * c->tree.order_n;
    if (is_leaf_metric(c)) {
        if (!c->d_work) HIPCALL(c, dalloc(c->d_work, kWorkWords));
        if (c->grid_waves == 0) {
            int cus = 0;
            HIPCALL(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
            c->grid_waves = std::max(cus, 1) * 32;
        }
        const int64_t npairs = (int64_t)c->n * (c->n > 0 ? c->n - 1 : 0) / 2;
        HIPCALL(c, hipEventRecord(c->ev0, st));
        if (use_leaf_merge(c)) {
            LeafMergeData md{c->d_rowptr, c->d_lment, c->d_lmstat, c->d_lmperm, c->n,
                             c->d_logf_bm, c->d_logf_blocks, c->d_logf_tab, std::max(c->logf_ntab, 0), c->logf_ln2,
                             c->d_logf_pieces};
            HIPCALL(c, launch_leaf_merge_all_pairs(md, c->opts.metric, d_sim, npairs, st));
        } else {
            HIPCALL(c, launch_zero_counter(c->d_work, st));
            HIPCALL(c, launch_leaf_all_pairs(leaf_data(c), c->opts.metric, c->d_work, d_sim, npairs, c->grid_waves, st));
        }
        HIPCALL(c, hipEventRecord(c->ev1, st));
        c->stats.pair_node_updates = c->stats.valu_ops = c->stats.bytes_moved = 0.0;
        c->kern_pending = true;
        return CMS_OK;
    }
    if (!use_sm(c)) HIPCALL(c, hipEventRecord(c->ev0, st));
    if (use_sm(c)) {
        const int64_t nb = (c->n + kSmBlock - 1) / kSmBlock;
        const int64_t ntiles = nb * (nb + 1) / 2;
        const int64_t npairs = (int64_t)c->n * (c->n > 0 ? c->n - 1 : 0) / 2;  // d_sim holds the whole triangle
        // the kernel time and visited rows include tuning launches that computed real tiles
        HIPCALL(c, hipEventRecord(c->ev0, st));
        HIPCALL(c, launch_zero_counters(c->d_counter, kCounters, st));
        std::vector<std::pair<int64_t, int64_t>> done;  // tile ranges the tuner computed for real
        rc = pick_variant(c, nb, ntiles, d_sim, npairs, 0, st, &done);
        if (rc) return rc;
        if (done.empty()) {
            HIPCALL(c, hipEventRecord(c->ev0, st));  // discarded tuning launches are not counted
            HIPCALL(c, launch_zero_counters(c->d_counter, kCounters, st));
        }
        // the tiles between the tuned ranges
        int64_t t0 = 0;
        done.push_back({ntiles, ntiles});
        for (const auto &r : done) {
            if (r.first > t0) {
                HIPCALL(c, launch_zero_counter(c->d_work, st));
                HIPCALL(c, launch_sm_persistent(sm_data(c), c->variant, nb, t0, r.first - t0, c->d_work, d_sim, npairs,
                                                0, 0, c->grid_waves, st));
            }
            t0 = std::max(t0, r.second);
        }
        c->stats.tiles = ntiles;
    } else {
        HIPCALL(c, launch_walk_all_pairs(walk_data(c), d_sim, st));
        const int64_t nb = (c->n + kWalkBlock - 1) / kWalkBlock;
        c->stats.tiles = nb * (nb + 1) / 2;
        c->stats.pair_node_updates = (double)c->stats.tiles * 64.0 * c->tree.order_n;
        // weighted: 2 min, 2 add, 4 sub, 4 mul, 2 add; unweighted: per child
        // 2 selects and 2 adds, 2 presence ORs
        c->stats.valu_ops = c->stats.pair_node_updates * (c->opts.metric == CMS_METRIC_UNWEIGHTED ? 10.0 : 14.0);
        const double elem = c->ab_storage == CMS_STORE_F32 ? 4.0 : 2.0;
        c->stats.bytes_moved = (double)c->stats.tiles * 16.0 * elem * (double)c->tree.leaf_stream.size();
    }
    HIPCALL(c, hipEventRecord(c->ev1, st));
    c->kern_pending = true;
    return CMS_OK;
}

int cms_copy_result(cms_ctx *c, float *host, int64_t count) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    if (count < 0 || (size_t)count > c->out_cap || (count > 0 && !host))
        return fail(c, CMS_E_INVAL, "cms_copy_result: count exceeds the result buffer");
    if (count == 0) return CMS_OK;
    HIPCALL(c, hipSetDevice(c->device));
    HIPCALL(c, hipMemcpyAsync(host, c->d_out, sizeof(float) * count, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return CMS_OK;
}

int cms_synchronize(cms_ctx *c) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    HIPCALL(c, hipSetDevice(c->device));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return CMS_OK;
}

int cms_device_count(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        set_global_error(std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
        return e == hipErrorNoDevice ? 0 : CMS_E_HIP;
    }
    return n;
}

int cms_calibrate_copy(cms_ctx *c, int64_t bytes, int reps, double *ms) {
    if (!c || bytes <= 0 || reps <= 0 || !ms) return fail(c, CMS_E_INVAL, "cms_calibrate_copy: bad argument");
    HIPCALL(c, hipSetDevice(c->device));
    uint8_t *a = nullptr, *b = nullptr;
    HIPCALL(c, hipMalloc((void **)&a, (size_t)bytes));
    hipError_t e = hipMalloc((void **)&b, (size_t)bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        HIPCALL(c, e);
    }
    HIPCALL(c, hipMemsetAsync(a, 1, (size_t)bytes, c->stream));
    HIPCALL(c, launch_calib_copy(a, b, bytes, c->stream));  // warm-up
    HIPCALL(c, hipEventRecord(c->evp0, c->stream));
    for (int r = 0; r < reps; r++) HIPCALL(c, launch_calib_copy(a, b, bytes, c->stream));
    HIPCALL(c, hipEventRecord(c->evp1, c->stream));
    HIPCALL(c, hipEventSynchronize(c->evp1));
    float t = 0;
    HIPCALL(c, hipEventElapsedTime(&t, c->evp0, c->evp1));
    *ms = (double)t / reps;
    dfree(a);
    dfree(b);
    return CMS_OK;
}

// Rebuild the device layouts of the current kernel from the uploaded samples
// (the walk layout; for the shared-minimum form also the embedding and the
// bitmaps), reusing allocations.
int cms_prepare(cms_ctx *c, int force) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    if (c->servers > 0) return fail(c, CMS_E_STATE, "the context is serving (destroy its server first)");
    if (force) {
        c->ab_storage = -1;
        c->sm_ready = false;
        c->leaf_ready = false;
    }
    return prepare(c);
}

int cms_all_pairs(cms_ctx *c, float *sim) {
    int rc = prepare(c);
    if (rc) return rc;
    const size_t np = (size_t)c->n * (c->n > 0 ? c->n - 1 : 0) / 2;
    if (np == 0) return CMS_OK;
    if (!sim) return fail(c, CMS_E_INVAL, "cms_all_pairs: null output");
    rc = ensure_out(c, np);
    if (rc) return rc;
    rc = cms_all_pairs_device(c, c->d_out, nullptr);
    if (rc) return rc;
    HIPCALL(c, hipMemcpyAsync(sim, c->d_out, sizeof(float) * np, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return CMS_OK;
}

int cms_sim_pairs(cms_ctx *c, int64_t n_pairs, const int32_t *row, const int32_t *col, float *sim) {
    int rc = prepare(c);
    if (rc) return rc;
    if (is_leaf_metric(c)) return fail(c, CMS_E_UNSUPPORTED, "-M 2..5 run through cms_all_pairs only");
    if (n_pairs < 0 || (n_pairs > 0 && (!row || !col || !sim))) return fail(c, CMS_E_INVAL, "cms_sim_pairs: bad arrays");
    if (n_pairs == 0) return CMS_OK;
    for (int64_t i = 0; i < n_pairs; i++)
        if (row[i] < 0 || row[i] >= c->n || col[i] < 0 || col[i] >= c->n)
            return fail(c, CMS_E_INVAL, "cms_sim_pairs: sample index out of range");
    // pair lists live in context buffers (grown, reused), so no exit path leaks
    HIPCALL(c, dreserve(c->d_prow, c->cap_prow, (size_t)n_pairs));
    HIPCALL(c, dreserve(c->d_pcol, c->cap_pcol, (size_t)n_pairs));
    rc = ensure_out(c, (size_t)n_pairs);
    if (rc) return rc;
    HIPCALL(c, hipMemcpyAsync(c->d_prow, row, sizeof(int32_t) * n_pairs, hipMemcpyHostToDevice, c->stream));
    HIPCALL(c, hipMemcpyAsync(c->d_pcol, col, sizeof(int32_t) * n_pairs, hipMemcpyHostToDevice, c->stream));
    if (use_sm(c))
        HIPCALL(c, launch_sm_pairs(sm_data(c), n_pairs, c->d_prow, c->d_pcol, c->d_out, c->stream));
    else
        HIPCALL(c, launch_walk_pairs(walk_data(c), n_pairs, c->d_prow, c->d_pcol, c->d_out, c->stream));
    HIPCALL(c, hipMemcpyAsync(sim, c->d_out, sizeof(float) * n_pairs, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return CMS_OK;
}

// cms_sim_pairs without device allocation: the layouts must already be
// built (cms_prepare, or any earlier computing call, on the same kernel and
// options); the pair lists and results are staged in the device's static
// scratch.
int cms_sim_pairs_noalloc(cms_ctx *c, int64_t n_pairs, const int32_t *row, const int32_t *col, float *sim) {
    if (!c) return fail(nullptr, CMS_E_INVAL, "null context");
    if (is_leaf_metric(c)) return fail(c, CMS_E_UNSUPPORTED, "-M 2..5 run through cms_all_pairs only");
    if (n_pairs < 0 || (n_pairs > 0 && (!row || !col || !sim)))
        return fail(c, CMS_E_INVAL, "cms_sim_pairs_noalloc: bad arrays");
    if (n_pairs > (int64_t)CMS_NOALLOC_MAX_PAIRS)
        return fail(c, CMS_E_INVAL, "cms_sim_pairs_noalloc: " + std::to_string(n_pairs) +
                                        " pairs exceed the scratch of CMS_NOALLOC_MAX_PAIRS = " +
                                        std::to_string((int64_t)CMS_NOALLOC_MAX_PAIRS));
    if (n_pairs == 0) return CMS_OK;
    if (!c->has_tree || c->n < 0) return fail(c, CMS_E_STATE, "cms_sim_pairs_noalloc: no tree or samples loaded");
    const bool ready = use_sm(c) ? (c->sm_ready && c->sm_metric == c->opts.metric && c->sm_binned == (c->opts.bin != 0))
                                 : (c->ab_storage == walk_storage(c) && !c->ab_binned);
    if (!ready)
        return fail(c, CMS_E_STATE,
                    "cms_sim_pairs_noalloc: device layouts not built for the current options (call cms_prepare first)");
    for (int64_t i = 0; i < n_pairs; i++)
        if (row[i] < 0 || row[i] >= c->n || col[i] < 0 || col[i] >= c->n)
            return fail(c, CMS_E_INVAL, "cms_sim_pairs_noalloc: sample index out of range");
    HIPCALL(c, hipSetDevice(c->device));
    int32_t *d_r = (int32_t *)c->scratch;
    int32_t *d_c = d_r + CMS_NOALLOC_MAX_PAIRS;
    float *d_o = (float *)(d_c + CMS_NOALLOC_MAX_PAIRS);
    HIPCALL(c, hipMemcpyAsync(d_r, row, sizeof(int32_t) * n_pairs, hipMemcpyHostToDevice, c->stream));
    HIPCALL(c, hipMemcpyAsync(d_c, col, sizeof(int32_t) * n_pairs, hipMemcpyHostToDevice, c->stream));
    if (use_sm(c))
        HIPCALL(c, launch_sm_pairs(sm_data(c), n_pairs, d_r, d_c, d_o, c->stream));
    else
        HIPCALL(c, launch_walk_pairs(walk_data(c), n_pairs, d_r, d_c, d_o, c->stream));
    HIPCALL(c, hipMemcpyAsync(sim, d_o, sizeof(float) * n_pairs, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return CMS_OK;
}

int cms_tile(cms_ctx *c, int r0, int nr, int c0, int nc, float *out) {
    int rc = prepare(c);
    if (rc) return rc;
    if (is_leaf_metric(c)) return fail(c, CMS_E_UNSUPPORTED, "-M 2..5 run through cms_all_pairs only");
    if (r0 < 0 || c0 < 0 || nr < 0 || nc < 0 || (int64_t)r0 + nr > c->n || (int64_t)c0 + nc > c->n)
        return fail(c, CMS_E_INVAL, "cms_tile: range outside the sample set");
    const size_t cnt = (size_t)nr * nc;
    if (cnt == 0) return CMS_OK;
    if (!out) return fail(c, CMS_E_INVAL, "cms_tile: null output");
    rc = ensure_out(c, cnt);
    if (rc) return rc;
    if (use_sm(c) && c->sm_binned) {
        // stored order differs from the caller's: go through the pair list
        std::vector<int32_t> pr(cnt), pc(cnt);
        for (int i = 0; i < nr; i++)
            for (int j = 0; j < nc; j++) {
                pr[(size_t)i * nc + j] = r0 + i;
                pc[(size_t)i * nc + j] = c0 + j;
            }
        rc = cms_sim_pairs(c, (int64_t)cnt, pr.data(), pc.data(), out);
        if (rc) return rc;
        for (int i = 0; i < nr; i++)
            for (int j = 0; j < nc; j++)
                if (r0 + i == c0 + j) out[(size_t)i * nc + j] = 1.0f;
        return CMS_OK;
    }
    if (use_sm(c))
        HIPCALL(c, launch_sm_rect(sm_data(c), r0, nr, c0, nc, c->d_out, c->stream));
    else
        HIPCALL(c, launch_walk_rect(walk_data(c), r0, nr, c0, nc, c->d_out, c->stream));
    HIPCALL(c, hipMemcpyAsync(out, c->d_out, sizeof(float) * cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return CMS_OK;
}

int cms_sample_order(cms_ctx *c, int32_t *order) {
    int rc = prepare(c);
    if (rc) return rc;
    if (!order && c->n > 0) return fail(c, CMS_E_INVAL, "cms_sample_order: null output");
    for (int p = 0; p < c->n; p++) order[p] = (use_sm(c) && c->sm_binned) ? c->h_perm[p] : p;
    return CMS_OK;
}

int cms_all_pairs_stream(cms_ctx *c, int64_t band_pairs, cms_band_fn fn, void *user) {
    int rc = prepare(c);
    if (rc) return rc;
    if (is_leaf_metric(c)) return fail(c, CMS_E_UNSUPPORTED, "-M 2..5 run through cms_all_pairs only");
    if (!use_sm(c)) return fail(c, CMS_E_UNSUPPORTED, "streamed all-pairs runs on the shared-minimum kernel");
    if (!fn) return fail(c, CMS_E_INVAL, "cms_all_pairs_stream: null callback");
    const int64_t n = c->n, nb = (n + kSmBlock - 1) / kSmBlock;
    if (n < 2) return CMS_OK;
    if (band_pairs <= 0) band_pairs = (int64_t)1 << 28;
    // pairs with stored row index in [0, r) = packed index of (r, r+1)
    auto row_start = [&](int64_t r) { return r * n - r * (r + 1) / 2; };
    auto tile_start = [&](int64_t rb) { return rb * nb - rb * (rb - 1) / 2; };
    // bands of whole row blocks, each at least one block, about band_pairs pairs
    std::vector<int64_t> cuts{0};
    while (cuts.back() < nb) {
        int64_t b = cuts.back() + 1;
        while (b < nb && row_start(std::min(b * kSmBlock, n)) - row_start(cuts.back() * kSmBlock) < band_pairs) b++;
        cuts.push_back(b);
    }
    int64_t cap = 0;
    for (size_t k = 0; k + 1 < cuts.size(); k++)
        cap = std::max(cap, row_start(std::min(cuts[k + 1] * kSmBlock, n)) - row_start(cuts[k] * kSmBlock));
    float *dbuf[2] = {nullptr, nullptr}, *hbuf[2] = {nullptr, nullptr};
    hipEvent_t done_k[2] = {nullptr, nullptr}, done_c[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        for (int i = 0; i < 2; i++) {
            if (dbuf[i]) (void)hipFree(dbuf[i]);
            if (hbuf[i]) (void)hipHostFree(hbuf[i]);
            if (done_k[i]) (void)hipEventDestroy(done_k[i]);
            if (done_c[i]) (void)hipEventDestroy(done_c[i]);
        }
    };
#define STREAMCALL(expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            cleanup();                                                                          \
            return fail(c, e_ == hipErrorOutOfMemory ? CMS_E_NOMEM : CMS_E_HIP,                 \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                     \
        }                                                                                       \
    } while (0)
    for (int i = 0; i < 2; i++) {
        STREAMCALL(hipMalloc((void **)&dbuf[i], sizeof(float) * cap));
        STREAMCALL(hipHostMalloc((void **)&hbuf[i], sizeof(float) * cap, hipHostMallocDefault));
        STREAMCALL(hipEventCreateWithFlags(&done_k[i], hipEventDisableTiming));
        STREAMCALL(hipEventCreateWithFlags(&done_c[i], hipEventDisableTiming));
    }
    if (!c->copy_stream) STREAMCALL(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    // choose the variant; the tuner's timings over the whole triangle are
    // written to the first band buffer and discarded (it is filled again)
    std::vector<std::pair<int64_t, int64_t>> tuned_ranges;
    rc = pick_variant(c, nb, tile_start(cuts[1]), dbuf[0], cap, 1, c->stream, &tuned_ranges);
    if (rc) {
        cleanup();
        return rc;
    }
    STREAMCALL(hipEventRecord(c->ev0, c->stream));
    STREAMCALL(launch_zero_counters(c->d_counter, kCounters, c->stream));
    const SmData d = sm_data(c);
    const int nbands = (int)cuts.size() - 1;
    std::vector<int64_t> first(nbands), count(nbands);
    for (int k = 0; k <= nbands; k++) {
        if (k < nbands) {
            const int b = k & 1;
            const int64_t r0 = cuts[k] * kSmBlock, r1 = std::min(cuts[k + 1] * kSmBlock, n);
            first[k] = row_start(r0);
            count[k] = row_start(r1) - first[k];
            const int64_t t0 = tile_start(cuts[k]), t1 = tile_start(cuts[k + 1]);
            if (k >= 2) STREAMCALL(hipStreamWaitEvent(c->stream, done_c[b], 0));  // device buffer free again
            if (c->opts.use_graph) {
                // capture this band's launches and replay them as a graph; later
                // bands update the instantiated graph in place
                hipGraph_t g = nullptr;
                STREAMCALL(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                hipError_t e1 = launch_zero_counter(c->d_work, c->stream);
                hipError_t e2 = launch_sm_persistent(d, c->variant, nb, t0, t1 - t0, c->d_work, dbuf[b], cap, 1, first[k],
                                                     c->grid_waves, c->stream);
                STREAMCALL(hipStreamEndCapture(c->stream, &g));
                STREAMCALL(e1);
                STREAMCALL(e2);
                bool updated = false;
                if (c->band_exec) {
                    hipGraphExecUpdateResult res;
                    hipGraphNode_t err_node = nullptr;
                    updated = hipGraphExecUpdate(c->band_exec, g, &err_node, &res) == hipSuccess &&
                              res == hipGraphExecUpdateSuccess;
                    if (!updated) {
                        (void)hipGraphExecDestroy(c->band_exec);
                        c->band_exec = nullptr;
                    }
                }
                if (!c->band_exec) STREAMCALL(hipGraphInstantiate(&c->band_exec, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
                STREAMCALL(hipGraphLaunch(c->band_exec, c->stream));
            } else {
                STREAMCALL(launch_zero_counter(c->d_work, c->stream));
                STREAMCALL(launch_sm_persistent(d, c->variant, nb, t0, t1 - t0, c->d_work, dbuf[b], cap, 1, first[k],
                                                c->grid_waves, c->stream));
            }
            STREAMCALL(hipEventRecord(done_k[b], c->stream));
            STREAMCALL(hipStreamWaitEvent(c->copy_stream, done_k[b], 0));
            STREAMCALL(hipMemcpyAsync(hbuf[b], dbuf[b], sizeof(float) * count[k], hipMemcpyDeviceToHost,
                                      c->copy_stream));
            STREAMCALL(hipEventRecord(done_c[b], c->copy_stream));
        }
        if (k >= 1) {
            const int pb = (k - 1) & 1;
            STREAMCALL(hipEventSynchronize(done_c[pb]));
            fn(user, first[k - 1], count[k - 1], hbuf[pb]);
        }
    }
    STREAMCALL(hipEventRecord(c->ev1, c->stream));
    STREAMCALL(hipStreamSynchronize(c->stream));
    STREAMCALL(hipStreamSynchronize(c->copy_stream));
#undef STREAMCALL
    cleanup();
    c->stats.tiles = nb * (nb + 1) / 2;
    c->stats.bands = nbands;
    c->stats.dense_updates = (double)n * (n - 1) / 2.0 * c->tree.order_n;
    c->kern_pending = true;
    return CMS_OK;
}

int cms_get_stats(cms_ctx *c, cms_stats *s) {
    if (!c || !s) return CMS_E_INVAL;
    if (c->kern_pending) {
        float ms = 0;
        HIPCALL(c, hipEventSynchronize(c->ev1));
        HIPCALL(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->stats.kernel_ms = ms;
        if (use_sm(c)) {
            unsigned long long cnt[kCounters] = {};
            HIPCALL(c, hipMemcpy(cnt, c->d_counter, sizeof cnt, hipMemcpyDeviceToHost));
            // a staged entry is one pair-entry for all 32 x 32 pairs of the
            // tile (17/16 VALU ops each, 2 x 32 values of Y); a sparse entry
            // costs one minimum and one add per pair with mass on both sides
            const double staged = (double)cnt[kCntStaged];
            const double split_pairs = (double)cnt[kCntSparsePairs] + (double)cnt[kCntHalfPairs];
            c->stats.pair_node_updates = staged * kSmBlock * kSmBlock + split_pairs;
            c->stats.valu_ops = staged * kSmBlock * kSmBlock * kSmOpsPerPairEntry + 2.0 * split_pairs;
            // staged entries read 2 x 32 values of Y, sparse ones two 16-byte
            // records, half ones one block's 32 values and one record
            c->stats.bytes_moved = staged * 2.0 * kSmBlock * 4.0 + (double)cnt[kCntSparseRows] * 32.0 +
                                   (double)cnt[kCntHalfRows] * (kSmBlock * 4.0 + 16.0);
            c->stats.entries_visited = (int64_t)cnt[kCntVisited];
            c->stats.entries_staged = (int64_t)cnt[kCntStaged];
            c->stats.entries_sparse = (int64_t)cnt[kCntSparseRows];
            c->stats.sparse_pairs = (int64_t)cnt[kCntSparsePairs];
            c->stats.entries_half = (int64_t)cnt[kCntHalfRows];
            c->stats.half_pairs = (int64_t)cnt[kCntHalfPairs];
        }
        c->kern_pending = false;
    }
    if (c->prep_pending) {
        float ms = 0;
        HIPCALL(c, hipEventSynchronize(c->evp1));
        HIPCALL(c, hipEventElapsedTime(&ms, c->evp0, c->evp1));
        c->stats.prep_ms = ms;
        c->prep_pending = false;
    }
    *s = c->stats;
    s->scale_bits = c->scale_bits;
    s->variant = c->variant + 1;
    return CMS_OK;
}

}  // extern "C"

// ------------------------------------------------- internal (serving loop)

namespace cms {

int ctx_prepare(cms_ctx *c) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (is_leaf_metric(c)) return fail(c, CMS_E_UNSUPPORTED, "serving runs -M 0 and -M 1 only");
    return ensure_sharedmin(c);
}

SmData ctx_sm_data(const cms_ctx *c) { return sm_data(c); }

int ctx_fail(cms_ctx *c, int code, const std::string &msg) { return fail(c, code, msg); }

void ctx_release_walk_layout(cms_ctx *c) {
    dfree(c->d_ab);
    c->cap_ab = 0;
    c->ab_storage = -1;
}

}  // namespace cms
