// Launch wrappers of the gfx950 kernels (host side). Device pointers only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cms_internal.h"
#include "tile_order.h"

namespace cms {

// Samples per walk block: a wave walks an 8 x 8 tile of pairs (lane = 8 r + c).
constexpr int kWalkBlock = 8;
// Leaf-stream positions staged in LDS per chunk.
constexpr int kWalkLeafChunk = 16;
// Waves (tiles) per walk workgroup.
constexpr int kWalkWaves = 4;
// Live register slots the walk tiles keep in VGPRs (a power of two; larger
// schedules keep them in LDS).
constexpr int kWalkVregSlots = 16;

// Layout of walk abundances: for sample block b (samples 8b..8b+7) and leaf
// stream position lp, value of sample 8b+k at ab[(b * nlp + lp) * 8 + k].
// Stored as float (F32) or as 16-bit words with a per-sample power-of-two
// scale (BF16 / F16).
struct WalkData {
    const WalkNode *nodes;
    const WalkChunk *chunks;
    int nchunks;
    int order_n;
    int nslots;
    int64_t nlp;        // leaf stream length
    const void *ab;     // block-leaf-major values
    const float *inv_scale;  // per-sample 2^-e (16-bit storage), may be null for F32
    int n;              // samples
    int storage;        // CMS_STORE_*
    int unweighted;     // Calc_sim_unweight (F32 storage only)
};

// Largest live-register count the walk kernels can hold in LDS.
int walk_max_slots();
// Largest live-register count of the embedding kernel.
constexpr int kEmbedMaxSlots = 56;
// Largest subtree segment (nodes) one embedding wave walks.
constexpr int kEmbedSegMax = 512;
// Largest transient F32 walk layout built for the embedding; larger sample
// sets are embedded in slices of stored positions.
constexpr size_t kEmbedSliceBytes = (size_t)16 << 30;

// Scatter CSR samples into the walk layout (buffer zeroed by the caller):
// layout position p in [0, cnt) receives sample perm[p0 + p] (p0 + p when
// perm is null), so a slice of stored positions can be laid out alone.
//   leaf_lp_ptr/leaf_lp: leaf -> leaf-stream positions (CSR over leaves)
hipError_t launch_walk_scatter(const int64_t *rowptr, const int32_t *col, const float *val,
                               const int32_t *leaf_lp_ptr, const int32_t *leaf_lp, int64_t nlp, int storage,
                               const float *scale, const int32_t *perm, int p0, int cnt, void *ab, hipStream_t st);
// Per-sample power-of-two scale for 16-bit storage from CSR values.
hipError_t launch_sample_scale(const int64_t *rowptr, const float *val, int n, int storage, float *scale,
                               float *inv_scale, hipStream_t st);

// All pairs of samples, packed upper triangle.
hipError_t launch_walk_all_pairs(const WalkData &d, float *out, hipStream_t st);
// Rectangle rows [r0, r0+nr) x cols [c0, c0+nc), row-major nr x nc.
hipError_t launch_walk_rect(const WalkData &d, int r0, int nr, int c0, int nc, float *out, hipStream_t st);
// Explicit pair list.
hipError_t launch_walk_pairs(const WalkData &d, int64_t n_pairs, const int32_t *row, const int32_t *col,
                             float *out, hipStream_t st);

// ---------------------------------------------------------- shared-minimum form
//
// 100 * sim(a, b) = sum over child edges e of min(y_a(e), y_b(e)), with
// y_s(e) = beta_e * x_s(e) >= 0 (see kern_sharedmin.hip). y is stored as
// fixed point, Y[row][sample] = round(y * 2^K) (uint32), rows = active
// entries in schedule order, samples padded to a multiple of the block size.

constexpr int kSmBlock = 32;   // samples per block (tile = 32 x 32 pairs per wave)
// Fixed-point ceiling of the shared-minimum embedding: every stored value and
// column sum stays below 2^31, so the tiles' |a - b| form is exact in uint32.
constexpr double kSmFixedMax = 2147483647.0;
// VALU lane-ops per pair-entry of the shared-minimum tiles: one v_sad_u32,
// plus one column-sum add per staged row and lane (16 pair-entries).
constexpr double kSmOpsPerPairEntry = 17.0 / 16.0;
constexpr int kSmWaves = 4;    // tiles per workgroup
constexpr int kSmChunk = 16;   // rows staged per LDS chunk

struct SmData {
    const uint32_t *Y;      // nrows x npad
    const uint32_t *bm;     // nblk x nwords bitmaps: bit r of block b = any Y[r][b*32..] > 0
    int64_t nrows;
    int64_t npad;
    int64_t nwords;
    int wide_rows;          // split tiles: 64-bit row addresses whatever the size (cms_opts.wide_rows)
    int n;                  // samples
    double inv_scale;       // 1 / (100 * 2^K)
    unsigned long long *counter;  // optional: entries visited (summed over tiles)
    int unweighted;               // epilogue I / (W_a + W_b - I) instead of I / (100 2^K)
    const unsigned long long *colsum;  // per-sample sum of Y (unweighted)
    const int32_t *perm;          // stored position -> caller's sample id (null = identity)
    const int32_t *pos;           // caller's sample id -> stored position (null = identity)
    // Light entries: bit r of lbm[b * nwords + r / 32] is set when block b
    // has 1 to kSmRecVals samples with mass at entry r; crl[croff[b] + k] =
    // {mask, v0, v1, v2} for block b's k-th light entry (bit j of mask:
    // Y[r][32 b + j] > 0, v = those values in ascending sample order, zeros
    // after the last). The blocks' records are packed: croff[b] = light
    // entries of the blocks before b.
    const uint32_t *lbm;          // null = none (the split tiles need them)
    const uint4 *crl;
    const int64_t *croff;         // nblocks + 1 offsets into crl
    const uint32_t *pl;           // pl[b * nwords + w] = light entries of block b before word w
    // An all-zero row of Y (index >= nrows; -1 = none): the split tiles
    // stage it in the slots of a short last chunk, so every chunk runs the
    // same unrolled accumulation.
    int64_t zrow;
};
constexpr int kSmRecVals = 3;
// Counters of a shared-minimum all-pairs run (SmData.counter[k]):
constexpr int kCntVisited = 0;    // entries set in both blocks' bitmaps, summed over tiles
constexpr int kCntStaged = 1;     // of those, entries accumulated through the staged |a - b| path
constexpr int kCntSparseRows = 2; // entries taken by the sparse path
constexpr int kCntSparsePairs = 3;  // pairs with mass on both sides in those entries
constexpr int kCntHalfRows = 4;     // entries light in one block only (half path)
constexpr int kCntHalfPairs = 5;    // pairs with mass on both sides in those entries
constexpr int kCounters = 8;
// Tile counters of the persistent shared-minimum kernels: one per XCD, 64
// bytes apart (kWorkWords words zeroed before each launch).
constexpr int kXcds = 8;
constexpr int kWorkStride = 8;
constexpr int kWorkWords = kXcds * kWorkStride;
constexpr unsigned kXcdRun = 256;  // consecutive tiles dealt to one XCD

// Persistent all-pairs launch over upper-triangle block tiles
// [t_begin, t_begin + t_count) of an nb x nb block grid. variant selects the
// LDS chunk / waves-per-workgroup instance (0..5); out_mode 0 = packed
// triangle in the caller's order, 1 = packed triangle of stored positions
// minus out_base. `work` is a zeroed device counter. out holds out_cap
// floats; a launch that could write outside them is refused on the host
// (hipErrorInvalidValue) before anything is queued.
// Variants: CHUNK x WAVES of the plain tiles (1..6) and of the split tiles
// (7..12), whose entries light in both blocks take the sparse path, one wave
// per tile (1,024 pairs per wave); and split tiles shared by PARTS waves of
// a workgroup (13..16: 512 or 256 pairs per wave), for sample sets with
// fewer tiles than the device has wave slots.
constexpr int kSmVariants = 16;
struct SmVariant {
    int chunk, waves, split, parts;  // parts: waves per tile
};
constexpr SmVariant kSmVariantTable[kSmVariants] = {
    {8, 4, 0, 1},  {16, 4, 0, 1}, {32, 4, 0, 1}, {16, 2, 0, 1}, {32, 2, 0, 1}, {16, 8, 0, 1},
    {8, 2, 1, 1},  {8, 1, 1, 1},  {8, 4, 1, 1},  {16, 2, 1, 1}, {16, 1, 1, 1}, {16, 4, 1, 1},
    {8, 2, 1, 2},  {8, 4, 1, 4},  {16, 2, 1, 2}, {16, 4, 1, 4}};
// Light-entry prefix of every block (PL[b][w] = light entries of block b
// before word w) and the packed record offsets (off[b] = light entries of
// blocks 0 .. b - 1; off[nblk] = all of them), from the light bitmaps.
hipError_t launch_light_prefix(const uint32_t *lbm, int64_t nblk, int64_t nwords, uint32_t *PL, int64_t *off,
                               hipStream_t st);
// The records of every block's light entries in rank order at crl + off[b],
// from Y and the light bitmaps (after launch_light_prefix); crl holds cap
// records, and records past them are not written.
hipError_t launch_light_records(const uint32_t *Y, int64_t npad, const uint32_t *lbm, int64_t nblk, int64_t nwords,
                                const uint32_t *PL, const int64_t *off, uint4 *crl, int64_t cap, hipStream_t st);
// The same for the listed blocks of a matrix of row stride `stride` (a
// multiple of 32), the columns whose bit of masks[k] is clear reading as
// zero (the serving query matrix: inactive slots).
hipError_t launch_light_records_listed(const uint32_t *Y, int64_t stride, const int32_t *blocks,
                                       const uint32_t *masks, int nblocks, const uint32_t *lbm, int64_t nwords,
                                       uint32_t *PL, uint4 *crl, int64_t crs, hipStream_t st);
constexpr int kSmCrossVariants = 8;   // the serving cross kernel: plain tiles 1..6, split tiles 7..8
constexpr int kSmDefaultVariant = 7;  // without autotune: split tiles, 8-entry chunk, 1 wave per workgroup
constexpr int kSmSharedTileVariant = 13;       // without autotune, below kSmSharedTileMaxTiles tiles: split
constexpr int64_t kSmSharedTileMaxTiles = 4096;  // tiles shared by 4 waves, 8-entry chunk
hipError_t launch_sm_persistent(const SmData &d, int variant, int64_t nb, int64_t t_begin, int64_t t_count,
                                unsigned long long *work, float *out, int64_t out_cap, int out_mode,
                                int64_t out_base, int grid_waves, hipStream_t st);

hipError_t launch_zero_counter(unsigned long long *counter, hipStream_t st);
hipError_t launch_zero_counters(unsigned long long *counter, int n, hipStream_t st);
hipError_t launch_calib_copy(const void *src, void *dst, int64_t bytes, hipStream_t st);

// Per-sample sum of leaf values over consumed leaf positions (FP64).
hipError_t launch_sample_mass(const int64_t *rowptr, const int32_t *col, const float *val, int n,
                              const int32_t *leaf_uses, double *mass, hipStream_t st);
// Embedding y per (row, sample) into Y (fixed point) and the block bitmaps.
//   ab: F32 walk layout (block-leaf-major) of the same samples
//   ent_row[e]: row of entry e in Y or -1 (inactive)
// One schedule node as the embedding kernel reads it (48 bytes): the walk
// node plus the Y rows and beta of its two child edges (row -1 = inactive).
struct EmbedNode {
    uint32_t code;
    int32_t leafpos;
    float w1, w2;
    int32_t row1, row2;
    int32_t pad0, pad1;
    double beta1, beta2;
};
static_assert(sizeof(EmbedNode) == 48, "EmbedNode must be 48 bytes");

// The samples' nonzero leaves in leaf-stream order, per stored sample p
// (built by launch_leaf_lists from the sample CSR): entries pos[i] / val[i]
// for i in [start, end[p]) with ascending positions; start[p * nseg + k] =
// the first entry at or after segment k's first leaf position; top[t *
// top_stride + p] = the value of top leaf t (the t-th leaf the nodes above
// the segments read, in schedule order), zero where the sample lacks it.
struct LeafLists {
    const int32_t *pos;
    const float *val;
    const int64_t *start;
    const int64_t *end;
    const float *top;
    int64_t top_stride;
};
// Largest leaf stream the list builder handles (two LDS words per 32
// positions in one workgroup, 62 KB); longer streams use the F32 walk layout.
constexpr int64_t kLeafListMaxPositions = 7936 * 32;
// Build the lists of stored samples [0, n) (perm: stored -> caller sample, or
// null); topidx[q] = top leaf index of stream position q, or -1; seg_lp[k] =
// segment k's first leaf position. ll.top must be zeroed by the caller.
hipError_t launch_leaf_lists(const int64_t *rowptr, const int32_t *col, const float *val, const int32_t *perm, int n,
                             const int32_t *leaf_lp_ptr, const int32_t *leaf_lp, int64_t nlp, const int32_t *seg_lp,
                             int nseg, const int32_t *topidx, int32_t *pos, float *lval, int64_t *start,
                             int64_t *end, float *top, int64_t top_stride, hipStream_t st);

struct EmbedData {
    const EmbedNode *nodes;  // pad0 = segment whose output this node is (-1: none)
    int order_n;
    int nslots;
    const int32_t *seg_first, *seg_last;  // nseg subtree segments
    int nseg;
    // records of the ntop nodes above the segments, in schedule order, each
    // with pad0 / pad1 for child 1 / 2: a leaf child's top leaf index, else
    // the segment whose output it is (-1: a register)
    const EmbedNode *top;
    int ntop;
    const float *ab;  // F32 walk layout (when sparse is 0)
    // read the leaf values from ll instead of ab; then the subtree segments
    // also store a row's entries of a 32-sample block only where the block
    // has mass in that row (its bitmap bit): the tiles, light records and
    // pair lists read no others, and the rectangle tiles zero what a lane's
    // own block lacks. With ab every entry is stored (the serving loop packs
    // whole query columns).
    int sparse;
    LeafLists ll;     // offset to the first sample embedded
    int64_t nlp;
    int n;
    int64_t npad;     // samples rounded up to 64 (one wave each); xroot stride
    int64_t nwords;
    int64_t ystride;  // row stride of Y (0 = npad)
    double scale;     // 2^K
    uint32_t *Y;
    uint32_t *bm;     // zeroed by the caller; bits are ORed in
    uint32_t *lbm;    // optional light-entry bitmaps (SmData.lbm), zeroed by the caller; bits are ORed in
    double *xroot;    // nseg x npad
    int unweighted;   // presence propagation, coefficient w (EmbedNode.beta* hold w)
    int sums_only;    // first pass: FP64 column sums of y into colsum_f64, nothing else
    unsigned long long *colsum;  // per-sample sum of the stored fixed-point y (atomic)
    double *colsum_f64;          // per-sample FP64 sum of y (sums_only pass)
};
hipError_t launch_embed(const EmbedData &e, hipStream_t st);
// All pairs, packed upper triangle.
// Rectangle over sample ranges aligned to nothing: rows [r0, r0+nr) x cols [c0, c0+nc).
hipError_t launch_sm_rect(const SmData &d, int r0, int nr, int c0, int nc, float *out, hipStream_t st);
// Explicit pair list.
hipError_t launch_sm_pairs(const SmData &d, int64_t n_pairs, const int32_t *row, const int32_t *col, float *out,
                           hipStream_t st);

// --------------------------------------------------- leaf metrics (-M 2..5)

// Normalised leaf abundances p = a / sum(a) as float, P[leaf][sample]
// (leaf_n x npad), the per-32-sample-block bitmaps of leaves with mass and
// per-sample totals stats[3 s + {0, 1, 2}] = sum a, sum p, sum p^2 (FP64).
struct LeafData {
    const float *P;
    const uint32_t *bm;
    const double *stats;
    int64_t npad, nwords;  // nwords = ceil(leaf_n / 32)
    int n, leaf_n;
};
// Fill P (zeroed by the caller), bm (zeroed) and stats from CSR samples.
hipError_t launch_leaf_prepare(const int64_t *rowptr, const int32_t *col, const float *val, int n, const LeafData &d,
                               hipStream_t st);
// All pairs of metric `mode` (CMS_METRIC_COS..BC), packed upper triangle of
// out_cap >= n (n - 1) / 2 floats; `work` is a zeroed device counter.
hipError_t launch_leaf_all_pairs(const LeafData &d, int mode, unsigned long long *work, float *out, int64_t out_cap,
                                 int grid_waves, hipStream_t st);

// The same metrics in the reference's FP32 operation order: per pair, a merge
// of the two samples' entry lists.
struct LeafMergeData {
    const int64_t *rowptr;  // the caller's samples, entries in ascending leaf order
    const uint2 *ent;       // per entry: leaf, bits of a (cosine) or of a / sum a (the others)
    const float2 *stat;     // per sample: FP32 sum of a, FP32 sum of a^2, both in leaf order
    const int32_t *perm;    // visiting position -> sample (most entries first)
    int n;
    // JSD: the host C library's logf exceptions (host_logf_table), on device
    const uint32_t *logf_bm;
    const uint32_t *logf_blocks;
    const unsigned long long *logf_tab;
    int logf_ntab;
    float logf_ln2;              // the host's logf(2.0f)
    const double *logf_pieces;   // table_log's invc[64], then logc[64]
};
// Fill ent (one per CSR entry) and stat (one per sample).
hipError_t launch_leaf_merge_prepare(const int64_t *rowptr, const int32_t *col, const float *val, int n,
                                     bool normalise, uint2 *ent, float2 *stat, hipStream_t st);
// All pairs of metric `mode`, packed upper triangle of out_cap >= n (n - 1) / 2.
hipError_t launch_leaf_merge_all_pairs(const LeafMergeData &d, int mode, float *out, int64_t out_cap,
                                       hipStream_t st);

// ------------------------------------------------------------------ serving

// Query slots (columns of QY, same rows as the references' Y) against the
// resident references.
struct CrossData {
    const uint32_t *QY;                 // nrows x qstride, slot = column
    int64_t qstride;                    // slots (multiple of 32)
    const uint32_t *qbm;                // (qstride / 32) x nwords query-block bitmaps
    const unsigned long long *qcolsum;  // per slot (unweighted)
    const uint32_t *Y;                  // references, nrows x npad
    const uint32_t *bm;                 // reference-block bitmaps
    const unsigned long long *colsum;   // per stored reference (unweighted)
    const int32_t *perm;                // stored reference -> caller's reference id (null = identity)
    int64_t npad;
    int64_t nwords;
    int n_ref;
    int unweighted;
    double inv_scale;                   // 1 / (100 * 2^K)
    unsigned long long *counter;        // optional: counters as SmData.counter (kCnt*)
    // split tiles (cross variants 7, 8): light bitmaps, prefixes and records
    // of both sides (the references' as SmData: lbm / pl / crl, croff; the
    // query blocks' records at qcrl + qb * qcrs), and a row index that is all
    // zero in QY and Y
    const uint32_t *qlbm, *qpl;
    const uint4 *qcrl;
    int64_t qcrs;
    const uint32_t *lbm, *pl;
    const uint4 *crl;
    const int64_t *croff;
    int64_t zrow;
};

// Step: tiles (active query block qblocks[k]) x (reference blocks
// [rb0, rb0 + nrb)); slot_out[slot] = the output row of the query in that slot
// (indexed by caller's reference id), null for an inactive slot. `work` is a
// zeroed device counter. Variants 0..5 as launch_sm_persistent's plain tiles;
// 6 and 7 split tiles (8-row chunk, 1 or 2 waves per workgroup), which need
// the light structures of CrossData. When the step has too few tiles to fill
// the device, tile rows are split over several waves whose sums meet in
// part_acc (part_tiles x 1024 zeroed uint32, left zeroed) and a second kernel
// writes the similarities.
hipError_t launch_sm_cross(const CrossData &x, int variant, const int32_t *qblocks, int nqa, int64_t rb0, int64_t nrb,
                           float *const *slot_out, uint32_t *part_acc, int64_t part_tiles, unsigned long long *work,
                           int grid_waves, hipStream_t st);
// Bitmaps of the listed query blocks from their active slots (masks[k] bit j
// = slot 32 blocks[k] + j is active), and their light bitmaps if qlbm is given.
hipError_t launch_qbm_build(const uint32_t *QY, int64_t qstride, int64_t nrows, int64_t nwords, const int32_t *blocks,
                            const uint32_t *masks, int nblocks, uint32_t *qbm, uint32_t *qlbm, hipStream_t st);
// QY[:, slots[q]] = Ys[:, q] and qcs[slots[q]] = cs[q] for q < nq.
hipError_t launch_pack(const uint32_t *Ys, int64_t sstride, int nq, const int32_t *slots, int64_t nrows, uint32_t *QY,
                       int64_t qstride, const unsigned long long *cs, unsigned long long *qcs, hipStream_t st);

}  // namespace cms
