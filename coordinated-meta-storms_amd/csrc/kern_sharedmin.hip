// Shared-minimum kernels: the same similarity as the walk of
// _Comp_Tree::Calc_sim (comp.h:332-375), from a per-sample embedding.
//
// Identity. Let node p of the schedule have children k (leaves or earlier
// nodes) with weights w_k = 1 - dist_k (MIN_DIST-clamped as in comp.h:343-344).
// Write x_s(k) for sample s's abundance under k decayed to k
// (x_s(leaf) = a_s(leaf), x_s(p) = sum_k w_k x_s(k)). The walk's leftovers
// satisfy RA(k) = x_a(k) - E_k and RB(k) = x_b(k) - E_k with the same E_k, so
// each shared minimum is min(x_a(k), x_b(k)) - E_k, and E_p = sum_k w_k
// min(x_a(k), x_b(k)). Summing over the schedule,
//
//     100 * sim = sum over child edges e of beta_e * min(x_a(e), x_b(e)),
//     beta_e = 1 - w_e if the parent's value is read later, 1 for a root.
//
// With beta >= 0 (no negative branch lengths) beta * min(x, y) =
// min(beta x, beta y), so the pair term is a plain histogram intersection of
// y_s(e) = beta_e x_s(e). Only entries where BOTH samples have mass (the
// intersection of their root paths) contribute, which is what the tiles below
// exploit: a tile of 32 x 32 samples visits only the entries set in both
// blocks' bitmaps.
//
// Arithmetic. y is computed in FP64 and stored as fixed point
// round(y * 2^K) in uint32 with K chosen on the host so that every sample's
// column sum stays below 2^31 (the column sum equals the sample's total leaf
// mass, about 100). Pair sums are then exact integer arithmetic (the tiles
// sum |a - b| with v_sad_u32 and recover the minima from the column sums:
// 1 + 1/16 VALU ops per pair-entry), order-independent and deterministic;
// the only error is the rounding of y (<= 2^-(K+1) per entry).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "sm_tile.h"

namespace cms {

typedef const __attribute__((address_space(4))) WalkNode *cnode_ptr;
typedef const __attribute__((address_space(4))) double *cdouble_ptr;
typedef const __attribute__((address_space(4))) int32_t *cint_ptr;

__device__ __forceinline__ void sm_decode_upper(int64_t t, int64_t nb, int64_t &rb, int64_t &cb) {
    double nn = (double)nb;
    double disc = (2.0 * nn + 1.0) * (2.0 * nn + 1.0) - 8.0 * (double)t;
    int64_t r = (int64_t)(((2.0 * nn + 1.0) - sqrt(disc)) * 0.5);
    if (r < 0) r = 0;
    while (r > 0 && r * nb - r * (r - 1) / 2 > t) r--;
    while ((r + 1) * nb - (r + 1) * r / 2 <= t) r++;
    rb = r;
    cb = r + (t - (r * nb - r * (r - 1) / 2));
}

// Block row ra and column cb of tile t in the order of a launch over whole
// block rows g (grouped_tile; row-major when g.r0 < 0).
__device__ __forceinline__ void sm_tile_at(int64_t t, int64_t nb, const TileRows &g, int64_t &ra, int64_t &cb) {
    int64_t rb, c;
    sm_decode_upper(t, nb, rb, c);
    grouped_tile(t, nb, g, rb, ra, cb);
}

__device__ __forceinline__ int64_t sm_xcd_remap(int64_t b, int64_t nwg) {
    const int64_t X = 8;
    int64_t xcd = b % X, idx = b / X, per = nwg / X, rem = nwg % X;
    return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// ------------------------------------------------------------ embedding

__global__ void sample_mass_kernel(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                   const float *__restrict__ val, int n, const int32_t *__restrict__ leaf_uses,
                                   double *__restrict__ mass) {
    const int s = blockIdx.x * blockDim.y + threadIdx.y;
    if (s >= n) return;
    double m = 0.0;
    for (int64_t k = rowptr[s] + threadIdx.x; k < rowptr[s + 1]; k += 64) m += (double)val[k] * leaf_uses[col[k]];
    for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off, 64);
    if (threadIdx.x == 0) mass[s] = m;
}

// Leaf lists (LeafLists): one workgroup per stored sample sorts its nonzero
// leaves by leaf-stream position with a counting sort in LDS: a bitmap of
// its positions, the exclusive popcount prefix over the bitmap words (rank
// of each position), then each entry written at its rank. The same ranks
// give the segments' first entries, and the top leaves' values go to their
// dense rows.
constexpr int kLeafListThreads = 256;
__global__ __launch_bounds__(kLeafListThreads) void leaf_lists_kernel(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col, const float *__restrict__ val,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ leaf_lp_ptr, const int32_t *__restrict__ leaf_lp,
    int64_t nlp, const int32_t *__restrict__ seg_lp, int nseg, const int32_t *__restrict__ topidx,
    int32_t *__restrict__ pos, float *__restrict__ lval, int64_t *__restrict__ start, int64_t *__restrict__ end,
    float *__restrict__ top, int64_t top_stride) {
    extern __shared__ uint32_t lsh[];  // bits[nwp], then pre[nwp]
    __shared__ uint32_t part[kLeafListThreads];
    const int64_t p = blockIdx.x;
    const int tid = threadIdx.x;
    const int nwp = (int)((nlp + 31) / 32);
    uint32_t *bits = lsh, *pre = lsh + nwp;
    const int64_t s = perm ? perm[p] : p;
    const int64_t k0 = rowptr[s], k1 = rowptr[s + 1];
    for (int w = tid; w < nwp; w += kLeafListThreads) bits[w] = 0u;
    __syncthreads();
    for (int64_t k = k0 + tid; k < k1; k += kLeafListThreads) {
        const int32_t leaf = col[k];
        for (int32_t t = leaf_lp_ptr[leaf]; t < leaf_lp_ptr[leaf + 1]; t++) {
            const int32_t q = leaf_lp[t];
            atomicOr(bits + (q >> 5), 1u << (q & 31));
        }
    }
    __syncthreads();
    // exclusive prefix of the words' popcounts: each thread a contiguous run
    const int per = (nwp + kLeafListThreads - 1) / kLeafListThreads;
    const int w0 = min(tid * per, nwp), w1 = min(w0 + per, nwp);
    uint32_t sum = 0;
    for (int w = w0; w < w1; w++) sum += __popc(bits[w]);
    part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < kLeafListThreads; off <<= 1) {
        const uint32_t add = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    uint32_t run = part[tid] - sum;
    for (int w = w0; w < w1; w++) {
        pre[w] = run;
        run += __popc(bits[w]);
    }
    __syncthreads();
    const uint32_t total = part[kLeafListThreads - 1];
    auto rank = [&](int32_t q) -> uint32_t {
        return pre[q >> 5] + __popc(bits[q >> 5] & ((1u << (q & 31)) - 1u));
    };
    for (int64_t k = k0 + tid; k < k1; k += kLeafListThreads) {
        const int32_t leaf = col[k];
        for (int32_t t = leaf_lp_ptr[leaf]; t < leaf_lp_ptr[leaf + 1]; t++) {
            const int32_t q = leaf_lp[t];
            const int64_t i = k0 + rank(q);
            pos[i] = q;
            lval[i] = val[k];
            const int32_t ti = topidx[q];
            if (ti >= 0) top[(int64_t)ti * top_stride + p] = val[k];
        }
    }
    for (int k = tid; k < nseg; k += kLeafListThreads) {
        const int32_t q = seg_lp[k];
        start[p * nseg + k] = k0 + (q < nlp ? rank(q) : total);
    }
    if (tid == 0) end[p] = k0 + total;
}

hipError_t launch_leaf_lists(const int64_t *rowptr, const int32_t *col, const float *val, const int32_t *perm, int n,
                             const int32_t *leaf_lp_ptr, const int32_t *leaf_lp, int64_t nlp, const int32_t *seg_lp,
                             int nseg, const int32_t *topidx, int32_t *pos, float *lval, int64_t *start,
                             int64_t *end, float *top, int64_t top_stride, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (nlp < 1 || nlp > kLeafListMaxPositions) return hipErrorInvalidValue;
    const size_t lds = 2 * sizeof(uint32_t) * (size_t)((nlp + 31) / 32);
    leaf_lists_kernel<<<n, kLeafListThreads, lds, st>>>(rowptr, col, val, perm, leaf_lp_ptr, leaf_lp, nlp, seg_lp,
                                                         nseg, topidx, pos, lval, start, end, top, top_stride);
    return hipGetLastError();
}

// Per-sample embedding. Every wave handles 64 consecutive samples
// (lane = sample) and walks part of the schedule in FP64:
//   embed_seg_kernel  one subtree segment per wave (grid.y = segment): its
//                     contiguous node range, staged in chunks of 64 node
//                     records plus the leaf values they consume;
//   embed_top_kernel  the nodes above the segments, reading segment outputs
//                     from xroot[segment][sample].
// Each node writes Y[row][s] for its active child edges and ORs the
// per-32-sample-block bits of those rows into the bitmaps.
constexpr int kEmbedNodeChunk = 8;  // node records (and 2x leaf positions) staged in LDS per step
constexpr int kEmbedLeafChunk = 2 * kEmbedNodeChunk;
static_assert(kEmbedLeafChunk % 8 == 0, "leaf values are staged in passes of 8");

struct BitmapWriter {
    uint32_t bits0 = 0, bits1 = 0;    // entries with mass in the wave's first / second block
    uint32_t light0 = 0, light1 = 0;  // of those, entries where the block has at most kSmRecVals samples with mass
    int64_t cur = -1;
    __device__ __forceinline__ void flush(uint32_t *bm, uint32_t *lbm, int64_t blk0, int64_t nwords, int lane) {
        if (cur >= 0 && lane == 0) {
            if (bits0) atomicOr(bm + blk0 * nwords + cur, bits0);
            if (bits1) atomicOr(bm + (blk0 + 1) * nwords + cur, bits1);
            if (lbm && light0) atomicOr(lbm + blk0 * nwords + cur, light0);
            if (lbm && light1) atomicOr(lbm + (blk0 + 1) * nwords + cur, light1);
        }
        bits0 = bits1 = light0 = light1 = 0;
    }
    __device__ __forceinline__ void add(int32_t row, uint32_t q, uint32_t *bm, uint32_t *lbm, int64_t blk0,
                                        int64_t nwords, int lane) {
        const uint64_t m = __ballot(q != 0u);
        const int64_t w = row >> 5;
        if (w != cur) {
            flush(bm, lbm, blk0, nwords, lane);
            cur = w;
        }
        const uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32), bit = 1u << (row & 31);
        if (lo) bits0 |= bit;
        if (hi) bits1 |= bit;
        if (lo && __popc(lo) <= kSmRecVals) light0 |= bit;
        if (hi && __popc(hi) <= kSmRecVals) light1 |= bit;
    }
};

// The same bitmaps with the per-entry work in vector registers: each lane
// keeps the bits of the current word where its own sample has mass (two
// vector ops per entry, no scalar ones), and at a word change the 32 lanes
// of each block sum their bits into per-bit counters saturating at 4 (bit
// planes c0, c1 and c2 = "4 or more", five shuffle-and-add steps), which
// give the block's word (count > 0) and its light word (count 1 to 3). Used
// where entries come in runs of consecutive rows (a subtree segment): there
// the scalar writer's ballot, popcounts and selects per entry kept the
// scalar unit busier than all the vector work (3.5x the scalar
// instructions of this form).
static_assert(kSmRecVals == 3, "the light counters saturate at 4");
// Value of lane (lane ^ X) within each 32-lane half: ds_swizzle in bit-mask
// mode (and-mask 0x1f, xor-mask X), no address register.
template <int X>
__device__ __forceinline__ uint32_t swizzle_xor(uint32_t v) {
    static_assert(X > 0 && X < 32, "xor within 32 lanes");
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (X << 10) | 0x1f);
}
// Counters of lanes lane and lane ^ X added (saturating at 4).
template <int X>
__device__ __forceinline__ void count_step(uint32_t &c0, uint32_t &c1, uint32_t &c2) {
    const uint32_t b0 = swizzle_xor<X>(c0), b1 = swizzle_xor<X>(c1), b2 = swizzle_xor<X>(c2);
    const uint32_t k0 = c0 & b0, x1 = c1 ^ b1;
    c2 = c2 | b2 | (c1 & b1) | (k0 & x1);
    c1 = x1 ^ k0;
    c0 ^= b0;
}
struct LaneBitmapWriter {
    uint32_t m = 0;    // this lane's sample has mass at entry cur * 32 + bit
    int64_t cur = -1;  // current bitmap word (wave-uniform)
    __device__ __forceinline__ void flush(uint32_t *bm, uint32_t *lbm, int64_t blk0, int64_t nwords, int lane) {
        if (cur < 0) return;
        uint32_t c0 = m, c1 = 0, c2 = 0;
        count_step<1>(c0, c1, c2);
        count_step<2>(c0, c1, c2);
        count_step<4>(c0, c1, c2);
        count_step<8>(c0, c1, c2);
        count_step<16>(c0, c1, c2);
        if ((lane & 31) == 0) {
            const int64_t w = (blk0 + (lane >> 5)) * nwords + cur;
            const uint32_t any = c0 | c1 | c2, light = (c0 | c1) & ~c2;
            if (any) atomicOr(bm + w, any);
            if (lbm && light) atomicOr(lbm + w, light);
        }
        m = 0;
    }
    __device__ __forceinline__ void add(int32_t row, uint32_t q, uint32_t *bm, uint32_t *lbm, int64_t blk0,
                                        int64_t nwords, int lane) {
        const int64_t w = row >> 5;
        if (w != cur) {
            flush(bm, lbm, blk0, nwords, lane);
            cur = w;
        }
        m |= min(q, 1u) << (row & 31);
    }
};

// UNW: unweighted form (comp.h:378-446): x is presence (leaf a > 0, parent =
// either child present) and the edge coefficient is w, so the pair sum is
// the weight shared by both samples. SUMS: only accumulate each sample's
// FP64 column sum of y (first pass, to choose the fixed-point scale).
// SPARSE: the leaf values come from the leaf lists (each lane walks its
// sample's list with a cursor, one entry prefetched), not the F32 walk
// layout: the layout's 4 bytes per sample and leaf, almost all zeros, are
// neither written nor read; and a row's entries of a 32-sample block are
// stored only where the block has mass (EmbedData.sparse).
template <bool UNW, bool SUMS, bool SPARSE>
__global__ __launch_bounds__(64) void embed_seg_kernel(const EmbedNode *__restrict__ nodes,
                                                       const int32_t *__restrict__ seg_first,
                                                       const int32_t *__restrict__ seg_last, int order_n,
                                                       const float *__restrict__ ab, LeafLists ll, int nseg,
                                                       int64_t nlp, int n, int64_t npad,
                                                       int64_t nwords, int64_t ystride, double scale,
                                                       uint32_t *__restrict__ Y,
                                                       uint32_t *__restrict__ bm, uint32_t *__restrict__ lbm, double *__restrict__ xroot,
                                                       unsigned long long *__restrict__ colsum,
                                                       double *__restrict__ colsum_f64) {
    extern __shared__ double eslots[];  // nslots * 64
    __shared__ float stage[kEmbedLeafChunk * 64];
    __shared__ EmbedNode snodes[kEmbedNodeChunk];
    const int lane = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.x * 64 + lane;
    const bool valid = s < n;
    const float *abs_ = ab + (valid ? (s >> 3) * nlp * 8 + (s & 7) : 0);
    const int64_t blk0 = (int64_t)blockIdx.x * 2;  // 32-sample blocks of this wave
    const int first = seg_first[blockIdx.y], last = seg_last[blockIdx.y];
    LaneBitmapWriter bw;
    unsigned long long qsum = 0;
    double ysum = 0.0;
    // SPARSE: the lane's list entry c (position q0, value v0) and the next
    // one (q1, v1); a position past the list is INT32_MAX
    int64_t c = 0, cend = 0;
    int32_t q0 = INT32_MAX, q1 = INT32_MAX;
    float v0 = 0.0f, v1 = 0.0f;
    if constexpr (SPARSE) {
        if (valid) {
            c = ll.start[s * nseg + blockIdx.y];
            cend = ll.end[s];
        }
        if (c < cend) {
            q0 = ll.pos[c];
            v0 = ll.val[c];
        }
        if (c + 1 < cend) {
            q1 = ll.pos[c + 1];
            v1 = ll.val[c + 1];
        }
    }

    for (int base = first; base <= last; base += kEmbedNodeChunk) {
        const int cnt = min(kEmbedNodeChunk, last + 1 - base);
        __builtin_amdgcn_wave_barrier();
        if (lane < cnt) snodes[lane] = nodes[base + lane];
        const int lb = nodes[base].leafpos;
        const int le = (base + cnt < order_n) ? nodes[base + cnt].leafpos : (int)nlp;
        const int nl = le - lb;  // <= 2 * cnt
        if constexpr (SPARSE) {
            // zeros, then the lane's entries in [lb, le) (positions below lb
            // were consumed by earlier chunks)
#pragma unroll
            for (int k = 0; k < kEmbedLeafChunk; k++) stage[k * 64 + lane] = 0.0f;
            while (q0 < le) {
                stage[(q0 - lb) * 64 + lane] = v0;
                q0 = q1;
                v0 = v1;
                c++;
                q1 = INT32_MAX;
                if (c + 1 < cend) {
                    q1 = ll.pos[c + 1];
                    v1 = ll.val[c + 1];
                }
            }
        }
        // leaf values in passes of 8 without per-value branches: a pass reads
        // 8 consecutive positions (the next chunk's leaves past this chunk's
        // last, into stage slots no node reads), clamped to the last
        // position only at the end of the leaf stream
        for (int j0 = 0; j0 < (SPARSE ? 0 : nl); j0 += 8) {
            const float *pa = abs_ + (int64_t)(lb + j0) * 8;
            float v[8];
            if (lb + j0 + 8 <= nlp) {
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = pa[k * 8];
            } else {
                const int lim = (int)(nlp - 1 - lb - j0);
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = pa[min(k, lim) * 8];
            }
#pragma unroll
            for (int k = 0; k < 8; k++) stage[(j0 + k) * 64 + lane] = valid ? v[k] : 0.0f;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int k = 0; k < cnt; k++) {
            const uint32_t code = __builtin_amdgcn_readfirstlane(snodes[k].code);
            int lp = __builtin_amdgcn_readfirstlane(snodes[k].leafpos) - lb;
            const int32_t rows[2] = {__builtin_amdgcn_readfirstlane(snodes[k].row1),
                                     __builtin_amdgcn_readfirstlane(snodes[k].row2)};
            const int32_t exp_seg = __builtin_amdgcn_readfirstlane(snodes[k].pad0);
            double x[2];
            if (code & kC1Leaf) {
                x[0] = (double)stage[lp * 64 + lane];
                if (UNW) x[0] = x[0] > 0.0 ? 1.0 : 0.0;
                lp++;
            } else {
                x[0] = eslots[(code & 0xffu) * 64 + lane];
            }
            if (code & kC2Leaf) {
                x[1] = (double)stage[lp * 64 + lane];
                if (UNW) x[1] = x[1] > 0.0 ? 1.0 : 0.0;
            } else {
                x[1] = eslots[((code >> 8) & 0xffu) * 64 + lane];
            }
            const double betas[2] = {snodes[k].beta1, snodes[k].beta2};
#pragma unroll
            for (int j = 0; j < 2; j++) {
                if (rows[j] < 0) continue;
                if (SUMS) {
                    ysum += betas[j] * x[j];
                    continue;
                }
                const uint32_t q = (uint32_t)rint(betas[j] * x[j] * scale);
                if constexpr (SPARSE) {
                    // the lanes of a block with mass in this row
                    const uint64_t has = __ballot(q != 0u);
                    if ((lane < 32 ? (uint32_t)has : (uint32_t)(has >> 32)) != 0u)
                        Y[(int64_t)rows[j] * ystride + s] = q;
                } else {
                    Y[(int64_t)rows[j] * ystride + s] = q;
                }
                qsum += q;
                bw.add(rows[j], q, bm, lbm, blk0, nwords, lane);
            }
            if (code & kDstLive) {
                const double xp = UNW ? fmax(x[0], x[1]) : (double)snodes[k].w1 * x[0] + (double)snodes[k].w2 * x[1];
                if (exp_seg >= 0)
                    xroot[(int64_t)exp_seg * npad + s] = xp;  // consumed by a top node
                else
                    eslots[((code >> 16) & 0xffu) * 64 + lane] = xp;
            }
        }
    }
    if (SUMS) {
        if (valid) atomicAdd(colsum_f64 + s, ysum);
        return;
    }
    bw.flush(bm, lbm, blk0, nwords, lane);
    if (valid && qsum) atomicAdd(colsum + s, qsum);
}

template <bool UNW, bool SUMS, bool SPARSE>
__global__ __launch_bounds__(64) void embed_top_kernel(const EmbedNode *__restrict__ top, int ntop,
                                                       const float *__restrict__ ab, LeafLists ll, int64_t nlp,
                                                       int n, int64_t npad,
                                                       int64_t nwords, int64_t ystride, double scale,
                                                       uint32_t *__restrict__ Y,
                                                       uint32_t *__restrict__ bm, uint32_t *__restrict__ lbm, const double *__restrict__ xroot,
                                                       unsigned long long *__restrict__ colsum,
                                                       double *__restrict__ colsum_f64) {
    extern __shared__ double eslots[];
    __shared__ EmbedNode tnodes[kEmbedNodeChunk];
    const int lane = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.x * 64 + lane;
    const bool valid = s < n;
    const float *abs_ = ab + (valid ? (s >> 3) * nlp * 8 + (s & 7) : 0);
    const int64_t blk0 = (int64_t)blockIdx.x * 2;
    BitmapWriter bw;
    unsigned long long qsum = 0;
    double ysum = 0.0;
    // Chunks of kEmbedNodeChunk nodes: their records staged in LDS (the next
    // chunk's fetched into registers while this one is walked), then the
    // loads of every child value held in memory (leaf abundances, segment
    // outputs) issued before the first node uses one, so a chunk waits on
    // memory about once instead of three times per node.
    const uint4 *recs = reinterpret_cast<const uint4 *>(top);  // 3 per record
    uint4 r0{}, r1{}, r2{};
    if (lane < min(kEmbedNodeChunk, ntop)) {
        r0 = recs[3 * lane];
        r1 = recs[3 * lane + 1];
        r2 = recs[3 * lane + 2];
    }
    for (int t0 = 0; t0 < ntop; t0 += kEmbedNodeChunk) {
        const int cnt = min(kEmbedNodeChunk, ntop - t0);
        __builtin_amdgcn_wave_barrier();
        if (lane < cnt) {
            uint4 *q = reinterpret_cast<uint4 *>(tnodes) + 3 * lane;
            q[0] = r0;
            q[1] = r1;
            q[2] = r2;
        }
        if (lane < min(kEmbedNodeChunk, ntop - t0 - kEmbedNodeChunk)) {
            const int64_t r = 3 * (int64_t)(t0 + kEmbedNodeChunk + lane);
            r0 = recs[r];
            r1 = recs[r + 1];
            r2 = recs[r + 2];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double xin[kEmbedNodeChunk][2];
#pragma unroll
        for (int k = 0; k < kEmbedNodeChunk; k++) {
            xin[k][0] = xin[k][1] = 0.0;
            if (k >= cnt) continue;
            const uint32_t code = __builtin_amdgcn_readfirstlane(tnodes[k].code);
            int64_t lp = __builtin_amdgcn_readfirstlane(tnodes[k].leafpos);
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int32_t src = __builtin_amdgcn_readfirstlane(j == 0 ? tnodes[k].pad0 : tnodes[k].pad1);
                if (code & (j == 0 ? kC1Leaf : kC2Leaf)) {
                    // SPARSE: src is the leaf's top leaf index
                    const float a = SPARSE ? ll.top[(int64_t)src * ll.top_stride + (valid ? s : 0)] : abs_[lp * 8];
                    xin[k][j] = valid ? (double)a : 0.0;
                    lp++;
                } else if (src >= 0) {
                    xin[k][j] = xroot[(int64_t)src * npad + s];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kEmbedNodeChunk; k++) {
            if (k >= cnt) break;
            const uint32_t code = __builtin_amdgcn_readfirstlane(tnodes[k].code);
            double x[2];
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int32_t src = __builtin_amdgcn_readfirstlane(j == 0 ? tnodes[k].pad0 : tnodes[k].pad1);
                if (code & (j == 0 ? kC1Leaf : kC2Leaf)) {
                    x[j] = UNW ? (xin[k][j] > 0.0 ? 1.0 : 0.0) : xin[k][j];
                } else if (src >= 0) {
                    x[j] = xin[k][j];
                } else {
                    x[j] = eslots[((code >> (8 * j)) & 0xffu) * 64 + lane];
                }
            }
            const int32_t rows[2] = {__builtin_amdgcn_readfirstlane(tnodes[k].row1),
                                     __builtin_amdgcn_readfirstlane(tnodes[k].row2)};
            const double betas[2] = {tnodes[k].beta1, tnodes[k].beta2};
#pragma unroll
            for (int j = 0; j < 2; j++) {
                if (rows[j] < 0) continue;
                if (SUMS) {
                    ysum += betas[j] * x[j];
                    continue;
                }
                const uint32_t q = (uint32_t)rint(betas[j] * x[j] * scale);
                Y[(int64_t)rows[j] * ystride + s] = q;
                qsum += q;
                bw.add(rows[j], q, bm, lbm, blk0, nwords, lane);
            }
            if (code & kDstLive)
                eslots[((code >> 16) & 0xffu) * 64 + lane] =
                    UNW ? fmax(x[0], x[1]) : (double)tnodes[k].w1 * x[0] + (double)tnodes[k].w2 * x[1];
        }
    }
    if (SUMS) {
        if (valid) atomicAdd(colsum_f64 + s, ysum);
        return;
    }
    bw.flush(bm, lbm, blk0, nwords, lane);
    if (valid && qsum) atomicAdd(colsum + s, qsum);
}

// ------------------------------------------------------------ pair tiles

// Square tile of the all-pairs matrix: lanes 0..31 stage the row block's
// columns of Y, lanes 32..63 the column block's. CHECK when a block may
// extend past the padded sample count (rectangles at arbitrary offsets).
template <int CHUNK, bool CHECK>
__device__ __forceinline__ uint32_t sm_square_tile(const SmData &d, const uint32_t *A0, const uint32_t *A1,
                                                   const uint32_t *B0, const uint32_t *B1, int64_t abase,
                                                   int64_t bbase, uint16_t *L, uint32_t *S, int lane,
                                                   uint32_t (&acc)[4][4]) {
    const int64_t my_s = (lane < 32) ? abase + lane : bbase + (lane - 32);
    SquareRows<CHECK> rows;
    rows.Y = d.Y;
    rows.stride = d.npad;
    rows.ok = my_s < d.npad;
    rows.col4 = (uint32_t)((!CHECK || rows.ok) ? my_s : 0) * 4u;
    rows.own = d.bm + (rows.ok ? my_s / kSmBlock : 0) * d.nwords;
    return sm_tile_accumulate<CHUNK>(A0, A1, B0, B1, 0, d.nwords, rows, L, S, lane, acc);
}

// Similarity of the pair at stored positions (pa, pb) from its shared sum.
__device__ __forceinline__ float sm_finish(const SmData &d, uint32_t acc, int64_t pa, int64_t pb) {
    return d.unweighted ? sm_similarity(true, d.inv_scale, acc, d.colsum[pa], d.colsum[pb])
                        : sm_similarity(false, d.inv_scale, acc, 0ull, 0ull);
}

// Rectangle with arbitrary (unpermuted) sample offsets: rows [r0, r0+nr) x
// cols [c0, c0+nc), row-major output, diagonal = 1.
__global__ __launch_bounds__(64 * kSmWaves) void sm_rect_kernel(SmData d, int64_t ntiles, int64_t nbc, int r0, int nr,
                                                                 int c0, int nc, float *__restrict__ out) {
    __shared__ uint16_t s_list[kSmWaves][64 * 32];
    __shared__ uint32_t s_stage[kSmWaves][kSmChunk * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t tile = sm_xcd_remap(blockIdx.x, gridDim.x) * kSmWaves + wave;
    if (tile >= ntiles) return;  // whole wave leaves; no workgroup barrier is used
    const int64_t nblk = d.npad / kSmBlock;
    const int64_t abase = r0 + (tile / nbc) * kSmBlock, bbase = c0 + (tile % nbc) * kSmBlock;
    const int64_t ba = abase / kSmBlock, bb = bbase / kSmBlock;
    const uint32_t *A0 = d.bm + ba * d.nwords, *B0 = d.bm + bb * d.nwords;
    const uint32_t *A1 = ((abase % kSmBlock) && ba + 1 < nblk) ? d.bm + (ba + 1) * d.nwords : nullptr;
    const uint32_t *B1 = ((bbase % kSmBlock) && bb + 1 < nblk) ? d.bm + (bb + 1) * d.nwords : nullptr;
    uint32_t acc[4][4];
    sm_square_tile<kSmChunk, true>(d, A0, A1, B0, B1, abase, bbase, s_list[wave], s_stage[wave], lane, acc);
    const int ty = lane >> 3, tx = lane & 7;
#pragma unroll
    for (int ii = 0; ii < 4; ii++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const int64_t row = abase + ty * 4 + ii, col = bbase + tx * 4 + jj;
            if (row < r0 + nr && col < c0 + nc)
                out[(row - r0) * (int64_t)nc + (col - c0)] =
                    (row == col) ? 1.0f : sm_finish(d, acc[ii][jj], row, col);
        }
}

// The similarity of stored pair (pa, pb) of an upper-triangle block tile
// (pairs at or below the diagonal and past the last sample skipped).
//   OUT 0: the packed upper triangle in the caller's sample order (stored
//          position -> sample id through d.perm, if any);
//   OUT 1: the packed triangle of stored positions, minus out_base.
// (Per pair: a helper taking the lane's 4 x 4 sums by reference put them in
// scratch memory.)
template <int OUT>
__device__ __forceinline__ void sm_store_pair(const SmData &d, uint32_t acc, int64_t pa, int64_t pb,
                                              float *__restrict__ out, int64_t out_base) {
    if (pa >= pb || pb >= d.n) return;
    const float sim = sm_finish(d, acc, pa, pb);
    if constexpr (OUT == 0) {
        int64_t a = d.perm ? d.perm[pa] : pa, b = d.perm ? d.perm[pb] : pb;
        if (a > b) {
            const int64_t x = a;
            a = b;
            b = x;
        }
        out[a * (int64_t)d.n + b - (a + 1) * (a + 2) / 2] = sim;
    } else {
        out[pa * (int64_t)d.n + pb - (pa + 1) * (pa + 2) / 2 - out_base] = sim;
    }
}

// Next tile of the persistent launches: tiles are dealt to the XCDs in runs
// of kXcdRun consecutive tiles (run q to XCD q mod 8; workgroups are
// dispatched to the XCDs round-robin), so the waves of one XCD work on few
// blocks at a time and share those blocks' Y lines in its L2 (at 10,000
// samples, row-major runs: L2 hits 15 % -> 40 % of requests, HBM reads
// -30 %). The calling
// wave (every lane; src and left stay wave-uniform) drains its own XCD's
// counter (src, initially blockIdx.x mod 8), then each other XCD's; -1 when
// all eight are past the range (left counts the counters not yet found
// empty).
__device__ __forceinline__ int64_t sm_next_tile(unsigned long long *__restrict__ work, int64_t t_count, int &src,
                                                int &left, int lane) {
    while (left > 0) {
        unsigned long long k = 0;
        if (lane == 0) k = atomicAdd(work + src * kWorkStride, 1ull);
        k = __shfl(k, 0, 64);
        const unsigned long long t = ((k / kXcdRun) * kXcds + (unsigned)src) * kXcdRun + k % kXcdRun;
        if (t < (unsigned long long)t_count) return (int64_t)t;
        if (--left > 0) src = (src + 1) & (kXcds - 1);
    }
    return -1;
}

// All pairs, persistent: a fixed grid of waves pulls upper-triangle block
// tiles [t_begin, t_begin + t_count) from per-XCD counters (tiles in the
// order of sm_tile_at: bands of kTileGroup block rows when the launch holds
// whole rows, row-major otherwise; with samples binned densest-first the
// heaviest tiles come first either way). Every wave exits when the counters
// pass t_count.
//   OUT 0: write the packed upper triangle in the caller's sample order
//          (stored position -> sample id through d.perm, if any);
//   OUT 1: write the packed triangle of stored positions, minus out_base
//          (a band buffer of a streamed run).
//   SPLIT: entries light in both blocks (d.lbm) take the sparse path of
//          sm_tile_split (records d.crl), the rest are staged.
template <int CHUNK, int WAVES, int OUT, bool SPLIT, bool SOFF = false>
__device__ __forceinline__ void sm_persistent_body(const SmData &d, int64_t nb, int64_t t_begin, int64_t t_count,
                                                   TileRows g, unsigned long long *__restrict__ work,
                                                   float *__restrict__ out, int64_t out_base) {
    // entry rows (uint16) of one window: 2,048 (split tiles: kSplitList);
    // split tiles also hold the window's word records and the accumulator
    __shared__ __align__(16) uint16_t s_list[WAVES][SPLIT ? kSplitList : 2048];
    __shared__ uint32_t s_stage[WAVES][CHUNK * 64];
    __shared__ __align__(16) uint4 s_words[SPLIT ? WAVES : 1][SPLIT ? kSplitWords : 1];
    __shared__ __align__(16) uint32_t s_acc[SPLIT ? WAVES : 1][SPLIT ? 32 * kAccStride : 4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // each wave takes its own tiles
    int src = (int)(blockIdx.x & (kXcds - 1)), left = kXcds;
    for (;;) {
        const int64_t t = sm_next_tile(work, t_count, src, left, lane);
        if (t < 0) break;
        int64_t ra, cb;
        sm_tile_at(t_begin + (int64_t)t, nb, g, ra, cb);
        // block indices are wave-uniform: keep them (and every address
        // derived from them) in scalar registers
        ra = __builtin_amdgcn_readfirstlane((int)ra);
        cb = __builtin_amdgcn_readfirstlane((int)cb);
        const int64_t abase = ra * kSmBlock, bbase = cb * kSmBlock;
        uint32_t acc[4][4];
        if constexpr (SPLIT) {
            SquareRows<false, SOFF> rows;
            rows.Y = d.Y;
            rows.stride = d.npad;
            rows.ok = true;
            rows.col4 = (uint32_t)((lane < 32) ? abase + lane : bbase + (lane - 32)) * 4u;
            SplitCounts sc;
            const HeavyRows H{d.Y + abase, d.npad, d.Y + bbase, d.npad};
            sm_tile_split<CHUNK, SquareRows<false, SOFF>>(d.bm + ra * d.nwords, d.bm + cb * d.nwords, d.lbm + ra * d.nwords,
                                                    d.lbm + cb * d.nwords, d.crl + d.croff[ra], d.crl + d.croff[cb],
                                                    d.nwords, d.zrow, rows, H, s_list[wave], s_words[wave],
                                                    s_stage[wave], s_acc[wave], lane, acc, sc);
            if (lane == 0) split_add_counts(d.counter, sc);
        } else {
            const uint32_t visited = sm_square_tile<CHUNK, false>(d, d.bm + ra * d.nwords, nullptr, d.bm + cb * d.nwords,
                                                                  nullptr, abase, bbase, s_list[wave], s_stage[wave],
                                                                  lane, acc);
            if (d.counter && lane == 0 && visited) {
                atomicAdd(d.counter + kCntVisited, (unsigned long long)visited);
                atomicAdd(d.counter + kCntStaged, (unsigned long long)visited);
            }
        }
        const int ty = lane >> 3, tx = lane & 7;
#pragma unroll
        for (int ii = 0; ii < 4; ii++)
#pragma unroll
            for (int jj = 0; jj < 4; jj++)
                sm_store_pair<OUT>(d, acc[ii][jj], abase + ty * 4 + ii, bbase + tx * 4 + jj, out, out_base);
    }
}

template <int CHUNK, int WAVES, int OUT>
__global__ __launch_bounds__(64 * WAVES) void sm_persistent_kernel(SmData d, int64_t nb, int64_t t_begin,
                                                                   int64_t t_count, TileRows g,
                                                                   unsigned long long *__restrict__ work,
                                                                   float *__restrict__ out, int64_t out_base) {
    sm_persistent_body<CHUNK, WAVES, OUT, false>(d, nb, t_begin, t_count, g, work, out, out_base);
}

// The split tiles held to 4 waves per SIMD (at most 128 VGPRs): their
// sparse entries wait on record loads, and occupancy hides that latency.
template <int CHUNK, int WAVES, int OUT, bool SOFF>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(CHUNK <= 8 && WAVES <= 2 ? 4 : 1))) void sm_persistent_split_kernel(
    SmData d, int64_t nb, int64_t t_begin, int64_t t_count, TileRows g, unsigned long long *__restrict__ work,
    float *__restrict__ out, int64_t out_base) {
    sm_persistent_body<CHUNK, WAVES, OUT, true, SOFF>(d, nb, t_begin, t_count, g, work, out, out_base);
}

// Split tiles with PARTS waves per tile (1024 / PARTS pairs per wave): a
// workgroup of PARTS waves takes one tile at a time, each wave walks every
// PARTS-th bitmap window, and the waves meet in the workgroup's accumulator:
// more waves per tile fill the device when the sample set has few tiles.
// Integer sums: every PARTS gives the same bits. Up to 168 VGPRs (3 waves per
// SIMD) for the 8-row chunks: at 128 their registers spill.
template <int CHUNK, int PARTS, int OUT>
__global__ __launch_bounds__(64 * PARTS) __attribute__((amdgpu_waves_per_eu(CHUNK <= 8 ? 3 : 1))) void sm_persistent_parts_kernel(
    SmData d, int64_t nb, int64_t t_begin, int64_t t_count, TileRows g, unsigned long long *__restrict__ work,
    float *__restrict__ out, int64_t out_base) {
    __shared__ __align__(16) uint16_t s_list[PARTS][kSplitList];
    __shared__ uint32_t s_stage[PARTS][CHUNK * 64];
    __shared__ __align__(16) uint4 s_words[PARTS][kSplitWords];
    __shared__ __align__(16) uint32_t s_acc[32 * kAccStride];
    __shared__ long long s_tile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ty = lane >> 3, tx = lane & 7;
    int src = (int)(blockIdx.x & (kXcds - 1)), left = kXcds;  // wave 0 deals the workgroup's tiles
    for (;;) {
        if (wave == 0) {
            const int64_t t = sm_next_tile(work, t_count, src, left, lane);
            if (lane == 0) s_tile = t;
        }
        for (int k = threadIdx.x; k < 32 * kAccStride; k += 64 * PARTS) s_acc[k] = 0u;
        __syncthreads();
        const long long t = s_tile;
        if (t < 0) break;
        int64_t ra, cb;
        sm_tile_at(t_begin + (int64_t)t, nb, g, ra, cb);
        ra = __builtin_amdgcn_readfirstlane((int)ra);
        cb = __builtin_amdgcn_readfirstlane((int)cb);
        const int64_t abase = ra * kSmBlock, bbase = cb * kSmBlock;
        SquareRows<false> rows;
        rows.Y = d.Y;
        rows.stride = d.npad;
        rows.ok = true;
        rows.col4 = (uint32_t)((lane < 32) ? abase + lane : bbase + (lane - 32)) * 4u;
        SplitCounts sc;
        uint32_t acc[4][4];
        const HeavyRows H{d.Y + abase, d.npad, d.Y + bbase, d.npad};
        sm_tile_split<CHUNK, SquareRows<false>, true>(
            d.bm + ra * d.nwords, d.bm + cb * d.nwords, d.lbm + ra * d.nwords, d.lbm + cb * d.nwords,
            d.crl + d.croff[ra], d.crl + d.croff[cb], d.nwords, d.zrow, rows, H, s_list[wave], s_words[wave],
            s_stage[wave], s_acc, lane, acc, sc, wave, PARTS, d.pl + ra * d.nwords, d.pl + cb * d.nwords);
        if (lane == 0) split_add_counts(d.counter, sc);
        // every wave's staged sums join the sparse and half sums
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (acc[i][j]) atomicAdd(s_acc + (ty * 4 + i) * kAccStride + tx * 4 + j, acc[i][j]);
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    sm_store_pair<OUT>(d, s_acc[(ty * 4 + i) * kAccStride + tx * 4 + j], abase + ty * 4 + i,
                                       bbase + tx * 4 + j, out, out_base);
        }
        // the accumulator and the tile id are rewritten for the next tile
        __syncthreads();
    }
}

// Explicit pair list (caller's sample ids; d.pos maps them to stored
// positions when samples are binned): lane = pair; 32-entry groups skipped
// when no lane's two samples share mass in them.
__global__ __launch_bounds__(256) void sm_pairs_kernel(SmData d, int64_t n_pairs, const int32_t *__restrict__ prow,
                                                       const int32_t *__restrict__ pcol, float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < n_pairs;
    int64_t a = ok ? prow[i] : 0, b = ok ? pcol[i] : 0;
    if (d.pos) {
        a = d.pos[a];
        b = d.pos[b];
    }
    const uint32_t *BA = d.bm + (a / kSmBlock) * d.nwords, *BB = d.bm + (b / kSmBlock) * d.nwords;
    uint32_t acc = 0u;
    for (int64_t g = 0; g < d.nwords; g++) {
        uint32_t m = ok ? (BA[g] & BB[g]) : 0u;
        if (__ballot(m != 0u) == 0) continue;
        while (m) {
            const int bit = __ffs(m) - 1;
            const int64_t r = g * 32 + bit;
            acc += min(d.Y[r * d.npad + a], d.Y[r * d.npad + b]);
            m &= m - 1u;
        }
    }
    const float sim = sm_finish(d, acc, a, b);
    if (ok) out[i] = sim;
}

// ------------------------------------------------------------ light records

// PL[b][w] = light entries of block b in words before w (one wave per block;
// block b = blocks[blockIdx.x], or blockIdx.x without a list).
__global__ __launch_bounds__(64) void light_prefix_kernel(const uint32_t *__restrict__ lbm, int64_t nwords,
                                                          const int32_t *__restrict__ blocks,
                                                          uint32_t *__restrict__ PL) {
    const int64_t b = blocks ? blocks[blockIdx.x] : blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t *Lw = lbm + b * nwords;
    uint32_t *P = PL + b * nwords;
    uint32_t base = 0;
    for (int64_t w0 = 0; w0 < nwords; w0 += 64) {
        const int64_t w = w0 + lane;
        const int c = (w < nwords) ? __popc(Lw[w]) : 0;
        const int incl = wave_incl_scan(c);
        if (w < nwords) P[w] = base + (uint32_t)(incl - c);
        base += (uint32_t)__builtin_amdgcn_readlane(incl, 63);
    }
}

// One wave per (block, 64-word window): the window's light entries are
// listed in LDS, then each half wave reads one entry's 32 values of Y (one
// 128-byte line; with a mask, the columns whose bit is clear read as zero)
// and writes its record {mask, values of the set bits in ascending sample
// order, zeros after them} at the entry's rank after the block's first
// record (crl + off[b] with packed offsets, crl + b * crs without). Y has row
// stride `stride`; block b = blocks[blockIdx.x] (mask masks[blockIdx.x]), or
// blockIdx.x without a list.
__global__ __launch_bounds__(64) void light_records_kernel(const uint32_t *__restrict__ Y, int64_t stride,
                                                           const int32_t *__restrict__ blocks,
                                                           const uint32_t *__restrict__ masks,
                                                           const uint32_t *__restrict__ lbm,
                                                           const uint32_t *__restrict__ PL, int64_t nwords,
                                                           const int64_t *__restrict__ off,
                                                           uint4 *__restrict__ crl, int64_t crs, int64_t cap) {
    __shared__ uint16_t rows[2048];
    __shared__ uint32_t wrank[64];
    const int64_t b = blocks ? blocks[blockIdx.x] : blockIdx.x, wb = (int64_t)blockIdx.y * 64;
    const uint32_t cols = masks ? masks[blockIdx.x] : 0xffffffffu;
    const int lane = threadIdx.x;
    const int64_t w = wb + lane;
    const uint32_t bits = (w < nwords) ? lbm[b * nwords + w] : 0u;
    const int c = __popc(bits);
    const int incl = wave_incl_scan(c);
    const int tot = __builtin_amdgcn_readlane(incl, 63);
    if (tot == 0) return;
    int pos = incl - c;
    // rank of the window's first light entry: the block prefix at its first word
    wrank[lane] = (w < nwords) ? PL[b * nwords + w] - (uint32_t)(incl - c) : 0u;
    for (uint32_t m = bits; m; m &= m - 1u) rows[pos++] = (uint16_t)(lane * 32 + __ffs(m) - 1);
    __syncthreads();
    const uint32_t rank0 = wrank[0];
    const int half = lane >> 5, col = lane & 31;
    const int64_t first = off ? off[b] : b * crs;
    uint4 *out = crl + first;
    const bool mine = (cols >> col) & 1u;
    const uint32_t *yc = Y + b * kSmBlock + col;
    // entries k0 + 2 u + half, u < kRecUnroll: the rows' loads are all
    // issued before the first is used
    constexpr int kRecUnroll = 8;
    for (int k0 = 0; k0 < tot; k0 += 2 * kRecUnroll) {
        uint32_t v[kRecUnroll];
#pragma unroll
        for (int u = 0; u < kRecUnroll; u++) {
            const int k = k0 + 2 * u + half;
            v[u] = 0u;
            if (mine && k < tot) v[u] = yc[(wb * 32 + rows[k]) * stride];
        }
#pragma unroll
        for (int u = 0; u < kRecUnroll; u++) {
            const int k = k0 + 2 * u + half;
            const bool in = k < tot;
            const uint64_t bal = __ballot(in && v[u] != 0u);
            const uint32_t mask = half ? (uint32_t)(bal >> 32) : (uint32_t)bal;
            const uint32_t below = mask & ((1u << col) - 1u);
            const int r = __popc(below), n = __popc(mask);
            if (in && first + rank0 + k < cap) {
                uint32_t *rec = reinterpret_cast<uint32_t *>(out + rank0 + (uint32_t)k);
                if (v[u] != 0u && r < kSmRecVals) rec[1 + r] = v[u];
                if (col == 0) {
                    rec[0] = mask;
                    for (int f = n; f < kSmRecVals; f++) rec[1 + f] = 0u;
                }
            }
        }
    }
}

// Packed record offsets: off[b] = light entries of blocks 0 .. b - 1 (each
// block's count from its last prefix word), off[nblk] = all. One wave.
__global__ __launch_bounds__(64) void light_offsets_kernel(const uint32_t *__restrict__ lbm,
                                                           const uint32_t *__restrict__ PL, int64_t nblk,
                                                           int64_t nwords, int64_t *__restrict__ off) {
    const int lane = threadIdx.x;
    int64_t base = 0;
    for (int64_t b0 = 0; b0 < nblk; b0 += 64) {
        const int64_t b = b0 + lane;
        const int64_t last = b * nwords + nwords - 1;
        // a block has at most nwords * 32 light entries; 64 of them sum below 2^31
        // for nwords < 2^20 (checked by the launcher)
        const int c = (b < nblk) ? (int)(PL[last] + (uint32_t)__popc(lbm[last])) : 0;
        const int incl = wave_incl_scan(c);
        if (b < nblk) off[b] = base + (int64_t)(incl - c);
        base += (int64_t)__builtin_amdgcn_readlane(incl, 63);
    }
    if (lane == 0) off[nblk] = base;
}

hipError_t launch_light_prefix(const uint32_t *lbm, int64_t nblk, int64_t nwords, uint32_t *PL, int64_t *off,
                               hipStream_t st) {
    if (nblk <= 0 || nwords <= 0) return hipSuccess;
    if (nwords >= ((int64_t)1 << 20)) return hipErrorInvalidValue;
    light_prefix_kernel<<<dim3((unsigned)nblk), 64, 0, st>>>(lbm, nwords, nullptr, PL);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    light_offsets_kernel<<<1, 64, 0, st>>>(lbm, PL, nblk, nwords, off);
    return hipGetLastError();
}

hipError_t launch_light_records(const uint32_t *Y, int64_t npad, const uint32_t *lbm, int64_t nblk, int64_t nwords,
                                const uint32_t *PL, const int64_t *off, uint4 *crl, int64_t cap, hipStream_t st) {
    if (nblk <= 0 || nwords <= 0) return hipSuccess;
    if (!off || npad < nblk * kSmBlock) return hipErrorInvalidValue;
    light_records_kernel<<<dim3((unsigned)nblk, (unsigned)((nwords + 63) / 64)), 64, 0, st>>>(
        Y, npad, nullptr, nullptr, lbm, PL, nwords, off, crl, 0, cap);
    return hipGetLastError();
}

hipError_t launch_light_records_listed(const uint32_t *Y, int64_t stride, const int32_t *blocks,
                                       const uint32_t *masks, int nblocks, const uint32_t *lbm, int64_t nwords,
                                       uint32_t *PL, uint4 *crl, int64_t crs, hipStream_t st) {
    if (nblocks <= 0 || nwords <= 0) return hipSuccess;
    if (crs < nwords * 32 || stride % kSmBlock) return hipErrorInvalidValue;
    light_prefix_kernel<<<dim3((unsigned)nblocks), 64, 0, st>>>(lbm, nwords, blocks, PL);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    light_records_kernel<<<dim3((unsigned)nblocks, (unsigned)((nwords + 63) / 64)), 64, 0, st>>>(
        Y, stride, blocks, masks, lbm, PL, nwords, nullptr, crl, crs, INT64_MAX);
    return hipGetLastError();
}

// ------------------------------------------------------------ launchers

__global__ void zero_counter_kernel(unsigned long long *c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

hipError_t launch_zero_counter(unsigned long long *counter, hipStream_t st) {
    zero_counter_kernel<<<1, 64, 0, st>>>(counter, kWorkWords);
    return hipGetLastError();
}

hipError_t launch_zero_counters(unsigned long long *counter, int n, hipStream_t st) {
    if (n <= 0 || n > 64) return hipErrorInvalidValue;
    zero_counter_kernel<<<1, 64, 0, st>>>(counter, n);
    return hipGetLastError();
}

// Plain streaming copy (16-byte loads and stores, grid-stride): the reference
// point that converts the FETCH_SIZE / WRITE_SIZE counters into bytes.
__global__ __launch_bounds__(256) void calib_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                         int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_calib_copy(const void *src, void *dst, int64_t bytes, hipStream_t st) {
    const int64_t n16 = bytes / 16;
    if (n16 <= 0) return hipSuccess;
    calib_copy_kernel<<<4096, 256, 0, st>>>((const uint4 *)src, (uint4 *)dst, n16);
    return hipGetLastError();
}

hipError_t launch_sample_mass(const int64_t *rowptr, const int32_t *col, const float *val, int n,
                              const int32_t *leaf_uses, double *mass, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    sample_mass_kernel<<<(n + 3) / 4, dim3(64, 4), 0, st>>>(rowptr, col, val, n, leaf_uses, mass);
    return hipGetLastError();
}

template <bool UNW, bool SUMS, bool SPARSE>
static hipError_t launch_embed_t(const EmbedData &e, hipStream_t st) {
    const int nslots = e.nslots > 0 ? e.nslots : 1;
    const size_t lds = (size_t)nslots * 64 * sizeof(double);
    if (lds + sizeof(float) * kEmbedLeafChunk * 64 + sizeof(EmbedNode) * kEmbedNodeChunk > 65536)
        return hipErrorInvalidConfiguration;
    const unsigned waves = (unsigned)(e.npad / 64);
    const int64_t ystride = e.ystride ? e.ystride : e.npad;
    if (e.nseg > 0) {
        embed_seg_kernel<UNW, SUMS, SPARSE><<<dim3(waves, (unsigned)e.nseg), 64, lds, st>>>(
            e.nodes, e.seg_first, e.seg_last, e.order_n, e.ab, e.ll, e.nseg, e.nlp, e.n, e.npad, e.nwords, ystride,
            e.scale, e.Y, e.bm, e.lbm, e.xroot, e.colsum, e.colsum_f64);
        hipError_t r = hipGetLastError();
        if (r != hipSuccess) return r;
    }
    if (e.ntop > 0)
        embed_top_kernel<UNW, SUMS, SPARSE><<<waves, 64, lds, st>>>(e.top, e.ntop, e.ab, e.ll, e.nlp, e.n, e.npad,
                                                                    e.nwords, ystride, e.scale, e.Y, e.bm, e.lbm,
                                                                    e.xroot, e.colsum, e.colsum_f64);
    return hipGetLastError();
}

template <bool SPARSE>
static hipError_t launch_embed_s(const EmbedData &e, hipStream_t st) {
    if (e.unweighted)
        return e.sums_only ? launch_embed_t<true, true, SPARSE>(e, st) : launch_embed_t<true, false, SPARSE>(e, st);
    return e.sums_only ? launch_embed_t<false, true, SPARSE>(e, st) : launch_embed_t<false, false, SPARSE>(e, st);
}

hipError_t launch_embed(const EmbedData &e, hipStream_t st) {
    if (e.n <= 0 || e.order_n <= 0) return hipSuccess;
    return e.sparse ? launch_embed_s<true>(e, st) : launch_embed_s<false>(e, st);
}

// Persistent launch: variant (chunk, waves) chosen by the caller / autotuner.
template <int OUT>
static hipError_t launch_persistent_t(const SmData &d, int variant, int64_t nb, int64_t t_begin, int64_t t_count,
                                      unsigned long long *work, float *out, int64_t out_base, int grid_waves,
                                      hipStream_t st) {
    if (t_count <= 0) return hipSuccess;
    const TileRows g = host_tile_rows(t_begin, t_count, nb);
    // the split tiles' rows as 32-bit byte offsets from the window's first
    // row: list rows (below kSplitList) times the stride times 4
    const bool r32 = !d.wide_rows && (uint64_t)kSplitList * (uint64_t)d.npad * 4ull < (1ull << 32);
#define CMS_PERSIST(CH, WV, SP)                                                                                  \
    {                                                                                                            \
        const int64_t wg = std::max<int64_t>(1, std::min<int64_t>((t_count + WV - 1) / WV, grid_waves / WV));   \
        if constexpr (SP) {                                                                                      \
            if (r32)                                                                                             \
                sm_persistent_split_kernel<CH, WV, OUT, true><<<dim3((unsigned)wg), 64 * WV, 0, st>>>(           \
                    d, nb, t_begin, t_count, g, work, out, out_base);                                            \
            else                                                                                                 \
                sm_persistent_split_kernel<CH, WV, OUT, false><<<dim3((unsigned)wg), 64 * WV, 0, st>>>(          \
                    d, nb, t_begin, t_count, g, work, out, out_base);                                            \
        } else                                                                                                   \
            sm_persistent_kernel<CH, WV, OUT><<<dim3((unsigned)wg), 64 * WV, 0, st>>>(d, nb, t_begin, t_count,   \
                                                                                   g, work, out, out_base);       \
    }
#define CMS_PARTS(CH, PT)                                                                                        \
    {                                                                                                            \
        const int64_t wg = std::max<int64_t>(1, std::min<int64_t>(t_count, grid_waves / PT));                    \
        sm_persistent_parts_kernel<CH, PT, OUT><<<dim3((unsigned)wg), 64 * PT, 0, st>>>(d, nb, t_begin, t_count,  \
                                                                                      g, work, out, out_base);    \
    }
    // the split tiles read the light records and the zero row; tiles shared
    // by several waves also the light prefixes
    if (variant >= 6 && (!d.lbm || !d.crl || !d.croff || d.zrow < d.nrows)) return hipErrorInvalidValue;
    if (variant >= 12 && !d.pl) return hipErrorInvalidValue;
    switch (variant) {
    case 0: CMS_PERSIST(8, 4, false); break;
    case 1: CMS_PERSIST(16, 4, false); break;
    case 2: CMS_PERSIST(32, 4, false); break;
    case 3: CMS_PERSIST(16, 2, false); break;
    case 4: CMS_PERSIST(32, 2, false); break;
    case 5: CMS_PERSIST(16, 8, false); break;
    case 6: CMS_PERSIST(8, 2, true); break;
    case 7: CMS_PERSIST(8, 1, true); break;
    case 8: CMS_PERSIST(8, 4, true); break;
    case 9: CMS_PERSIST(16, 2, true); break;
    case 10: CMS_PERSIST(16, 1, true); break;
    case 11: CMS_PERSIST(16, 4, true); break;
    case 12: CMS_PARTS(8, 2); break;
    case 13: CMS_PARTS(8, 4); break;
    case 14: CMS_PARTS(16, 2); break;
    default: CMS_PARTS(16, 4); break;
    }
#undef CMS_PERSIST
#undef CMS_PARTS
    return hipGetLastError();
}


hipError_t launch_sm_persistent(const SmData &d, int variant, int64_t nb, int64_t t_begin, int64_t t_count,
                                unsigned long long *work, float *out, int64_t out_cap, int out_mode,
                                int64_t out_base, int grid_waves, hipStream_t st) {
    if (t_count <= 0) return hipSuccess;
    if (t_begin < 0 || t_begin + t_count > nb * (nb + 1) / 2 || nb * kSmBlock > d.npad) return hipErrorInvalidValue;
    // every packed index the launch writes must fall inside out[0, out_cap)
    const int64_t n = d.n;
    auto row_start = [n](int64_t r) { return r * n - r * (r + 1) / 2; };
    if (out_mode == 0) {
        if (out_cap < n * (n - 1) / 2) return hipErrorInvalidValue;
    } else {
        int64_t rb0, cb0, rb1, cb1;
        host_decode_upper(t_begin, nb, rb0, cb0);
        host_decode_upper(t_begin + t_count - 1, nb, rb1, cb1);
        const int64_t lo = row_start(std::min(rb0 * kSmBlock, n)), hi = row_start(std::min((rb1 + 1) * kSmBlock, n));
        if (lo < out_base || hi - out_base > out_cap) return hipErrorInvalidValue;
    }
    return out_mode == 0 ? launch_persistent_t<0>(d, variant, nb, t_begin, t_count, work, out, out_base, grid_waves, st)
                         : launch_persistent_t<1>(d, variant, nb, t_begin, t_count, work, out, out_base, grid_waves, st);
}

hipError_t launch_sm_rect(const SmData &d, int r0, int nr, int c0, int nc, float *out, hipStream_t st) {
    const int64_t nbr = (nr + kSmBlock - 1) / kSmBlock, nbc = (nc + kSmBlock - 1) / kSmBlock;
    const int64_t ntiles = nbr * nbc;
    if (ntiles <= 0) return hipSuccess;
    const int64_t nwg = (ntiles + kSmWaves - 1) / kSmWaves;
    sm_rect_kernel<<<dim3((unsigned)nwg), 64 * kSmWaves, 0, st>>>(d, ntiles, nbc, r0, nr, c0, nc, out);
    return hipGetLastError();
}

hipError_t launch_sm_pairs(const SmData &d, int64_t n_pairs, const int32_t *row, const int32_t *col, float *out,
                           hipStream_t st) {
    if (n_pairs <= 0) return hipSuccess;
    sm_pairs_kernel<<<dim3((unsigned)((n_pairs + 255) / 256)), 256, 0, st>>>(d, n_pairs, row, col, out);
    return hipGetLastError();
}

}  // namespace cms
