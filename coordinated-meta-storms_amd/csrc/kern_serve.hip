// Serving kernels: query sets against the resident references.
//
//   sm_cross_kernel   persistent; tiles = (active 32-slot query block) x
//                     (32-sample reference block of the step's band); the
//                     shared-minimum accumulation of the all-pairs kernels
//                     with the row side read from the query matrix QY and the
//                     column side from the references' Y;
//   qbm_build_kernel  per 32-slot query block, the bitmap of rows where an
//                     active slot of the block has mass;
//   pack_kernel       copies a query set's embedded columns (staging matrix)
//                     into its slots of QY, with their column sums.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"
#include "sm_tile.h"

namespace cms {

// Output of one tile element: query slot x reference stored position p.
__device__ __forceinline__ void cross_store(const CrossData &x, float *const *slot_out, int64_t slot, int64_t p,
                                            uint32_t acc) {
    float *o = slot_out[slot];
    if (!o || p >= x.n_ref) return;
    const unsigned long long wq = x.unweighted ? x.qcolsum[slot] : 0ull;
    const unsigned long long wr = x.unweighted ? x.colsum[p] : 0ull;
    o[x.perm ? x.perm[p] : p] = sm_similarity(x.unweighted, x.inv_scale, acc, wq, wr);
}

// Work unit u = tile t * ks + part c; tiles are numbered reference-block-major
// (t = rb_local * nqa + k), so the query blocks of one reference block are
// pulled together and share its Y columns in L2. Part c visits bitmap words
// [c * cw, (c + 1) * cw) of the tile (cw a multiple of 64). With ks == 1
// (SPLIT false) a unit is a whole tile and writes the similarities; with
// SPLIT its partial sums are added (exact integer atomics) into
// part_acc[t][k][lane] and cross_finish_kernel writes them. Every wave exits
// when the counter passes the unit count.
template <int CHUNK, int WAVES, bool SPLIT>
__global__ __launch_bounds__(64 * WAVES) void sm_cross_kernel(CrossData x, const int32_t *__restrict__ qblocks,
                                                              int nqa, int64_t rb0, int64_t nunits, int ks,
                                                              int64_t cw, float *const *__restrict__ slot_out,
                                                              uint32_t *__restrict__ part_acc,
                                                              unsigned long long *__restrict__ work) {
    __shared__ uint16_t s_list[WAVES][64 * 32];
    __shared__ uint32_t s_stage[WAVES][CHUNK * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ty = lane >> 3, tx = lane & 7;
    uint32_t visited = 0;
    for (;;) {
        unsigned long long u = 0;
        if (lane == 0) u = atomicAdd(work, 1ull);
        u = __shfl(u, 0, 64);
        if (u >= (unsigned long long)nunits) break;
        const int64_t t = (int64_t)u / ks, c = (int64_t)u % ks;
        const int64_t w0 = c * cw, w1 = min(x.nwords, w0 + cw);
        if (w0 >= w1) continue;
        const int64_t rb = rb0 + t / nqa;
        const int64_t qb = qblocks[t % nqa];
        CrossRows rows;
        rows.QY = x.QY + qb * kSmBlock;
        rows.qs = x.qstride;
        rows.Y = x.Y + rb * kSmBlock;
        rows.ys = x.npad;
        rows.col4 = (uint32_t)(lane & 31) * 4u;
        rows.lo = lane < 32;
        uint32_t acc[4][4];
        visited += sm_tile_accumulate<CHUNK>(x.qbm + qb * x.nwords, nullptr, x.bm + rb * x.nwords, nullptr, w0, w1,
                                             rows, s_list[wave], s_stage[wave], lane, acc);
        if constexpr (SPLIT) {
            uint32_t *pa = part_acc + t * 1024 + lane;
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    if (acc[ii][jj]) atomicAdd(pa + (ii * 4 + jj) * 64, acc[ii][jj]);
        } else {
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    cross_store(x, slot_out, qb * kSmBlock + ty * 4 + ii, rb * kSmBlock + tx * 4 + jj, acc[ii][jj]);
        }
    }
    if (x.counter && lane == 0 && visited) {
        atomicAdd(x.counter + kCntVisited, (unsigned long long)visited);
        atomicAdd(x.counter + kCntStaged, (unsigned long long)visited);
    }
}

// The same step with split tiles (sm_tile_split): entries light in a block
// are read from the light records (the query blocks' are rebuilt when their
// slots change). With SPLIT, part c of ks takes the tile's windows c, c + ks,
// ... and adds its shared sums (staged rows + its accumulator) into
// part_acc.
template <int CHUNK, int WAVES, bool SPLIT>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(4))) void sm_cross_split_kernel(
    CrossData x, const int32_t *__restrict__ qblocks, int nqa, int64_t rb0, int64_t nunits, int ks,
    float *const *__restrict__ slot_out, uint32_t *__restrict__ part_acc, unsigned long long *__restrict__ work) {
    __shared__ __align__(16) uint16_t s_list[WAVES][kSplitList];
    __shared__ uint32_t s_stage[WAVES][CHUNK * 64];
    __shared__ __align__(16) uint4 s_words[WAVES][kSplitWords];
    __shared__ __align__(16) uint32_t s_acc[WAVES][32 * kAccStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ty = lane >> 3, tx = lane & 7;
    for (;;) {
        unsigned long long u = 0;
        if (lane == 0) u = atomicAdd(work, 1ull);
        u = __shfl(u, 0, 64);
        if (u >= (unsigned long long)nunits) break;
        const int64_t t = (int64_t)u / ks;
        const int c = (int)((int64_t)u % ks);
        const int64_t rb = __builtin_amdgcn_readfirstlane((int)(rb0 + t / nqa));
        const int64_t qb = __builtin_amdgcn_readfirstlane(qblocks[t % nqa]);
        CrossRows rows;
        rows.QY = x.QY + qb * kSmBlock;
        rows.qs = x.qstride;
        rows.Y = x.Y + rb * kSmBlock;
        rows.ys = x.npad;
        rows.col4 = (uint32_t)(lane & 31) * 4u;
        rows.lo = lane < 32;
        const HeavyRows H{x.QY + qb * kSmBlock, x.qstride, x.Y + rb * kSmBlock, x.npad};
        const int64_t nw = x.nwords;
        SplitCounts sc;
        uint32_t acc[4][4];
        if constexpr (SPLIT) {
            uint32_t *ACC = s_acc[wave];
            for (int k = lane; k < 32 * kAccStride; k += 64) ACC[k] = 0u;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            sm_tile_split<CHUNK, CrossRows, true>(x.qbm + qb * nw, x.bm + rb * nw, x.qlbm + qb * nw, x.lbm + rb * nw,
                                                  x.qcrl + qb * x.qcrs, x.crl + x.croff[rb], nw, x.zrow, rows, H,
                                                  s_list[wave], s_words[wave], s_stage[wave], ACC, lane, acc, sc, c,
                                                  ks, x.qpl + qb * nw, x.pl + rb * nw);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            uint32_t *pa = part_acc + t * 1024 + lane;
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const uint32_t v = acc[ii][jj] + ACC[(ty * 4 + ii) * kAccStride + tx * 4 + jj];
                    if (v) atomicAdd(pa + (ii * 4 + jj) * 64, v);
                }
        } else {
            sm_tile_split<CHUNK, CrossRows, false>(x.qbm + qb * nw, x.bm + rb * nw, x.qlbm + qb * nw, x.lbm + rb * nw,
                                                   x.qcrl + qb * x.qcrs, x.crl + x.croff[rb], nw, x.zrow, rows, H,
                                                   s_list[wave], s_words[wave], s_stage[wave], s_acc[wave], lane, acc,
                                                   sc);
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    cross_store(x, slot_out, qb * kSmBlock + ty * 4 + ii, rb * kSmBlock + tx * 4 + jj, acc[ii][jj]);
        }
        if (lane == 0) split_add_counts(x.counter, sc);
    }
}

// Similarities of split tiles from their summed parts; clears the sums for
// the next step. One thread per tile element (t, k, lane).
__global__ __launch_bounds__(256) void cross_finish_kernel(CrossData x, const int32_t *__restrict__ qblocks, int nqa,
                                                           int64_t rb0, int64_t ntiles,
                                                           float *const *__restrict__ slot_out,
                                                           uint32_t *__restrict__ part_acc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntiles * 1024) return;
    const int64_t t = i >> 10;
    const int k = (int)((i >> 6) & 15), lane = (int)(i & 63);
    const uint32_t acc = part_acc[i];
    part_acc[i] = 0u;
    const int64_t rb = rb0 + t / nqa, qb = qblocks[t % nqa];
    const int ty = lane >> 3, tx = lane & 7;
    cross_store(x, slot_out, qb * kSmBlock + ty * 4 + (k >> 2), rb * kSmBlock + tx * 4 + (k & 3), acc);
}

// One thread per (row, listed block): does an active slot of the block have
// mass in that row (qbm), and do 1 to kSmRecVals of them (qlbm, if given: the
// split tiles' light bitmap)? A wave covers 64 consecutive rows = two bitmap
// words.
__global__ __launch_bounds__(256) void qbm_build_kernel(const uint32_t *__restrict__ QY, int64_t qstride,
                                                        int64_t nrows, int64_t nwords,
                                                        const int32_t *__restrict__ blocks,
                                                        const uint32_t *__restrict__ masks,
                                                        uint32_t *__restrict__ qbm, uint32_t *__restrict__ qlbm) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blocks[blockIdx.y];
    const uint32_t mask = masks[blockIdx.y];
    int cnt = 0;  // active slots with mass
    if (r < nrows && mask) {
        const uint4 *row = reinterpret_cast<const uint4 *>(QY + r * qstride + b * kSmBlock);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint4 v = row[i];
            const uint32_t m = mask >> (4 * i);
            cnt += ((m & 1u) && v.x) + ((m & 2u) && v.y) + ((m & 4u) && v.z) + ((m & 8u) && v.w);
        }
    }
    const uint64_t bits = __ballot(cnt > 0);
    const uint64_t light = __ballot(cnt > 0 && cnt <= kSmRecVals);
    const int lane = threadIdx.x & 63;
    const int64_t w0 = (r - lane) / 32;  // first word of this wave's 64 rows
    if (lane == 0 && w0 < nwords) {
        qbm[b * nwords + w0] = (uint32_t)bits;
        if (qlbm) qlbm[b * nwords + w0] = (uint32_t)light;
    }
    if (lane == 1 && w0 + 1 < nwords) {
        qbm[b * nwords + w0 + 1] = (uint32_t)(bits >> 32);
        if (qlbm) qlbm[b * nwords + w0 + 1] = (uint32_t)(light >> 32);
    }
}

__global__ __launch_bounds__(256) void pack_kernel(const uint32_t *__restrict__ Ys, int64_t sstride, int nq,
                                                   const int32_t *__restrict__ slots, int64_t nrows,
                                                   uint32_t *__restrict__ QY, int64_t qstride,
                                                   const unsigned long long *__restrict__ cs,
                                                   unsigned long long *__restrict__ qcs) {
    const int64_t total = nrows * nq;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / nq;
        const int q = (int)(i - r * nq);
        QY[r * qstride + slots[q]] = Ys[r * sstride + q];
        if (r == 0) qcs[slots[q]] = cs[q];
    }
}

// ------------------------------------------------------------ launchers

hipError_t launch_sm_cross(const CrossData &x, int variant, const int32_t *qblocks, int nqa, int64_t rb0, int64_t nrb,
                           float *const *slot_out, uint32_t *part_acc, int64_t part_tiles, unsigned long long *work,
                           int grid_waves, hipStream_t st) {
    const int64_t ntiles = (int64_t)nqa * nrb;
    if (ntiles <= 0) return hipSuccess;
    if (rb0 < 0 || (rb0 + nrb) * kSmBlock > x.npad || x.qstride % kSmBlock) return hipErrorInvalidValue;
    const bool split_tiles = variant >= 6;
    if (split_tiles && (!x.qlbm || !x.qpl || !x.qcrl || !x.lbm || !x.pl || !x.crl || !x.croff || x.zrow < 0))
        return hipErrorInvalidValue;
    // split the rows of each tile until the step has about four units per
    // resident wave (parts of at least 64 bitmap words; split tiles: every
    // ks-th window of kSplitWords words)
    const int64_t groups = split_tiles ? (x.nwords + kSplitWords - 1) / kSplitWords : (x.nwords + 63) / 64;
    int64_t ks = std::min<int64_t>(groups, std::max<int64_t>(1, (4 * (int64_t)grid_waves + ntiles - 1) / ntiles));
    if (!part_acc || ntiles > part_tiles) ks = 1;
    const int64_t cw = ((groups + ks - 1) / ks) * 64;
    if (!split_tiles) ks = std::max<int64_t>(1, (x.nwords + cw - 1) / cw);
    const int64_t nunits = ntiles * ks;
#define CMS_CROSS_SPLIT(CH, WV)                                                                                  \
    {                                                                                                            \
        const int64_t wg = std::max<int64_t>(1, std::min<int64_t>((nunits + WV - 1) / WV, grid_waves / WV));   \
        if (ks > 1)                                                                                              \
            sm_cross_split_kernel<CH, WV, true><<<dim3((unsigned)wg), 64 * WV, 0, st>>>(                        \
                x, qblocks, nqa, rb0, nunits, (int)ks, slot_out, part_acc, work);                                \
        else                                                                                                     \
            sm_cross_split_kernel<CH, WV, false><<<dim3((unsigned)wg), 64 * WV, 0, st>>>(                       \
                x, qblocks, nqa, rb0, nunits, 1, slot_out, part_acc, work);                                      \
    }
#define CMS_CROSS(CH, WV)                                                                                        \
    {                                                                                                            \
        const int64_t wg = std::max<int64_t>(1, std::min<int64_t>((nunits + WV - 1) / WV, grid_waves / WV));   \
        if (ks > 1)                                                                                              \
            sm_cross_kernel<CH, WV, true><<<dim3((unsigned)wg), 64 * WV, 0, st>>>(                              \
                x, qblocks, nqa, rb0, nunits, (int)ks, cw, slot_out, part_acc, work);                            \
        else                                                                                                     \
            sm_cross_kernel<CH, WV, false><<<dim3((unsigned)wg), 64 * WV, 0, st>>>(                             \
                x, qblocks, nqa, rb0, nunits, 1, cw, slot_out, part_acc, work);                                  \
    }
    switch (variant) {
    case 0: CMS_CROSS(8, 4); break;
    case 1: CMS_CROSS(16, 4); break;
    case 2: CMS_CROSS(32, 4); break;
    case 3: CMS_CROSS(16, 2); break;
    case 4: CMS_CROSS(32, 2); break;
    case 5: CMS_CROSS(16, 8); break;
    case 6: CMS_CROSS_SPLIT(8, 1); break;
    default: CMS_CROSS_SPLIT(8, 2); break;
    }
#undef CMS_CROSS
#undef CMS_CROSS_SPLIT
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ks == 1) return e;
    const int64_t nthreads = ntiles * 1024;
    cross_finish_kernel<<<dim3((unsigned)((nthreads + 255) / 256)), 256, 0, st>>>(x, qblocks, nqa, rb0, ntiles,
                                                                                  slot_out, part_acc);
    return hipGetLastError();
}

hipError_t launch_qbm_build(const uint32_t *QY, int64_t qstride, int64_t nrows, int64_t nwords, const int32_t *blocks,
                            const uint32_t *masks, int nblocks, uint32_t *qbm, uint32_t *qlbm, hipStream_t st) {
    if (nblocks <= 0 || nrows <= 0) return hipSuccess;
    if (qstride % kSmBlock) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((nrows + 255) / 256), (unsigned)nblocks);
    qbm_build_kernel<<<grid, 256, 0, st>>>(QY, qstride, nrows, nwords, blocks, masks, qbm, qlbm);
    return hipGetLastError();
}

hipError_t launch_pack(const uint32_t *Ys, int64_t sstride, int nq, const int32_t *slots, int64_t nrows, uint32_t *QY,
                       int64_t qstride, const unsigned long long *cs, unsigned long long *qcs, hipStream_t st) {
    if (nq <= 0 || nrows <= 0) return hipSuccess;
    const int64_t total = nrows * nq;
    const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 8192);
    pack_kernel<<<grid, 256, 0, st>>>(Ys, sstride, nq, slots, nrows, QY, qstride, cs, qcs);
    return hipGetLastError();
}

}  // namespace cms
