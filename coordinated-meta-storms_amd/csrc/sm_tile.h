// Pair-tile accumulation of the shared-minimum kernels, shared by the
// all-pairs kernels and the serving cross kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cms {

// Row loaders: value of this lane's sample at row rbase + r of Y, with
// window(rbase) called once per 64-word bitmap window and r < 2048. Row
// addresses come from the wave-uniform row index in scalar registers (64-bit
// r * stride, so any stride the allocation allows is addressed correctly),
// and the lane's column is a 32-bit byte offset, so a staged row costs one
// vector load and no vector address arithmetic.
//
// One matrix (all-pairs tiles): lane column `col` (row block sample for lanes
// 0..31, column block sample for lanes 32..63) of Y with row stride `stride`.
// CHECK: a lane whose column lies outside the matrix (ok false) loads column
// 0 and stages zero.
template <bool CHECK>
struct SquareRows {
    const uint32_t *Y;
    int64_t stride;
    uint32_t col4;  // column * 4 (bytes)
    bool ok;
    const uint32_t *win = nullptr;
    __device__ __forceinline__ void window(int64_t rbase) { win = Y + rbase * stride; }
    __device__ __forceinline__ uint32_t load(uint32_t r) const {
        const char *row = reinterpret_cast<const char *>(win + (uint64_t)r * (uint64_t)stride);
        const uint32_t v = *reinterpret_cast<const uint32_t *>(row + col4);
        return (!CHECK || ok) ? v : 0u;
    }
};

// Two matrices (serving): lanes 0..31 read QY (query slots of one block,
// row stride qs), lanes 32..63 read Y (references of one block, stride ys).
struct CrossRows {
    const uint32_t *QY;
    int64_t qs;
    const uint32_t *Y;
    int64_t ys;
    uint32_t col4;  // (lane & 31) * 4
    bool lo;        // lane < 32
    const uint32_t *qwin = nullptr, *ywin = nullptr;
    __device__ __forceinline__ void window(int64_t rbase) {
        qwin = QY + rbase * qs;
        ywin = Y + rbase * ys;
    }
    __device__ __forceinline__ uint32_t load(uint32_t r) const {
        uint32_t v;
        if (lo) {
            const char *row = reinterpret_cast<const char *>(qwin + (uint64_t)r * (uint64_t)qs);
            v = *reinterpret_cast<const uint32_t *>(row + col4);
        } else {
            const char *row = reinterpret_cast<const char *>(ywin + (uint64_t)r * (uint64_t)ys);
            v = *reinterpret_cast<const uint32_t *>(row + col4);
        }
        return v;
    }
};

// Per-entry pair operation of the tiles: acc += f(a0, b0) + f(a1, b1) for
// two staged rows. kColSum: the tile also sums each lane's own column over
// the visited rows and hands both to finish().
//
// Shared minimum, direct: integer minima added (one v_min_u32 per pair-entry
// and one v_add3_u32 per two: 1.5 VALU ops per pair-entry).
struct MinAddOp {
    typedef uint32_t Acc;
    static constexpr bool kColSum = false;
    __device__ __forceinline__ static void add2(Acc &acc, uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1) {
        acc = acc + min(a0, b0) + min(a1, b1);
    }
};

// Shared minimum through absolute differences: over the visited rows I,
//   sum_I min(a, b) = (sum_I a + sum_I b - sum_I |a - b|) / 2,
// one v_sad_u32 (|a - b| + acc) per pair-entry plus one add per staged row
// and lane for the column sums (1 + 1/16 VALU ops per pair-entry). Exact in
// uint32 when every value and column sum is below 2^31: then a - b does not
// overflow as a signed difference and sum a + sum b < 2^32.
struct SadOp {
    typedef uint32_t Acc;
    static constexpr bool kColSum = true;
    __device__ __forceinline__ static uint32_t sad(uint32_t a, uint32_t b, uint32_t c) {
        uint32_t d;
        asm("v_sad_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
        return d;
    }
    __device__ __forceinline__ static void add2(Acc &acc, uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1) {
        acc = sad(a1, b1, sad(a0, b0, acc));
    }
    // acc = sum |a - b| -> sum min(a, b), given both columns' sums
    __device__ __forceinline__ static uint32_t finish(uint32_t acc, uint32_t sa, uint32_t sb) {
        return (sa + sb - acc) >> 1;
    }
};

// The shared-minimum kernels' pair operation.
typedef SadOp SmOp;

// One wave = 32 x 32 pairs, lane (ty, tx) = 4 rows x 4 columns.
//
// Visit the entries set in both blocks' bitmaps (OR of two blocks when a
// side is not block-aligned): per 64 bitmap words, compact the set bits into
// an LDS list, stage their 2 x 32 Y values in LDS CHUNK rows at a time
// (through `rows`, see the loaders above), and accumulate Op into the lane's
// 4 x 4 registers (for the shared-minimum ops: the 4 x 4 shared sums).
// Returns the number of entries visited.
//
// Only bitmap words [w_begin, w_end) are visited (w_begin a multiple of 64):
// a tile's rows can be split over several waves whose integer partial sums
// add up exactly (each part's sums are shared minima on their own).
template <int CHUNK, class Rows, class Op = SmOp>
__device__ __forceinline__ uint32_t sm_tile_accumulate(const uint32_t *A0, const uint32_t *A1, const uint32_t *B0,
                                                       const uint32_t *B1, int64_t w_begin, int64_t w_end,
                                                       const Rows &rows, uint16_t *L, uint32_t *S, int lane,
                                                       typename Op::Acc (&acc)[4][4]) {
    static_assert(CHUNK % 2 == 0, "rows are staged in pairs");
    const int ty = lane >> 3, tx = lane & 7;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;
    uint32_t visited = 0;
    uint32_t cs = 0;  // kColSum: this lane's column summed over the visited rows
    for (int64_t wb = w_begin; wb < w_end; wb += 64) {
        const int64_t wi = wb + lane;
        uint32_t w = 0u;
        if (wi < w_end) {
            uint32_t a = A0[wi], b = B0[wi];
            if (A1) a |= A1[wi];
            if (B1) b |= B1[wi];
            w = a & b;
        }
        const int cnt = __popc(w);
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const int tot = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
        if (tot == 0) continue;
        visited += (uint32_t)tot;
        int pos = incl - cnt;
        __builtin_amdgcn_wave_barrier();
        while (w) {
            const int b = __ffs(w) - 1;
            L[pos++] = (uint16_t)(lane * 32 + b);
            w &= w - 1u;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        Rows win = rows;
        win.window(wb * 32);
        // Stage CHUNK listed rows into registers. Row ids are wave-uniform
        // (every lane reads the same LDS word); slots past the list end load
        // the chunk's first row (a valid address) and are never accumulated.
        uint32_t v[CHUNK];
        auto load_chunk = [&](int j) {
            const uint32_t first =
                (uint32_t)__builtin_amdgcn_readfirstlane((int)*reinterpret_cast<const uint32_t *>(L + j)) & 0xffffu;
#pragma unroll
            for (int k = 0; k < CHUNK; k += 2) {
                const uint32_t pr =
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)*reinterpret_cast<const uint32_t *>(L + j + k));
                const uint32_t r0 = (j + k < tot) ? (pr & 0xffffu) : first;
                const uint32_t r1 = (j + k + 1 < tot) ? (pr >> 16) : first;
                v[k] = win.load(r0);
                v[k + 1] = win.load(r1);
            }
        };
        load_chunk(0);
        for (int j = 0; j < tot; j += CHUNK) {
            const int kn = min(CHUNK, tot - j);
            if constexpr (Op::kColSum) {
                if (kn == CHUNK) {
#pragma unroll
                    for (int k = 0; k < CHUNK; k++) cs += v[k];
                } else {
#pragma unroll
                    for (int k = 0; k < CHUNK; k++) cs += (k < kn) ? v[k] : 0u;
                }
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < CHUNK; k++) S[k * 64 + lane] = v[k];
            if (kn & 1) S[kn * 64 + lane] = 0u;  // rows are accumulated in pairs
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the next chunk's loads are in flight while this one accumulates
            if (j + CHUNK < tot) load_chunk(j + CHUNK);
            // two staged rows: 4 x 4 pair-entries each
            auto rows2 = [&](int k) {
                const uint4 a0 = *reinterpret_cast<const uint4 *>(S + k * 64 + ty * 4);
                const uint4 b0 = *reinterpret_cast<const uint4 *>(S + k * 64 + 32 + tx * 4);
                const uint4 a1 = *reinterpret_cast<const uint4 *>(S + (k + 1) * 64 + ty * 4);
                const uint4 b1 = *reinterpret_cast<const uint4 *>(S + (k + 1) * 64 + 32 + tx * 4);
                const uint32_t av0[4] = {a0.x, a0.y, a0.z, a0.w}, bv0[4] = {b0.x, b0.y, b0.z, b0.w};
                const uint32_t av1[4] = {a1.x, a1.y, a1.z, a1.w}, bv1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
                for (int ii = 0; ii < 4; ii++)
#pragma unroll
                    for (int jj = 0; jj < 4; jj++)
                        Op::add2(acc[ii][jj], av0[ii], bv0[jj], av1[ii], bv1[jj]);
            };
            if (Op::kColSum && kn == CHUNK) {
                // full chunk of the integer tiles: unrolled, LDS offsets are
                // immediates and the reads can be issued ahead of the
                // arithmetic (the FP64 leaf ops keep the loop)
#pragma unroll
                for (int k = 0; k < CHUNK; k += 2) rows2(k);
            } else {
                for (int k = 0; k < kn; k += 2) rows2(k);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    if constexpr (Op::kColSum) {
        // lanes 0..31 hold the row block's columns, 32..63 the column block's
        uint32_t sa[4], sb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            sa[i] = (uint32_t)__shfl((int)cs, ty * 4 + i, 64);
            sb[i] = (uint32_t)__shfl((int)cs, 32 + tx * 4 + i, 64);
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = Op::finish(acc[i][j], sa[i], sb[j]);
    }
    return visited;
}

// Similarity from a pair's shared sum: weighted, acc / (100 * 2^K)
// (inv_scale = 1 / (100 * 2^K)); unweighted, the shared weight over the
// weight of either sample, acc / (W_a + W_b - acc) with W the column sums of
// the stored y. Clamped to 1 (rounding of y can push a self-like pair over).
__device__ __forceinline__ float sm_similarity(bool unweighted, double inv_scale, uint32_t acc,
                                               unsigned long long wa, unsigned long long wb) {
    double t;
    if (unweighted) {
        const unsigned long long den = wa + wb - acc;
        t = den ? (double)acc / (double)den : __builtin_nan("");
    } else {
        t = (double)acc * inv_scale;
    }
    t = t > 1.0 ? 1.0 : t;
    return (float)t;
}

}  // namespace cms
