// Pair-tile accumulation of the shared-minimum kernels, shared by the
// all-pairs kernels and the serving cross kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cms {

// One wave = 32 x 32 pairs, lane (ty, tx) = 4 rows x 4 columns.
//
// Visit the entries set in both blocks' bitmaps (OR of two blocks when a
// side is not block-aligned): per 64 bitmap words, compact the set bits into
// an LDS list, stage their 2 x 32 Y values in LDS CHUNK rows at a time, and
// accumulate min/add into the lane's 4 x 4 registers. Returns the number of
// entries visited.
//
// Ycol/ystride: the Y column this lane stages (lanes 0..31 the 32 row
// samples, lanes 32..63 the 32 column samples) and its row stride; s_ok
// false for a lane whose column lies outside the matrix (it stages zeros).
//
// Only bitmap words [w_begin, w_end) are visited (w_begin a multiple of 64):
// a tile's rows can be split over several waves whose integer partial sums
// add up exactly.
template <int CHUNK>
__device__ __forceinline__ uint32_t sm_tile_accumulate(const uint32_t *A0, const uint32_t *A1, const uint32_t *B0,
                                                       const uint32_t *B1, int64_t w_begin, int64_t w_end,
                                                       const uint32_t *Ycol, int64_t ystride, bool s_ok,
                                                       uint16_t *L, uint32_t *S, int lane, uint32_t (&acc)[4][4]) {
    const int ty = lane >> 3, tx = lane & 7;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0u;
    uint32_t visited = 0;
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
        const int tot = __shfl(incl, 63, 64);
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
        const uint32_t *Ywb = Ycol + wb * 32 * ystride;
        for (int j = 0; j < tot; j += CHUNK) {
            uint32_t v[CHUNK];
#pragma unroll
            for (int k = 0; k < CHUNK; k++) {
                v[k] = 0u;
                if (j + k < tot && s_ok) v[k] = Ywb[(int64_t)L[j + k] * ystride];
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < CHUNK; k++) S[k * 64 + lane] = v[k];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int kn = min(CHUNK, tot - j);
            for (int k = 0; k < kn; k += 2) {
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
                        acc[ii][jj] = acc[ii][jj] + min(av0[ii], bv0[jj]) + min(av1[ii], bv1[jj]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
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
