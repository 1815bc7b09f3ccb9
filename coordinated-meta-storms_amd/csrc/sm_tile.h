// Pair-tile accumulation of the shared-minimum kernels, shared by the
// all-pairs kernels and the serving cross kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.h"

namespace cms {

// Word 3 of a raw buffer resource on gfx9-family parts (32-bit data format,
// no swizzle, no index stride): buffer loads through it address base +
// offset bytes like a global load.
constexpr int kRawBufferFlags = 0x00020000;

// Row loaders: value of this lane's sample at row rbase + r of Y, with
// window(rbase) called once per 64-word bitmap window and r < 2048. Row
// addresses come from the wave-uniform row index in scalar registers (64-bit
// r * stride, so any stride the allocation allows is addressed correctly),
// and the lane's column is a 32-bit byte offset, so a staged row costs one
// vector load and no vector address arithmetic.
//
// One matrix (all-pairs tiles): lane column `col` (row block sample for lanes
// 0..31, column block sample for lanes 32..63) of Y with row stride `stride`.
// SOFF: the rows are read through one buffer resource per window (its first
// row) with the row's byte offset in the instruction's 32-bit scalar offset,
// which needs every list row (below kSplitList) times the stride times 4 to
// fit in 32 bits (the launcher checks); the tiles then stage no zero row
// (kMaskTail: a short last chunk loads the window's first row in its empty
// slots and clears them). Otherwise each row gets its own resource from a
// 64-bit row address, about ten more scalar instructions per staged row.
// CHECK (tiles at any sample offset, whose rows are those of the union of
// two blocks' bitmaps): a lane whose column lies outside the matrix (ok
// false) loads column 0 and stages zero, and a lane stages zero for rows its
// own block's bitmap (own) lacks, where the embedding stored nothing.
template <bool CHECK, bool SOFF = false>
struct SquareRows {
    static constexpr bool kMaskTail = SOFF;
    const uint32_t *Y;
    int64_t stride;
    uint32_t col4;  // column * 4 (bytes)
    bool ok;
    const uint32_t *own = nullptr;  // CHECK: the bitmap of this lane's block
    const uint32_t *win = nullptr;
    int64_t rb = 0;
    __device__ __forceinline__ void window(int64_t rbase) {
        win = Y + rbase * stride;
        rb = rbase;
    }
    __device__ __forceinline__ uint32_t load(uint32_t r) const {
        // a raw buffer whose base is the row (scalar registers) and whose
        // offset is the lane's column: the address costs scalar
        // instructions only (a global load of the same row would add a 64-bit
        // vector address per load)
        uint32_t v;
        if constexpr (SOFF) {
            const __amdgpu_buffer_rsrc_t w = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<uint32_t *>(win), (short)0, (int)0xffffffff, kRawBufferFlags);
            v = __builtin_amdgcn_raw_buffer_load_b32(w, (int)col4, (int)(r * ((uint32_t)stride * 4u)), 0);
        } else {
            const __amdgpu_buffer_rsrc_t row = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<uint32_t *>(win + (uint64_t)r * (uint64_t)stride), (short)0, (int)0xffffffff,
                kRawBufferFlags);
            v = __builtin_amdgcn_raw_buffer_load_b32(row, (int)col4, 0, 0);
        }
        if constexpr (CHECK) {
            const int64_t e = rb + r;
            return (ok && ((own[e >> 5] >> (e & 31)) & 1u)) ? v : 0u;
        }
        return v;
    }
};

// Two matrices (serving): lanes 0..31 read QY (query slots of one block,
// row stride qs), lanes 32..63 read Y (references of one block, stride ys).
struct CrossRows {
    static constexpr bool kMaskTail = false;
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

// Inclusive prefix sum over the 64 lanes of a wave through DPP moves (row
// shifts by 1, 2, 4 and 8 within each row of 16 lanes, then lane 15 of each
// row added to the next row and lane 31 to the upper half): VALU work and no
// LDS traffic, where a shuffle-based scan takes an LDS permute per step.
__device__ __forceinline__ int wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1 and 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2 and 3
    return x;
}

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
        return (a > b ? a - b : b - a) + c;
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
        const int incl = wave_incl_scan(cnt);
        const int tot = __builtin_amdgcn_readlane(incl, 63);
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

// Split tile (aligned 32 x 32 blocks only): the entries set in both blocks'
// bitmaps split by the blocks' light bitmaps LA, LB (bit set: the block has
// at most kSmRecVals samples with mass at that entry).
//
//   - sparse entries, light in both blocks: at most 3 x 3 pairs with a
//     shared minimum, where staging would read 2 x 32 values of Y. Their
//     records {mask, values} sit in rank order of the light bits (RA[k] for
//     the k-th light entry of block A), so a tile's sparse entries read
//     nearby 16-byte records and never touch Y. One lane per entry adds the
//     minima into the wave's accumulator ACC (LDS atomics);
//   - half entries, light in one block only: one lane per entry reads the
//     other block's 32 values of Y (one 128-byte line) and the light block's
//     record and adds the (at most 3 x 32) minima into ACC: 128 + 16 bytes
//     instead of 256;
//   - staged entries, the rest: the |a - b| path of sm_tile_accumulate.
//
// ACC is 32 x 33 uint32 (row = row-block sample, padded against bank
// conflicts). Windows of kSplitWords bitmap words (kSplitList entries; lanes
// past kSplitWords hold empty words). L (kSplitList uint16) holds the staged
// entries' rows from the front, then the half entries light in A, then those
// light in B, and the sparse entries' rows from the back (disjoint subsets of
// the window's entries); W (kSplitWords uint4) holds each word's light bits
// and the ranks of the light entries before it, from which record positions
// follow. Integer sums:
// the pair's shared sum is (sa + sb - sum |a - b|) / 2 over the staged
// entries plus ACC.
//
// STRIDED: the tile's windows are shared by nparts waves (of a workgroup, or
// separate work units); this one takes windows part, part + nparts, ...
// (rank bases from the light prefixes PLA, PLB: light entries of each block
// before a word), adds into ACC, which the caller zeroes before and reads
// after, and returns only its staged rows' shared sums, exact on their own
// (|a - b| and a + b have the same parity). Otherwise the wave walks every
// window, zeroes ACC and returns the tile's shared sums.
//
// H locates the two blocks' rows for the half entries' heavy side: row r of
// block A at H.A + r * H.as, of block B at H.B + r * H.bs (the all-pairs
// tiles read both from Y; the serving tiles A from the query matrix).
struct SplitCounts {
    // tile totals: entries, sparse entries and their pairs with mass, half
    // entries and the pairs they evaluate (32 per light sample)
    uint32_t visited = 0, sparse = 0, pairs = 0, half = 0, half_pairs = 0;
};

// A split tile's counts (one wave's share) into counter[kCnt*] (if given),
// added once per tile: wave totals over thousands of tiles would not fit 32
// bits.
__device__ __forceinline__ void split_add_counts(unsigned long long *counter, const SplitCounts &sc) {
    if (!counter || !sc.visited) return;
    atomicAdd(counter + kCntVisited, (unsigned long long)sc.visited);
    atomicAdd(counter + kCntStaged, (unsigned long long)(sc.visited - sc.sparse - sc.half));
    if (sc.sparse) {
        atomicAdd(counter + kCntSparseRows, (unsigned long long)sc.sparse);
        atomicAdd(counter + kCntSparsePairs, (unsigned long long)sc.pairs);
    }
    if (sc.half) {
        atomicAdd(counter + kCntHalfRows, (unsigned long long)sc.half);
        atomicAdd(counter + kCntHalfPairs, (unsigned long long)sc.half_pairs);
    }
}

constexpr int kAccStride = 33;
// Split tiles walk the bitmaps in windows of 48 words (1,536 entries): the
// window's entry list, word records, staging rows and accumulator then take
// 10,112 bytes of LDS per wave, so 16 waves fit in a CU's 160 KB (4 per SIMD,
// as many as 128 VGPRs allow).
constexpr int kSplitWords = 48;
constexpr int kSplitList = kSplitWords * 32;

struct HeavyRows {
    const uint32_t *A;  // block A's first column
    int64_t as;         // its row stride (words)
    const uint32_t *B;
    int64_t bs;
};

template <int CHUNK, class Rows, bool STRIDED = false>
__device__ __forceinline__ void sm_tile_split(const uint32_t *A0, const uint32_t *B0, const uint32_t *LA,
                                              const uint32_t *LB, const uint4 *RA, const uint4 *RB, int64_t nwords,
                                              int64_t zrow, const Rows &rows, const HeavyRows &H, uint16_t *L,
                                              uint4 *W, uint32_t *S, uint32_t *ACC, int lane, uint32_t (&acc)[4][4],
                                              SplitCounts &cnt, int part = 0, int nparts = 1,
                                              const uint32_t *PLA = nullptr, const uint32_t *PLB = nullptr) {
    static_assert(CHUNK % 2 == 0, "rows are staged in pairs");
    // words from one of the wave's windows to the next
    const int64_t kStep = STRIDED ? (int64_t)nparts * kSplitWords : (int64_t)kSplitWords;
    const int ty = lane >> 3, tx = lane & 7;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;
    if constexpr (!STRIDED)
        for (int k = lane; k < 32 * kAccStride; k += 64) ACC[k] = 0u;
    // the lane's discard word for the sparse and half loops: the staging
    // buffer, which only the staged loop after them uses
    uint32_t *const trash = S + lane;
    uint32_t cs = 0;                // this lane's column summed over the staged rows
    uint32_t baseA = 0, baseB = 0;  // light entries of each block before the window (ranks)
    // this lane's sparse pairs with mass and half pairs evaluated, summed
    // over the tile's windows and reduced over the wave once at the end (at
    // most 24 entries x 96 pairs per window and lane: far below 2^32)
    uint32_t pairs_lane = 0;
    uint32_t hp_lane = 0;
    // the four bitmap words of the lane in the next window, loaded one window
    // ahead so their latency hides behind the current window's work
    uint32_t na = 0u, nb = 0u, nla = 0u, nlb = 0u;
    const int64_t w0 = STRIDED ? (int64_t)part * kSplitWords : 0;
    if (lane < kSplitWords && w0 + lane < nwords) {
        na = A0[w0 + lane];
        nb = B0[w0 + lane];
        nla = LA[w0 + lane];
        nlb = LB[w0 + lane];
    }
    for (int64_t wb = w0; wb < nwords; wb += kStep) {
        const uint32_t a = na, b = nb, la = nla, lb = nlb;
        const int64_t wn = wb + kStep + lane;
        na = nb = nla = nlb = 0u;
        if (lane < kSplitWords && wn < nwords) {
            na = A0[wn];
            nb = B0[wn];
            nla = LA[wn];
            nlb = LB[wn];
        }
        const uint32_t both = a & b, sw = both & la & lb, hwa = both & la & ~lb, hwb = both & lb & ~la,
                       dw = both & ~(la | lb);
        // 16-bit counts packed two per scan (totals <= 2,048): (staged | half
        // light in A << 16), (light A | light B << 16), (sparse | half light in B << 16)
        const int c1 = __popc(dw) | (__popc(hwa) << 16), c2 = __popc(la) | (__popc(lb) << 16),
                  c3 = __popc(sw) | (__popc(hwb) << 16);
        const int i1 = wave_incl_scan(c1), i2 = wave_incl_scan(c2), i3 = wave_incl_scan(c3);
        const int t1 = __builtin_amdgcn_readlane(i1, 63);
        const int t2 = __builtin_amdgcn_readlane(i2, 63);
        const int t3 = __builtin_amdgcn_readlane(i3, 63);
        const int totd = t1 & 0xffff, tha = t1 >> 16, tots = t3 & 0xffff, thb = t3 >> 16, toth = tha + thb;
        uint32_t winA, winB;
        if constexpr (!STRIDED) {
            winA = baseA;
            winB = baseB;
            baseA += (uint32_t)(t2 & 0xffff);
            baseB += (uint32_t)(t2 >> 16);
        } else {
            winA = PLA[wb];
            winB = PLB[wb];
        }
        if (totd + toth + tots == 0) continue;
        cnt.visited += (uint32_t)(totd + toth + tots);
        cnt.sparse += (uint32_t)tots;
        cnt.half += (uint32_t)toth;
        const int e1 = i1 - c1, e2 = i2 - c2, e3 = i3 - c3;  // this lane's exclusive offsets
        int posd = e1 & 0xffff, posa = totd + (e1 >> 16), posb = totd + tha + (e3 >> 16), poss = e3 & 0xffff;
        __builtin_amdgcn_wave_barrier();
        if ((tots + toth) && lane < kSplitWords) W[lane] = make_uint4(la, lb, (uint32_t)(e2 & 0xffff), (uint32_t)(e2 >> 16));
        for (uint32_t m = dw; m; m &= m - 1u) L[posd++] = (uint16_t)(lane * 32 + __ffs(m) - 1);
        for (uint32_t m = hwa; m; m &= m - 1u) L[posa++] = (uint16_t)(lane * 32 + __ffs(m) - 1);
        for (uint32_t m = hwb; m; m &= m - 1u) L[posb++] = (uint16_t)(lane * 32 + __ffs(m) - 1);
        for (uint32_t m = sw; m; m &= m - 1u) L[kSplitList - 1 - poss++] = (uint16_t)(lane * 32 + __ffs(m) - 1);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // staged entries: as sm_tile_accumulate, from L[0, totd); the first
        // chunk's loads are issued before the sparse and half entries are
        // summed, so their latency is hidden behind that work
        Rows win = rows;
        win.window(wb * 32);
        uint32_t v[CHUNK];
        // slots past the list end stage the zero row (relative to the window)
        const uint32_t zr = (uint32_t)(zrow - wb * 32);
        auto load_chunk = [&](int j) {
#pragma unroll
            for (int k = 0; k < CHUNK; k += 2) {
                const uint32_t pr =
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)*reinterpret_cast<const uint32_t *>(L + j + k));
                const uint32_t zslot = Rows::kMaskTail ? 0u : zr;
                const uint32_t r0 = (j + k < totd) ? (pr & 0xffffu) : zslot;
                const uint32_t r1 = (j + k + 1 < totd) ? (pr >> 16) : zslot;
                v[k] = win.load(r0);
                v[k + 1] = win.load(r1);
            }
        };
        if (totd) load_chunk(0);
        // sparse entries: one per lane, lane l taking entries 64 j + l
        // (neighbouring lanes read neighbouring records, which share lines)
        const int sniter = (tots + 63) >> 6;
        for (int j = 0; j < sniter; j++) {
            const int k = j * 64 + lane;
            if (k < tots) {
                const uint32_t r = L[kSplitList - 1 - k];
                const uint4 wd = W[r >> 5];
                const uint32_t below = (1u << (r & 31u)) - 1u;
                const uint4 qa = RA[winA + wd.z + __popc(wd.x & below)];
                const uint4 qb = RB[winB + wd.w + __popc(wd.y & below)];
                const uint32_t va[3] = {qa.y, qa.z, qa.w}, vb[3] = {qb.y, qb.z, qb.w};
                const int na = __popc(qa.x), nb = __popc(qb.x);
                uint32_t ia[3], ib[3], ma = qa.x, mb = qb.x;
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    ia[t] = (uint32_t)(__ffs(ma) - 1) * kAccStride;
                    ma &= ma - 1u;
                    ib[t] = (uint32_t)(__ffs(mb) - 1);
                    mb &= mb - 1u;
                }
                // every lane adds 9 values; the pairs it does not have go
                // to its own discard word
#pragma unroll
                for (int t = 0; t < 3; t++)
#pragma unroll
                    for (int u = 0; u < 3; u++)
                        atomicAdd((t < na && u < nb) ? ACC + ia[t] + ib[u] : trash, min(va[t], vb[u]));
                pairs_lane += (uint32_t)(na * nb);
            }
        }
        // half entries: one per lane, which reads the heavy block's 32 values
        // of the row (one 128-byte line) and the light block's record; first
        // those light in A (ACC row = the light sample), then those light in
        // B (ACC column = the light sample). Lanes whose entries share a light
        // sample add into the same cells (light samples cluster at the
        // block's densest, first, samples), which serialises the atomics.
        auto half = [&](int first, int count, auto light_b) {
            constexpr bool BL = decltype(light_b)::value;
            constexpr uint32_t kStep = BL ? (uint32_t)kAccStride : 1u;  // ACC words from heavy sample c to c + 1
            // lane l takes entries l * niter + j: the entries of one step
            // come from all over the window (different subtrees, so fewer
            // lanes share a light sample), each lane's from one stretch
            const int niter = (count + 63) >> 6;
            for (int j = 0; j < niter; j++) {
                const int k = lane * niter + j;
                if (k < count) {
                    const uint32_t r = L[first + k];
                    const uint4 wd = W[r >> 5];
                    const uint32_t below = (1u << (r & 31u)) - 1u;
                    const uint4 q = BL ? RB[winB + wd.w + __popc(wd.y & below)] : RA[winA + wd.z + __popc(wd.x & below)];
                    const int64_t hr = wb * 32 + r;
                    const uint4 *hrow = reinterpret_cast<const uint4 *>(BL ? H.A + hr * H.as : H.B + hr * H.bs);
                    uint4 h[8];  // heavy samples 4g .. 4g + 3
#pragma unroll
                    for (int g = 0; g < 8; g++) h[g] = hrow[g];
                    const uint32_t vl[3] = {q.y, q.z, q.w};  // zero past the record's samples
                    const int nl = __popc(q.x);
                    // the record's light samples; slots past nl repeat the
                    // first and are masked off below
                    uint32_t ml = q.x, il[3];
                    il[0] = (uint32_t)(__ffs(ml) - 1);
                    ml &= ml - 1u;
#pragma unroll
                    for (int t = 1; t < 3; t++) {
                        il[t] = ml ? (uint32_t)(__ffs(ml) - 1) : il[0];
                        ml &= ml - 1u;
                    }
                    // the cells are opaque to the optimiser, so the 3 x 32
                    // cells are the three bases plus immediate offsets
                    // instead of registers of precomputed addresses
                    uint32_t ci[3];
#pragma unroll
                    for (int t = 0; t < 3; t++) ci[t] = BL ? il[t] : il[t] * kAccStride;
                    asm volatile("" : "+v"(ci[0]), "+v"(ci[1]), "+v"(ci[2]));
                    const bool l1 = nl > 1, l2 = nl > 2;
                    // per heavy sample with mass: its minimum with each light
                    // sample; lanes whose heavy value is zero sit out (exec
                    // mask), which also keeps them off cells other lanes add
                    // to, and so do lanes past their record's light samples
#pragma unroll
                    for (int g = 0; g < 8; g++) {
                        const uint32_t hq[4] = {h[g].x, h[g].y, h[g].z, h[g].w};
#pragma unroll
                        for (int e = 0; e < 4; e++)
                            if (hq[e]) {
                                atomicAdd(ACC + ci[0] + (g * 4 + e) * kStep, min(vl[0], hq[e]));
                                if (l1) atomicAdd(ACC + ci[1] + (g * 4 + e) * kStep, min(vl[1], hq[e]));
                                if (l2) atomicAdd(ACC + ci[2] + (g * 4 + e) * kStep, min(vl[2], hq[e]));
                            }
                    }
                    hp_lane += 32u * (uint32_t)nl;
                }
            }
        };
        half(totd, tha, std::false_type{});
        half(totd + tha, thb, std::true_type{});
        if (totd == 0) continue;
        for (int j = 0; j < totd; j += CHUNK) {
            // a short last chunk holds zero rows past the list end (|0 - 0|
            // and the column sums gain nothing), so every chunk runs the same
            // unrolled accumulation
            if constexpr (Rows::kMaskTail) {
                if (j + CHUNK > totd) {
#pragma unroll
                    for (int k = 0; k < CHUNK; k++)
                        if (j + k >= totd) v[k] = 0u;
                }
            }
#pragma unroll
            for (int k = 0; k < CHUNK; k++) cs += v[k];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < CHUNK; k++) S[k * 64 + lane] = v[k];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the column sum is complete here, before the next chunk's loads
            // reuse this chunk's registers (otherwise the sum is sunk below
            // the loads and the loaded values are copied back at the end of
            // every chunk)
            asm volatile("" ::"v"(cs));
            if (j + CHUNK < totd) load_chunk(j + CHUNK);
#pragma unroll
            for (int k = 0; k < CHUNK; k += 2) {
                const uint4 a0 = *reinterpret_cast<const uint4 *>(S + k * 64 + ty * 4);
                const uint4 b0 = *reinterpret_cast<const uint4 *>(S + k * 64 + 32 + tx * 4);
                const uint4 a1 = *reinterpret_cast<const uint4 *>(S + (k + 1) * 64 + ty * 4);
                const uint4 b1 = *reinterpret_cast<const uint4 *>(S + (k + 1) * 64 + 32 + tx * 4);
                const uint32_t av0[4] = {a0.x, a0.y, a0.z, a0.w}, bv0[4] = {b0.x, b0.y, b0.z, b0.w};
                const uint32_t av1[4] = {a1.x, a1.y, a1.z, a1.w}, bv1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
                for (int ii = 0; ii < 4; ii++)
#pragma unroll
                    for (int jj = 0; jj < 4; jj++) SmOp::add2(acc[ii][jj], av0[ii], bv0[jj], av1[ii], bv1[jj]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    {
        uint32_t p = pairs_lane, h = hp_lane;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            p += (uint32_t)__shfl_xor((int)p, off, 64);
            h += (uint32_t)__shfl_xor((int)h, off, 64);
        }
        cnt.pairs += (uint32_t)__builtin_amdgcn_readfirstlane((int)p);
        cnt.half_pairs += (uint32_t)__builtin_amdgcn_readfirstlane((int)h);
    }
    uint32_t sa[4], sb[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        sa[i] = (uint32_t)__shfl((int)cs, ty * 4 + i, 64);
        sb[i] = (uint32_t)__shfl((int)cs, 32 + tx * 4 + i, 64);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc[i][j] = SmOp::finish(acc[i][j], sa[i], sb[j]);
            if constexpr (!STRIDED) acc[i][j] += ACC[(ty * 4 + i) * kAccStride + tx * 4 + j];
        }
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
