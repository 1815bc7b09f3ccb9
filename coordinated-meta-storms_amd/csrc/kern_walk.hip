// Walk kernels: the per-pair post-order Meta-Storms walk of
// _Comp_Tree::Calc_sim (comp.h:332-375), as the reference GPU kernels
// CalcSimOfDiagonalSquare / CalcSimOfNormalRectangle (cms_cuda.cu:52-936)
// compute it, re-designed for gfx950:
//
//  * one wave walks an 8 x 8 tile of pairs (lane = 8 r + c): all 64 lanes
//    execute the same node, so the schedule is read with scalar loads and the
//    leaf/register branches are wave-uniform;
//  * abundances are stored block-leaf-major (8 samples per block, leaves in
//    schedule order), staged per chunk of 32 leaf positions into LDS with a
//    one-chunk register prefetch, so the walk reads LDS only;
//  * the live register slots (re-allocated by liveness on the host, ~10 for
//    GreenGenes-sized trees instead of 150) sit in LDS as two planes (sample
//    a, sample b), one dword per lane at slot * 64 + lane: bank-conflict free;
//  * no FMA contraction (built with -ffp-contract=off) and the reference's
//    order of the two total additions, so F32 results are bit-identical to
//    the CPU `comp -M 0`;
//  * 16-bit storage variants (BF16 / F16 with a per-sample power-of-two
//    scale) unpack to FP32 at staging time and accumulate with Kahan
//    compensation; the final division runs in FP64.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace cms {

typedef const __attribute__((address_space(4))) WalkNode *cnode_ptr;
typedef const __attribute__((address_space(4))) WalkChunk *cchunk_ptr;

template <int STORE>
__device__ __forceinline__ float load_ab(const void *__restrict__ ab, int64_t idx, float inv) {
    if constexpr (STORE == CMS_STORE_F32) {
        return static_cast<const float *>(ab)[idx];
    } else if constexpr (STORE == CMS_STORE_BF16) {
        uint32_t u = static_cast<const uint16_t *>(ab)[idx];
        return __uint_as_float(u << 16) * inv;
    } else {
        _Float16 h = static_cast<const _Float16 *>(ab)[idx];
        return (float)h * inv;
    }
}

// Upper-triangle tile t of an nb x nb block grid (row-major, diagonal
// included) -> (rb, cb), rb <= cb.
__device__ __forceinline__ void decode_upper(int64_t t, int64_t nb, int64_t &rb, int64_t &cb) {
    double nn = (double)nb;
    double disc = (2.0 * nn + 1.0) * (2.0 * nn + 1.0) - 8.0 * (double)t;
    int64_t r = (int64_t)(((2.0 * nn + 1.0) - sqrt(disc)) * 0.5);
    if (r < 0) r = 0;
    // start(r) = r*nb - r*(r-1)/2
    while (r > 0 && r * nb - r * (r - 1) / 2 > t) r--;
    while ((r + 1) * nb - (r + 1) * r / 2 <= t) r++;
    rb = r;
    cb = r + (t - (r * nb - r * (r - 1) / 2));
}

// Workgroup id remapped so that each XCD (workgroups are dealt round-robin
// over the 8 XCDs) receives a contiguous range of tiles, keeping the shared
// row block of neighbouring tiles in that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nwg) {
    const int64_t X = 8;
    int64_t xcd = b % X, idx = b / X, per = nwg / X, rem = nwg % X;
    return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// total / count, clamped to [0, 1] (comp.h:437-440); two samples with no
// present leaf divide 0 by 0, which gives the sign-set NaN on the host.
__device__ __forceinline__ float unweighted_sim(float total, float count) {
    if (count == 0.0f && total == 0.0f) return __uint_as_float(0xffc00000u);
    float t = __fdiv_rn(total, count);
    t = (t > 1.0f) ? 1.0f : t;
    t = (t < 0.0f) ? 0.0f : t;
    return t;
}

// MODE 0: all pairs (upper-triangle tiles, packed output).
// MODE 1: rectangle r0.. x c0.. (row-major output, diagonal = 1).
// VREG: the live register slots (at most kWalkVregSlots) are VGPR arrays
// indexed by the wave-uniform slot number (s_set_gpr_idx: one v_mov per
// access, no LDS); otherwise they sit in LDS as two planes (sample a,
// sample b), one dword per lane at slot * 64 + lane (bank-conflict free).
// UNW: the unweighted walk of Calc_sim_unweight (comp.h:378-446): presence
// flags propagate, and each child edge adds its weight to the shared total
// when both samples are present and to the count when either is; sim =
// total / count, clamped (F32 storage only).
template <int STORE, int MODE, bool VREG, bool UNW = false>
__global__ __launch_bounds__(64 * kWalkWaves) void walk_tile_kernel(WalkData d, int64_t ntiles, int64_t nb,
                                                                     int r0, int nr, int c0, int nc,
                                                                     float *__restrict__ out) {
    extern __shared__ float lds[];
    constexpr bool KAHAN = (STORE != CMS_STORE_F32);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int waves_per_wg = (int)(blockDim.x >> 6);
    const int64_t tile = xcd_remap(blockIdx.x, gridDim.x) * waves_per_wg + wave;
    if (tile >= ntiles) return;  // whole wave leaves; no workgroup barrier is used
    const int per_wave = 16 * kWalkLeafChunk + (VREG ? 0 : 2 * 64 * d.nslots);
    float *stage = lds + (size_t)wave * per_wave;
    float *slotA = stage + 16 * kWalkLeafChunk;
    float *slotB = slotA + 64 * d.nslots;
    float vA[VREG ? kWalkVregSlots : 1], vB[VREG ? kWalkVregSlots : 1];
    if constexpr (VREG) {
#pragma unroll
        for (int k = 0; k < kWalkVregSlots; k++) vA[k] = vB[k] = 0.0f;
    }
    constexpr uint32_t kSlotMask = VREG ? kWalkVregSlots - 1 : 0xffu;
    auto slot_read = [&](uint32_t sl, float &a, float &b) {
        if constexpr (VREG) {
            a = vA[sl & kSlotMask];
            b = vB[sl & kSlotMask];
        } else {
            a = slotA[sl * 64 + lane];
            b = slotB[sl * 64 + lane];
        }
    };
    auto slot_write = [&](uint32_t sl, float a, float b) {
        if constexpr (VREG) {
            vA[sl & kSlotMask] = a;
            vB[sl & kSlotMask] = b;
        } else {
            slotA[sl * 64 + lane] = a;
            slotB[sl * 64 + lane] = b;
        }
    };

    int64_t rowbase, colbase;
    int64_t rowlim, collim;
    if constexpr (MODE == 0) {
        int64_t rb, cb;
        decode_upper(tile, nb, rb, cb);
        rowbase = rb * kWalkBlock;
        colbase = cb * kWalkBlock;
        rowlim = collim = d.n;
    } else {
        int64_t tr = tile / nb, tc = tile % nb;
        rowbase = r0 + tr * kWalkBlock;
        colbase = c0 + tc * kWalkBlock;
        rowlim = r0 + nr;
        collim = c0 + nc;
    }
    const int r = lane >> 3, c = lane & 7;
    const int64_t my_row = rowbase + r, my_col = colbase + c;

    // Staging assignment: lane loads stage[idx] for idx = lane + 64 k,
    // idx = j * 16 + q (leaf position j of the chunk, sample q of the tile).
    int64_t src_off[16 * kWalkLeafChunk / 64];
    float src_inv[16 * kWalkLeafChunk / 64];
    bool src_ok[16 * kWalkLeafChunk / 64];
#pragma unroll
    for (int k = 0; k < 16 * kWalkLeafChunk / 64; k++) {
        int idx = lane + 64 * k, q = idx & 15;
        int64_t s = (q < 8) ? rowbase + q : colbase + (q - 8);
        bool ok = (q < 8) ? (s < rowlim && s < d.n) : (s < collim && s < d.n);
        src_ok[k] = ok;
        s = ok ? s : 0;
        src_off[k] = (s >> 3) * d.nlp * 8 + (s & 7);
        src_inv[k] = (KAHAN && ok) ? d.inv_scale[s] : 1.0f;
    }
    cchunk_ptr chunks = (cchunk_ptr)d.chunks;
    cnode_ptr nodes = (cnode_ptr)d.nodes;

    float pf[16 * kWalkLeafChunk / 64];
    auto fetch = [&](int ch) {
        const int lb = chunks[ch].leaf_begin, le = chunks[ch].leaf_end;
#pragma unroll
        for (int k = 0; k < 16 * kWalkLeafChunk / 64; k++) {
            int idx = lane + 64 * k, j = idx >> 4;
            bool ok = src_ok[k] && (lb + j < le);
            pf[k] = ok ? load_ab<STORE>(d.ab, src_off[k] + (int64_t)(lb + j) * 8, src_inv[k]) : 0.0f;
        }
    };

    float total = 0.0f, comp = 0.0f;
    if (d.nchunks > 0) fetch(0);
    for (int ch = 0; ch < d.nchunks; ch++) {
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < 16 * kWalkLeafChunk / 64; k++) stage[lane + 64 * k] = pf[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int nb_ = chunks[ch].node_begin, ne = chunks[ch].node_end, lb = chunks[ch].leaf_begin;
        if (ch + 1 < d.nchunks) fetch(ch + 1);
        // unrolled: the loop's scalar instructions are shared by several
        // nodes (the walk is scalar-issue bound as much as vector-issue
        // bound); 16-bit storage unrolls less to keep 8 waves per SIMD
        constexpr int kNodeUnroll = KAHAN ? 2 : 4;
#pragma unroll kNodeUnroll
        for (int i = nb_; i < ne; i++) {
            const uint32_t code = nodes[i].code;
            int lp = nodes[i].leafpos - lb;
            const float w1 = nodes[i].w1, w2 = nodes[i].w2;
            float a1, b1, a2, b2;
            if (code & kC1Leaf) {
                a1 = stage[lp * 16 + r];
                b1 = stage[lp * 16 + 8 + c];
                lp++;
            } else {
                slot_read(code & 0xffu, a1, b1);
            }
            if (code & kC2Leaf) {
                a2 = stage[lp * 16 + r];
                b2 = stage[lp * 16 + 8 + c];
            } else {
                slot_read((code >> 8) & 0xffu, a2, b2);
            }
            if constexpr (UNW) {
                // leaf values are abundances, register values presence flags
                const bool p1a = a1 > 0.0f, p1b = b1 > 0.0f, p2a = a2 > 0.0f, p2b = b2 > 0.0f;
                total += (p1a && p1b) ? w1 : 0.0f;
                comp += (p1a || p1b) ? w1 : 0.0f;  // the count of comp.h:382
                total += (p2a && p2b) ? w2 : 0.0f;
                comp += (p2a || p2b) ? w2 : 0.0f;
                if (code & kDstLive)
                    slot_write((code >> 16) & 0xffu, (p1a || p2a) ? 1.0f : 0.0f, (p1b || p2b) ? 1.0f : 0.0f);
                continue;
            }
            const float m1 = fminf(a1, b1), m2 = fminf(a2, b2);
            if constexpr (KAHAN) {
                float y = m1 - comp, t = total + y;
                comp = (t - total) - y;
                total = t;
                y = m2 - comp;
                t = total + y;
                comp = (t - total) - y;
                total = t;
            } else {
                total += m1;
                total += m2;
            }
            if (code & kDstLive) {
                float ra = (a1 - m1) * w1 + (a2 - m2) * w2;
                float rb_ = (b1 - m1) * w1 + (b2 - m2) * w2;
                slot_write((code >> 16) & 0xffu, ra, rb_);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }

    float sim;
    if constexpr (UNW) {
        sim = unweighted_sim(total, comp);
    } else if constexpr (KAHAN) {
        double t = ((double)total - (double)comp) / 100.0;
        t = t > 1.0 ? 1.0 : t;
        t = t < 0.0 ? 0.0 : t;
        sim = (float)t;
    } else {
        float t = total;
        t /= 100.0;  // (float)((double)t / 100.0), comp.h:371
        t = (t > 1.0) ? 1 : t;
        t = (t < 0.0) ? 0 : t;
        sim = t;
    }
    if constexpr (MODE == 0) {
        if (my_row < my_col && my_col < d.n) {
            int64_t p = my_row * (int64_t)d.n + my_col - (my_row + 1) * (my_row + 2) / 2;
            out[p] = sim;
        }
    } else {
        if (my_row < rowlim && my_col < collim)
            out[(my_row - r0) * (int64_t)nc + (my_col - c0)] = (my_row == my_col) ? 1.0f : sim;
    }
}

// Explicit pair list: one lane per pair, leaf values gathered from HBM.
template <int STORE, bool UNW = false>
__global__ __launch_bounds__(256) void walk_pairs_kernel(WalkData d, int64_t n_pairs, const int32_t *__restrict__ prow,
                                                         const int32_t *__restrict__ pcol, float *__restrict__ out) {
    extern __shared__ float lds[];
    constexpr bool KAHAN = (STORE != CMS_STORE_F32);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
    if (i0 + wave * 64 >= n_pairs) return;
    float *slotA = lds + (size_t)wave * 2 * 64 * d.nslots;
    float *slotB = slotA + 64 * d.nslots;
    const int64_t pi = i0 + threadIdx.x;
    const bool ok = pi < n_pairs;
    const int64_t ra = ok ? prow[pi] : 0, rc = ok ? pcol[pi] : 0;
    const int64_t offa = (ra >> 3) * d.nlp * 8 + (ra & 7), offb = (rc >> 3) * d.nlp * 8 + (rc & 7);
    const float inva = KAHAN ? d.inv_scale[ra] : 1.0f, invb = KAHAN ? d.inv_scale[rc] : 1.0f;
    cnode_ptr nodes = (cnode_ptr)d.nodes;
    float total = 0.0f, comp = 0.0f;
    const int order_n = d.order_n;
    for (int i = 0; i < order_n; i++) {
        const uint32_t code = nodes[i].code;
        int64_t lp = nodes[i].leafpos;
        const float w1 = nodes[i].w1, w2 = nodes[i].w2;
        float a1, b1, a2, b2;
        if (code & kC1Leaf) {
            a1 = load_ab<STORE>(d.ab, offa + lp * 8, inva);
            b1 = load_ab<STORE>(d.ab, offb + lp * 8, invb);
            lp++;
        } else {
            a1 = slotA[(code & 0xffu) * 64 + lane];
            b1 = slotB[(code & 0xffu) * 64 + lane];
        }
        if (code & kC2Leaf) {
            a2 = load_ab<STORE>(d.ab, offa + lp * 8, inva);
            b2 = load_ab<STORE>(d.ab, offb + lp * 8, invb);
        } else {
            a2 = slotA[((code >> 8) & 0xffu) * 64 + lane];
            b2 = slotB[((code >> 8) & 0xffu) * 64 + lane];
        }
        if constexpr (UNW) {
            const bool p1a = a1 > 0.0f, p1b = b1 > 0.0f, p2a = a2 > 0.0f, p2b = b2 > 0.0f;
            total += (p1a && p1b) ? w1 : 0.0f;
            comp += (p1a || p1b) ? w1 : 0.0f;
            total += (p2a && p2b) ? w2 : 0.0f;
            comp += (p2a || p2b) ? w2 : 0.0f;
            if (code & kDstLive) {
                slotA[((code >> 16) & 0xffu) * 64 + lane] = (p1a || p2a) ? 1.0f : 0.0f;
                slotB[((code >> 16) & 0xffu) * 64 + lane] = (p1b || p2b) ? 1.0f : 0.0f;
            }
            continue;
        }
        const float m1 = fminf(a1, b1), m2 = fminf(a2, b2);
        if constexpr (KAHAN) {
            float y = m1 - comp, t = total + y;
            comp = (t - total) - y;
            total = t;
            y = m2 - comp;
            t = total + y;
            comp = (t - total) - y;
            total = t;
        } else {
            total += m1;
            total += m2;
        }
        if (code & kDstLive) {
            float xa = (a1 - m1) * w1 + (a2 - m2) * w2;
            float xb = (b1 - m1) * w1 + (b2 - m2) * w2;
            slotA[((code >> 16) & 0xffu) * 64 + lane] = xa;
            slotB[((code >> 16) & 0xffu) * 64 + lane] = xb;
        }
    }
    float sim;
    if constexpr (UNW) {
        sim = unweighted_sim(total, comp);
    } else if constexpr (KAHAN) {
        double t = ((double)total - (double)comp) / 100.0;
        t = t > 1.0 ? 1.0 : t;
        t = t < 0.0 ? 0.0 : t;
        sim = (float)t;
    } else {
        float t = total;
        t /= 100.0;
        t = (t > 1.0) ? 1 : t;
        t = (t < 0.0) ? 0 : t;
        sim = t;
    }
    if (ok) out[pi] = sim;
}

__global__ void sample_scale_kernel(const int64_t *__restrict__ rowptr, const float *__restrict__ val, int n,
                                    int storage, float *__restrict__ scale, float *__restrict__ inv_scale) {
    const int s = blockIdx.x * blockDim.y + threadIdx.y;
    if (s >= n) return;
    float mx = 0.0f;
    for (int64_t k = rowptr[s] + threadIdx.x; k < rowptr[s + 1]; k += 64) mx = fmaxf(mx, val[k]);
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if (threadIdx.x == 0) {
        // power of two e with max * 2^e in [2^14, 2^15): F16 keeps its
        // largest values below 65504 and small ones away from subnormals.
        int e = 0;
        if (mx > 0.0f) e = 14 - ilogbf(mx);
        (void)storage;
        scale[s] = ldexpf(1.0f, e);
        inv_scale[s] = ldexpf(1.0f, -e);
    }
}

__global__ void walk_scatter_kernel(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                    const float *__restrict__ val, const int32_t *__restrict__ leaf_lp_ptr,
                                    const int32_t *__restrict__ leaf_lp, int64_t nlp, int storage,
                                    const float *__restrict__ scale, const int32_t *__restrict__ perm, int p0,
                                    int cnt, void *__restrict__ ab) {
    const int p = blockIdx.x;  // position in the layout
    if (p >= cnt) return;
    const int s = perm ? perm[p0 + p] : p0 + p;  // sample stored there
    const int64_t base = (int64_t)(p >> 3) * nlp * 8 + (p & 7);
    const float sc = (storage == CMS_STORE_F32) ? 1.0f : scale[s];
    for (int64_t k = rowptr[s] + threadIdx.x; k < rowptr[s + 1]; k += blockDim.x) {
        const int32_t leaf = col[k];
        const float v = val[k] * sc;
        for (int32_t t = leaf_lp_ptr[leaf]; t < leaf_lp_ptr[leaf + 1]; t++) {
            const int64_t idx = base + (int64_t)leaf_lp[t] * 8;
            if (storage == CMS_STORE_F32) {
                static_cast<float *>(ab)[idx] = v;
            } else if (storage == CMS_STORE_BF16) {
                uint32_t u = __float_as_uint(v);
                u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
                static_cast<uint16_t *>(ab)[idx] = (uint16_t)(u >> 16);
            } else {
                static_cast<_Float16 *>(ab)[idx] = (_Float16)v;
            }
        }
    }
}

// ------------------------------------------------------------------ launchers

static inline bool walk_vreg(int nslots) { return nslots <= kWalkVregSlots; }

static inline size_t walk_lds_per_wave(int nslots) {
    return (size_t)(16 * kWalkLeafChunk + (walk_vreg(nslots) ? 0 : 2 * 64 * nslots)) * sizeof(float);
}

hipError_t launch_sample_scale(const int64_t *rowptr, const float *val, int n, int storage, float *scale,
                               float *inv_scale, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    dim3 blk(64, 4);
    sample_scale_kernel<<<(n + 3) / 4, blk, 0, st>>>(rowptr, val, n, storage, scale, inv_scale);
    return hipGetLastError();
}

hipError_t launch_walk_scatter(const int64_t *rowptr, const int32_t *col, const float *val,
                               const int32_t *leaf_lp_ptr, const int32_t *leaf_lp, int64_t nlp, int storage,
                               const float *scale, const int32_t *perm, int p0, int cnt, void *ab, hipStream_t st) {
    if (cnt <= 0) return hipSuccess;
    walk_scatter_kernel<<<cnt, 256, 0, st>>>(rowptr, col, val, leaf_lp_ptr, leaf_lp, nlp, storage, scale, perm, p0, cnt,
                                             ab);
    return hipGetLastError();
}

// Waves per workgroup that fit the 64 KiB workgroup LDS limit (0 = none).
static inline int waves_for_lds(size_t per_wave) {
    const size_t limit = 65536;
    int w = kWalkWaves;
    while (w > 0 && per_wave * (size_t)w > limit) w--;
    return w;
}

int walk_max_slots() { return (int)((65536 / sizeof(float) - 16 * kWalkLeafChunk) / 128); }

template <int MODE>
static hipError_t launch_tiles(const WalkData &d, int64_t ntiles, int64_t nb, int r0, int nr, int c0, int nc,
                               float *out, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    const int waves = waves_for_lds(walk_lds_per_wave(d.nslots));
    if (waves == 0) return hipErrorInvalidConfiguration;
    const int64_t nwg = (ntiles + waves - 1) / waves;
    const size_t lds = walk_lds_per_wave(d.nslots) * waves;
    const dim3 grid((unsigned)nwg), block(64 * waves);
#define CMS_WALK_TILES(VR)                                                                                  \
    switch (d.storage) {                                                                                    \
    case CMS_STORE_F32:                                                                                     \
        if (d.unweighted)                                                                                   \
            walk_tile_kernel<CMS_STORE_F32, MODE, VR, true>                                                 \
                <<<grid, block, lds, st>>>(d, ntiles, nb, r0, nr, c0, nc, out);                             \
        else                                                                                                \
            walk_tile_kernel<CMS_STORE_F32, MODE, VR><<<grid, block, lds, st>>>(d, ntiles, nb, r0, nr, c0, nc, out); \
        break;                                                                                              \
    case CMS_STORE_BF16:                                                                                    \
        walk_tile_kernel<CMS_STORE_BF16, MODE, VR><<<grid, block, lds, st>>>(d, ntiles, nb, r0, nr, c0, nc, out); \
        break;                                                                                              \
    default:                                                                                                \
        walk_tile_kernel<CMS_STORE_F16, MODE, VR><<<grid, block, lds, st>>>(d, ntiles, nb, r0, nr, c0, nc, out);  \
        break;                                                                                              \
    }
    if (walk_vreg(d.nslots)) {
        CMS_WALK_TILES(true)
    } else {
        CMS_WALK_TILES(false)
    }
#undef CMS_WALK_TILES
    return hipGetLastError();
}

hipError_t launch_walk_all_pairs(const WalkData &d, float *out, hipStream_t st) {
    const int64_t nb = (d.n + kWalkBlock - 1) / kWalkBlock;
    const int64_t ntiles = nb * (nb + 1) / 2;
    return launch_tiles<0>(d, ntiles, nb, 0, 0, 0, 0, out, st);
}

hipError_t launch_walk_rect(const WalkData &d, int r0, int nr, int c0, int nc, float *out, hipStream_t st) {
    const int64_t nbr = (nr + kWalkBlock - 1) / kWalkBlock, nbc = (nc + kWalkBlock - 1) / kWalkBlock;
    return launch_tiles<1>(d, nbr * nbc, nbc, r0, nr, c0, nc, out, st);
}

hipError_t launch_walk_pairs(const WalkData &d, int64_t n_pairs, const int32_t *row, const int32_t *col, float *out,
                             hipStream_t st) {
    if (n_pairs <= 0) return hipSuccess;
    const int waves = waves_for_lds((size_t)64 * d.nslots * sizeof(float2));
    if (waves == 0) return hipErrorInvalidConfiguration;
    const int64_t nwg = (n_pairs + 64 * waves - 1) / (64 * waves);
    const size_t lds = (size_t)waves * 64 * d.nslots * sizeof(float2);
    const dim3 grid((unsigned)nwg), block(64 * waves);
    switch (d.storage) {
    case CMS_STORE_F32:
        if (d.unweighted)
            walk_pairs_kernel<CMS_STORE_F32, true><<<grid, block, lds, st>>>(d, n_pairs, row, col, out);
        else
            walk_pairs_kernel<CMS_STORE_F32><<<grid, block, lds, st>>>(d, n_pairs, row, col, out);
        break;
    case CMS_STORE_BF16:
        walk_pairs_kernel<CMS_STORE_BF16><<<grid, block, lds, st>>>(d, n_pairs, row, col, out);
        break;
    default:
        walk_pairs_kernel<CMS_STORE_F16><<<grid, block, lds, st>>>(d, n_pairs, row, col, out);
        break;
    }
    return hipGetLastError();
}

}  // namespace cms
