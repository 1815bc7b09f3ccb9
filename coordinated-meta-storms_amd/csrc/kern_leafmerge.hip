// Non-tree metrics of comp -M 2..5 in the reference's FP32 arithmetic:
// Calc_Dist_Cos, Calc_Dist_E, Calc_Dist_JSD and Calc_Dist_Bray_Curtis
// (dist.h:13-159), similarity = (float)(1.0 - dist) (comp.h:313-326).
//
// The reference loops over all leaves. At a leaf where both samples are zero
// each of its sums gains +0 (cosine, Euclidean, Bray-Curtis: sums that start
// at +0 and only gain non-negative terms are unchanged) or the leaf is skipped
// (JSD), so a lane that merges the two samples' leaf lists in ascending leaf
// order and applies the reference's step at every leaf of the union computes
// the same FP32 values in the same order. Per-sample quantities that the
// reference recomputes inside each call (sum of a, sum of a^2, a / sum a) are
// computed once per sample with the same FP32 operations in leaf order.
// All four are bit-identical to the reference: JSD's logf is the host C
// library's, reproduced on the device from log in double and the host's list
// of inputs where the two differ (host_logf.cpp).
// All code here is built without FMA contraction; divisions and square roots
// are the correctly rounded ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace cms {

namespace {

constexpr uint32_t kEnd = 0xffffffffu;  // past the last leaf of a list (leaf ids are < 2^31)
constexpr int kMergeWavesPerBlock = 4;
constexpr int kMergeK = 32;  // entries of each list staged per window (one per lane of a half wave)
// LDS stride of a staged list: one spare entry after the kMergeK staged ones.
// It holds a sentinel (leaf kEnd), so a merge reads a list's next entry
// without a bounds check, and it shifts list l to bank 2 l, so lanes reading
// the same position of different lists do not conflict.
constexpr int kMergeStride = kMergeK + 1;
// JSD: leaves of a window that both samples of a pair have, queued per wave
// (up to kJsdQueue at a time) so that the wave evaluates their logs with
// every lane busy. The size keeps a JSD workgroup's LDS at 21.5 KB, so 7 fit
// in a CU's 160 KB: the kernel waits on LDS and memory for most of its
// cycles and gains from every wave more than it loses to the extra rounds.
constexpr int kJsdQueue = 96;
// Workgroups per CU that a merge kernel is built for: 8 (8 waves per SIMD, at
// most 64 VGPRs) except JSD's 7 (at most 72 VGPRs; 64 would spill more and
// its LDS would need a smaller queue).
template <int MODE>
constexpr int merge_min_blocks() {
    return MODE == CMS_METRIC_JSD ? 7 : 8;
}

// One thread per sample: sum of a and sum of a^2 in FP32, leaf order
// (dist.h:20-30 and 49-57), and the entry list (leaf, a) or (leaf, a / sum a)
// (dist.h:62-65).
template <bool NORM>
__global__ __launch_bounds__(64) void leaf_merge_prepare_kernel(const int64_t *__restrict__ rowptr,
                                                                const int32_t *__restrict__ col,
                                                                const float *__restrict__ val, int n,
                                                                uint2 *__restrict__ ent, float2 *__restrict__ stat) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= n) return;
    const int64_t b = rowptr[s], e = rowptr[s + 1];
    float sum = 0.0f, sq = 0.0f;
    for (int64_t k = b; k < e; k++) {
        const float v = val[k];
        sum = sum + v;
        sq = sq + v * v;
    }
    for (int64_t k = b; k < e; k++) {
        const float v = val[k];
        ent[k] = make_uint2((uint32_t)col[k], __float_as_uint(NORM ? __fdiv_rn(v, sum) : v));
    }
    stat[s] = make_float2(sum, sq);
}

// The host C library's logf on (0, 4]: table_log(x) rounded to float,
// except for the inputs the host listed (host_logf_table: where its logf
// rounds otherwise, or the double lies too close to a rounding midpoint),
// whose results it gave.
// log(x) in double, the same operations as the host's table_log
// (host_logf.cpp): x = 2^k z with z in [0.699, 1.398) cut into 64 pieces,
// log c_i from the table, log1p(z / c_i - 1) by a degree-8 polynomial.
__device__ __forceinline__ double table_log(float x, const double *invc, const double *logc) {
    uint32_t u = __float_as_uint(x);
    int32_t k0 = 0;
    if (u < 0x00800000u) {  // subnormal: scale into the normal range (exact)
        x *= 0x1p23f;
        u = __float_as_uint(x);
        k0 = -23;
    }
    const uint32_t tmp = u - kLogOff;
    const int i = (int)((tmp >> 17) & (kLogPieces - 1));
    const int32_t k = ((int32_t)tmp >> 23) + k0;
    const double z = (double)__uint_as_float(u - (tmp & 0xff800000u));
    const double r = fma(z, invc[i], -1.0);
    double q = fma(r, -0.125, 1.0 / 7.0);
    q = fma(r, q, -1.0 / 6.0);
    q = fma(r, q, 0.2);
    q = fma(r, q, -0.25);
    q = fma(r, q, 1.0 / 3.0);
    q = fma(r, q, -0.5);
    const double lp = fma(r * r, q, r);
    return fma((double)k, 0x1.62e42fefa39efp-1, logc[i]) + lp;
}

struct HostLogf {
    const uint32_t *bm, *blocks;
    const unsigned long long *tab;
    int ntab;
    float ln2;  // the host's logf(2.0f): a leaf that one sample misses gives x / h = 2
    const double *invc, *logc;  // table_log's pieces (in LDS)
    // The listed input u is in the table.
    __device__ __forceinline__ float listed(uint32_t u) const {
        int lo = 0, hi = ntab - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((uint32_t)(tab[mid] >> 32) < u)
                lo = mid + 1;
            else
                hi = mid;
        }
        return __uint_as_float((uint32_t)tab[lo]);
    }
    __device__ __forceinline__ float operator()(float x) const {
        if (x == 2.0f) return ln2;
        const uint32_t u = __float_as_uint(x), b = u >> 8;
        // outside table_log's domain, the C library's values: logf(+-0) =
        // -inf, logf(+inf) = +inf (x / h is +inf when h rounds to 0 next to
        // a zero), a NaN is returned as it is, and a negative input gives
        // the x86 default NaN
        if ((u & 0x7fffffffu) == 0u) return -__builtin_inff();
        if (u >= 0x7f800000u) return (u <= 0x7fffffffu || u > 0xff800000u) ? x : __uint_as_float(0xffc00000u);
        if (u <= kLogfMaxBits && ((blocks[b >> 5] >> (b & 31u)) & 1u) && ((bm[u >> 5] >> (u & 31u)) & 1u))
            return listed(u);
        return __double2float_rn(table_log(x, invc, logc));
    }
};

// One leaf of the union, p and q being the two samples' values there (zero
// where a sample has no entry).
template <int MODE>
struct MergeAcc {
    HostLogf lg;
    float f = 0.0f, g = 0.0f;  // cos: f; Euclidean: f; Bray-Curtis: sum, diff; JSD: dkl_m, dkl_n
    __device__ __forceinline__ void step(float p, float q) {
        if constexpr (MODE == CMS_METRIC_COS) {
            f = f + p * q;  // dist.h:25
        } else if constexpr (MODE == CMS_METRIC_EU) {
            const float d = p - q;  // pow(d, 2) is exact in double (dist.h:71-72)
            f = __double2float_rn(__dadd_rn((double)f, (double)d * (double)d));
        } else if constexpr (MODE == CMS_METRIC_BC) {
            f = f + (p + q);  // dist.h:146-150
            float d = p - q;
            if (d < 0.0f) d = -d;
            g = g + d;
        }
    }
    // JSD (dist.h:107-118): x log(x / h) for x = p or q, h = (p + q) / 2
    // computed as (p + q) * 0.5 (one rounding of the same value).
    __device__ __forceinline__ float jsd_term(float x, float h) const { return x * lg(__fdiv_rn(x, h)); }
    // One leaf of the union: p's term into f and q's into g, for the values
    // that are positive. A leaf where both are positive has its two terms
    // evaluated already (tq). A leaf where one value is positive and the
    // other zero has h = x / 2 exactly (x >= 2^-125), so x / h = 2 and its
    // term is x * logf(2). A sum without a term at this leaf adds +0, which
    // leaves it unchanged: it starts at +0 and a sum is -0 only when both
    // operands are, so it is never -0.
    __device__ __forceinline__ void jsd_leaf(float p, float q, bool both, float2 tq) {
        const bool pp = p > 0.0f, single = !both && (pp || q > 0.0f);
        const float x = pp ? p : q, o = pp ? q : p;
        float t = x * lg.ln2;
        if (single && !(o == 0.0f && x >= 0x1p-125f)) t = jsd_term(x, (p + q) * 0.5f);
        f = f + (both ? tq.x : (single && pp ? t : 0.0f));
        g = g + (both ? tq.y : (single && !pp ? t : 0.0f));
    }
};

// The reference's distance for a pair whose totals are degenerate, or a
// negative value when the lists must be merged.
template <int MODE>
__device__ __forceinline__ bool degenerate(float2 xa, float2 xb, float &dist) {
    if constexpr (MODE == CMS_METRIC_COS) {
        if (xa.y * xb.y == 0.0f) {  // dist.h:33
            dist = 1.0f;
            return true;
        }
    } else if constexpr (MODE == CMS_METRIC_EU || MODE == CMS_METRIC_BC) {
        if (xa.x <= 0.0f || xb.x <= 0.0f) {  // dist.h:59-60, 138-139
            dist = 1.0f;
            return true;
        }
    } else {
        // JSD has no guard: an all-zero sample divides 0 by 0 at every leaf.
        // Both all-zero: every term is skipped (NaN > 0 is false), dist 0.
        // One all-zero: the other sample's first leaf adds q * log(q / NaN),
        // the x86 default NaN (sign set) that 0/0 gives.
        if (xa.x == 0.0f && xb.x == 0.0f) {
            dist = 0.0f;
            return true;
        }
        if (xa.x == 0.0f || xb.x == 0.0f) {
            dist = __uint_as_float(0xffc00000u);
            return true;
        }
    }
    return false;
}

template <int MODE>
__device__ __forceinline__ float finish(const MergeAcc<MODE> &acc, float2 xa, float2 xb) {
    if constexpr (MODE == CMS_METRIC_COS) {
        return 1.0f - __fdiv_rn(acc.f, sqrtf(xa.y * xb.y));  // dist.h:35-37
    } else if constexpr (MODE == CMS_METRIC_EU) {
        return sqrtf(acc.f);  // dist.h:74
    } else if constexpr (MODE == CMS_METRIC_BC) {
        return acc.f <= 0.0f ? 1.0f : __fdiv_rn(acc.g, acc.f);  // dist.h:152-153
    } else {
        return __double2float_rn(__dsqrt_rn((double)(acc.f + acc.g) / 2.0));  // dist.h:124
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The next leaf m of the union of lists A and B (from positions ia, ib), and
// which lists hold it; false when m is at or past lim. The values are tied
// here so that each entry is one 64-bit LDS read (not a leaf read, then a
// value read after the exit test).
__device__ __forceinline__ bool merge_next(const uint2 *A, const uint2 *B, int ia, int ib, uint32_t lim,
                                           uint32_t &m, bool &ta, bool &tb, uint32_t &va, uint32_t &vb) {
    const uint2 ea = A[ia], eb = B[ib];
    va = ea.y;
    vb = eb.y;
    asm volatile("" : "+v"(va), "+v"(vb));
    m = min(ea.x, eb.x);
    ta = ea.x == m;
    tb = eb.x == m;
    return m < lim;
}

// Lanes below this one in mask.
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// One JSD window of the wave's tile, in rounds. Pass 1 merges every lane's
// two lists towards lim and queues the leaves where both values are positive,
// in wave step order (slot = leaves queued at earlier steps + lanes below at
// this step); the wave stops at the first step whose leaves would not fit.
// Pass 2 evaluates the queued leaves' two terms, 64 at a time. Pass 3 merges
// the same steps again (up to where pass 1 stopped, so with the same slots)
// and adds every leaf's terms to the sums in leaf order. The next round
// resumes pass 1 where it stopped.
__device__ __forceinline__ void jsd_window(const uint2 *A, const uint2 *B, uint32_t lim, bool merge,
                                           MergeAcc<CMS_METRIC_JSD> &acc, float2 *queue, int lane) {
    int ia1 = 0, ib1 = 0, ia3 = 0, ib3 = 0;
    bool done = !merge;  // pass 1 of this lane has reached lim
    for (;;) {
        uint32_t nq = 0;
        bool full = false;
        if (!done) {
            for (;;) {
                uint32_t m, va, vb;
                bool ta, tb;
                if (!merge_next(A, B, ia1, ib1, lim, m, ta, tb, va, vb)) {
                    done = true;
                    break;
                }
                const float p = ta ? __uint_as_float(va) : 0.0f, q = tb ? __uint_as_float(vb) : 0.0f;
                const bool both = p > 0.0f && q > 0.0f;
                const unsigned long long mask = __ballot(both);
                const uint32_t cnt = (uint32_t)__popcll(mask);
                // nq and cnt are the same in every lane still merging
                if (nq + cnt > (uint32_t)kJsdQueue) {
                    full = true;
                    break;
                }
                if (both) queue[nq + lanes_below(mask)] = make_float2(p, q);
                nq += cnt;
                ia1 += ta;
                ib1 += tb;
            }
        }
        // the lane that merged longest counted every queued leaf
        for (int off = 32; off > 0; off >>= 1) nq = max(nq, (uint32_t)__shfl_xor((int)nq, off, 64));
        wave_sync();
        for (uint32_t k = lane; k < nq; k += 64) {
            const float2 pq = queue[k];
            const float h = (pq.x + pq.y) * 0.5f;
            queue[k] = make_float2(acc.jsd_term(pq.x, h), acc.jsd_term(pq.y, h));
        }
        wave_sync();
        if (merge) {
            uint32_t n3 = 0;
            for (;;) {
                if (ia3 == ia1 && ib3 == ib1) break;
                uint32_t m, va, vb;
                bool ta, tb;
                merge_next(A, B, ia3, ib3, lim, m, ta, tb, va, vb);
                const float p = ta ? __uint_as_float(va) : 0.0f, q = tb ? __uint_as_float(vb) : 0.0f;
                const bool both = p > 0.0f && q > 0.0f;
                const unsigned long long mask = __ballot(both);
                const float2 t = both ? queue[n3 + lanes_below(mask)] : make_float2(0.0f, 0.0f);
                acc.jsd_leaf(p, q, both, t);
                n3 += (uint32_t)__popcll(mask);
                ia3 += ta;
                ib3 += tb;
            }
        }
        if (__ballot(full) == 0ull) break;
        wave_sync();  // pass 3's queue reads before the next round's writes
    }
}

// Upper-triangle tiles of 8 x 8 pairs over the visiting order (samples with
// the most entries first, so the lanes of a wave merge lists of similar
// length), one lane per pair, grid-stride over tiles. The tile's 16 entry
// lists (8 row samples, 8 column samples) are staged into LDS kMergeK
// entries at a time with coalesced loads. A window ends below the smallest
// leaf that some list has not staged yet (lim); every lane merges its two
// lists up to lim from LDS, and each list then advances by its staged
// entries below lim, so the windows split the leaf axis and every lane sees
// its pair's union in ascending leaf order. The list whose last staged entry
// set lim advances by kMergeK entries, so the windows always progress; when
// every list is staged to its end, lim is past all leaves and the tile ends.
template <int MODE>
__global__ __launch_bounds__(64 * kMergeWavesPerBlock, merge_min_blocks<MODE>()) void leaf_merge_kernel(LeafMergeData d, int64_t nb,
                                                                              int64_t t_count,
                                                                              float *__restrict__ out) {
    __shared__ uint2 s_ent[kMergeWavesPerBlock][16 * kMergeStride];
    __shared__ int64_t s_pos[kMergeWavesPerBlock][16], s_end[kMergeWavesPerBlock][16];
    __shared__ double s_pieces[MODE == CMS_METRIC_JSD ? 2 * kLogPieces : 1];
    __shared__ float2 s_queue[MODE == CMS_METRIC_JSD ? kMergeWavesPerBlock : 1]
                             [MODE == CMS_METRIC_JSD ? kJsdQueue : 1];
    if constexpr (MODE == CMS_METRIC_JSD) {
        for (int k = threadIdx.x; k < 2 * kLogPieces; k += blockDim.x) s_pieces[k] = d.logf_pieces[k];
        __syncthreads();  // every wave of the block reaches this before its tile loop
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint2 *sbuf = s_ent[wave];
    int64_t *lpos = s_pos[wave], *lend = s_end[wave];
    const int r = lane >> 3, c = lane & 7;
    if (lane < 16) sbuf[lane * kMergeStride + kMergeK] = make_uint2(kEnd, 0u);  // never overwritten
    const int64_t stride = (int64_t)gridDim.x * kMergeWavesPerBlock;
    for (int64_t t = (int64_t)blockIdx.x * kMergeWavesPerBlock + wave; t < t_count; t += stride) {
        const double nn = (double)nb;
        int64_t ra = (int64_t)(((2.0 * nn + 1.0) - sqrt((2.0 * nn + 1.0) * (2.0 * nn + 1.0) - 8.0 * (double)t)) * 0.5);
        if (ra < 0) ra = 0;
        while (ra > 0 && ra * nb - ra * (ra - 1) / 2 > t) ra--;
        while ((ra + 1) * nb - (ra + 1) * ra / 2 <= t) ra++;
        const int64_t cb = ra + (t - (ra * nb - ra * (ra - 1) / 2));
        // list l < 8: visiting position 8 ra + l; l >= 8: 8 cb + l - 8
        if (lane < 16) {
            const int64_t p = lane < 8 ? ra * 8 + lane : cb * 8 + (lane - 8);
            int64_t b = 0, e = 0;
            if (p < d.n) {
                const int s = d.perm[p];
                b = d.rowptr[s];
                e = d.rowptr[s + 1];
            }
            lpos[lane] = b;
            lend[lane] = e;
        }
        wave_sync();
        // this lane's pair: list r (sample a) and list 8 + c (sample b)
        const int64_t pa = ra * 8 + r, pb = cb * 8 + c;
        const bool active = pa < pb && pb < d.n;
        int sa = 0, sb = 0;
        float dist = 0.0f;
        bool merge = false;
        float2 xa = make_float2(0.0f, 0.0f), xb = make_float2(0.0f, 0.0f);
        if (active) {
            sa = d.perm[pa];
            sb = d.perm[pb];
            xa = d.stat[sa];
            xb = d.stat[sb];
            merge = !degenerate<MODE>(xa, xb, dist);
        }
        const uint2 *A = sbuf + r * kMergeStride, *B = sbuf + (8 + c) * kMergeStride;
        MergeAcc<MODE> acc;
        acc.lg = HostLogf{d.logf_bm, d.logf_blocks, d.logf_tab, d.logf_ntab, d.logf_ln2, s_pieces,
                          s_pieces + (MODE == CMS_METRIC_JSD ? kLogPieces : 0)};
        for (;;) {
            // stage: lists 2k (lanes 0-31) and 2k + 1 (lanes 32-63), entry lane & 31
            uint32_t cap = kEnd;
#pragma unroll 4
            for (int k = 0; k < 8; k++) {
                const int l = 2 * k + (lane >> 5), e = lane & 31;
                const int64_t p = lpos[l] + e, pe = lend[l];
                const uint2 v = p < pe ? d.ent[p] : make_uint2(kEnd, 0u);
                sbuf[l * kMergeStride + e] = v;
                if (e == kMergeK - 1 && p + 1 < pe) cap = min(cap, v.x + 1u);
            }
            for (int off = 32; off > 0; off >>= 1) cap = min(cap, (uint32_t)__shfl_xor((int)cap, off, 64));
            const uint32_t lim = cap;
            wave_sync();
            // every list advances by its staged entries below lim (read by
            // the next window's staging, after this window's merge)
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int l = 2 * k + (lane >> 5), e = lane & 31;
                const unsigned long long below = __ballot(sbuf[l * kMergeStride + e].x < lim);
                if (lane == 0) {
                    lpos[2 * k] += __popcll(below & 0xffffffffull);
                    lpos[2 * k + 1] += __popcll(below >> 32);
                }
            }
            wave_sync();
            if constexpr (MODE == CMS_METRIC_JSD) {
                jsd_window(A, B, lim, merge, acc, s_queue[wave], lane);
            } else if (merge) {
                // a list never advances past a sentinel (leaf kEnd >= lim), so
                // ia, ib <= kMergeK
                int ia = 0, ib = 0;
                for (;;) {
                    uint32_t m, va, vb;
                    bool ta, tb;
                    if (!merge_next(A, B, ia, ib, lim, m, ta, tb, va, vb)) break;
                    acc.step(ta ? __uint_as_float(va) : 0.0f, tb ? __uint_as_float(vb) : 0.0f);
                    ia += ta;
                    ib += tb;
                }
            }
            wave_sync();
            if (lim == kEnd) break;
        }
        if (active) {
            if (merge) dist = finish<MODE>(acc, xa, xb);
            if (sa > sb) {
                const int x = sa;
                sa = sb;
                sb = x;
            }
            // an x86 invalid operation gives the default NaN (0xffc00000) and
            // 1.0 - NaN keeps it; the device's arithmetic returns another NaN
            out[(int64_t)sa * d.n + sb - ((int64_t)sa + 1) * ((int64_t)sa + 2) / 2] =
                dist != dist ? __uint_as_float(0xffc00000u) : __double2float_rn(1.0 - (double)dist);
        }
        wave_sync();
    }
}

}  // namespace

hipError_t launch_leaf_merge_prepare(const int64_t *rowptr, const int32_t *col, const float *val, int n,
                                     bool normalise, uint2 *ent, float2 *stat, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 63) / 64);
    if (normalise)
        leaf_merge_prepare_kernel<true><<<blocks, 64, 0, st>>>(rowptr, col, val, n, ent, stat);
    else
        leaf_merge_prepare_kernel<false><<<blocks, 64, 0, st>>>(rowptr, col, val, n, ent, stat);
    return hipGetLastError();
}

hipError_t launch_leaf_merge_all_pairs(const LeafMergeData &d, int mode, float *out, int64_t out_cap,
                                       hipStream_t st) {
    const int64_t n = d.n;
    if (n < 2) return hipSuccess;
    if (out_cap < n * (n - 1) / 2) return hipErrorInvalidValue;
    const int64_t nb = (n + 7) / 8;
    const int64_t ntiles = nb * (nb + 1) / 2;
    const unsigned blocks =
        (unsigned)std::min<int64_t>((ntiles + kMergeWavesPerBlock - 1) / kMergeWavesPerBlock, (int64_t)1 << 20);
    const unsigned threads = 64 * kMergeWavesPerBlock;
    switch (mode) {
    case CMS_METRIC_COS:
        leaf_merge_kernel<CMS_METRIC_COS><<<blocks, threads, 0, st>>>(d, nb, ntiles, out);
        break;
    case CMS_METRIC_EU:
        leaf_merge_kernel<CMS_METRIC_EU><<<blocks, threads, 0, st>>>(d, nb, ntiles, out);
        break;
    case CMS_METRIC_JSD:
        leaf_merge_kernel<CMS_METRIC_JSD><<<blocks, threads, 0, st>>>(d, nb, ntiles, out);
        break;
    case CMS_METRIC_BC:
        leaf_merge_kernel<CMS_METRIC_BC><<<blocks, threads, 0, st>>>(d, nb, ntiles, out);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cms
