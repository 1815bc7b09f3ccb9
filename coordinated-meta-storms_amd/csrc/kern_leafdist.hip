// Non-tree metrics of comp -M 2..5 (dist.h:13-159 via _Comp_Tree::Calc_sim
// mode 2-5, comp.h:296-330): cosine, Euclidean, Jensen-Shannon and
// Bray-Curtis over the leaf abundances, similarity = 1 - dist.
//
// Each is a function of per-sample totals and of one sum over the leaves where
// BOTH samples have mass (I), with p = a / sum(a) stored as float
// (P[leaf][sample]), Pa = sum of a sample's stored p, Qa = sum of p^2 (FP64):
//   cosine      sum_I p q / sqrt(Qa Qb)
//   Euclidean   1 - sqrt(Qa + Qb - 2 sum_I p q)
//   Bray-Curtis 1 - sum|p - q| / sum(p + q) = 2 sum_I min(p, q) / (Pa + Pb)
//   JSD         1 - sqrt(S / 2), S = sum over leaves of p log(2p/(p+q)) +
//               q log(2q/(p+q)) = ln2 (Pa + Pb - sum_I (p + q))
//               + sum_I (p+q)/2 phi(r), r = (p-q)/(p+q),
//               phi(r) = (1+r) ln(1+r) + (1-r) ln(1-r)
// (a leaf with mass on one side only contributes p ln 2). The I-sums use the
// tiles of the shared-minimum kernels over leaf rows (bitmap of leaves with
// mass per 32-sample block) with FP64 accumulators; phi(r), by its series
// for small |r|, keeps the JSD free of cancellation for near-identical
// samples (phi(0) = 0 exactly).
// Degenerate samples follow the reference: an all-zero sample gives 0 for
// cosine, Euclidean and Bray-Curtis; JSD is 1 for two all-zero samples and
// NaN (bits 0xffc00000, as 0/0 gives on the host) when exactly one is.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "kernels.h"
#include "sm_tile.h"

namespace cms {

namespace {

__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }

struct DotOp {  // cosine, Euclidean: sum_I p q
    typedef double Acc;
    static constexpr bool kColSum = false;
    __device__ __forceinline__ static void add2(Acc &acc, uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1) {
        acc = fma((double)as_f(a0), (double)as_f(b0), acc);
        acc = fma((double)as_f(a1), (double)as_f(b1), acc);
    }
};

struct MinOp {  // Bray-Curtis: sum_I min(p, q)
    typedef double Acc;
    static constexpr bool kColSum = false;
    __device__ __forceinline__ static void add2(Acc &acc, uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1) {
        // p >= 0 stored as float: the unsigned order of the bits is the float order
        acc += (double)as_f(min(a0, b0)) + (double)as_f(min(a1, b1));
    }
};

struct JsdOp {  // JSD: sum_I [(p+q)/2 phi(r) - ln2 (p + q)]
    typedef double Acc;
    static constexpr bool kColSum = false;
    // phi(r) = (1+r) ln(1+r) + (1-r) ln(1-r) for p, q > 0, r = (p-q)/(p+q):
    //   |r| <= 1/4: sum_k u^k / (k (2k-1)), u = r^2, six terms (the next is
    //               below 7e-10 relative): no cancellation for near-equal p, q
    //   |r| >  1/4: x ln x + y ln y with x = 2p/(p+q), y = 2q/(p+q) (native
    //               log2; x ln x -> 0 below 1e-30)
    // The ln2 (p + q) part and the sum stay in FP64 (p + q exact there).
    __device__ __forceinline__ static double term(uint32_t ua, uint32_t ub) {
        const float a = as_f(ua), b = as_f(ub);
        const float t = a + b;
        const float rt = __builtin_amdgcn_rcpf(t);
        const float r = (a - b) * rt;
        const float u = r * r;
        float ps = fmaf(u, 1.0f / 66.0f, 1.0f / 45.0f);
        ps = fmaf(u, ps, 1.0f / 28.0f);
        ps = fmaf(u, ps, 1.0f / 15.0f);
        ps = fmaf(u, ps, 1.0f / 6.0f);
        ps = fmaf(u, ps, 1.0f);
        const float x = 2.0f * a * rt, y = 2.0f * b * rt;
        const float xl = x > 1e-30f ? x * __log2f(x) : 0.0f;
        const float yl = y > 1e-30f ? y * __log2f(y) : 0.0f;
        const float phi = u <= 0.0625f ? u * ps : 0.69314718f * (xl + yl);
        const double g = ((double)a + (double)b) * (0.5 * (double)phi - 0.69314718055994530942);
        return (ua != 0u && ub != 0u) ? g : 0.0;
    }
    __device__ __forceinline__ static void add2(Acc &acc, uint32_t a0, uint32_t b0, uint32_t a1, uint32_t b1) {
        acc += term(a0, b0) + term(a1, b1);
    }
};

// One wave per sample: p = a / sum(a) into P, the block bitmap bits, and
// the totals (sum a, sum of stored p, sum of stored p^2).
__global__ __launch_bounds__(64) void leaf_prepare_kernel(const int64_t *__restrict__ rowptr,
                                                          const int32_t *__restrict__ col,
                                                          const float *__restrict__ val, int n, int leaf_n,
                                                          int64_t npad, int64_t nwords, float *__restrict__ P,
                                                          uint32_t *__restrict__ bm, double *__restrict__ stats) {
    const int s = blockIdx.x;
    const int lane = threadIdx.x;
    if (s >= n) return;
    const int64_t b = rowptr[s], e = rowptr[s + 1];
    double sa = 0.0;
    for (int64_t k = b + lane; k < e; k += 64) sa += (double)val[k];
    for (int off = 32; off > 0; off >>= 1) sa += __shfl_xor(sa, off, 64);
    double pa = 0.0, qa = 0.0;
    if (sa > 0.0) {
        for (int64_t k = b + lane; k < e; k += 64) {
            const int32_t leaf = col[k];
            const float v = val[k];
            if (!(v > 0.0f) || leaf < 0 || leaf >= leaf_n) continue;
            const float p = (float)((double)v / sa);
            P[(int64_t)leaf * npad + s] = p;
            pa += (double)p;
            qa += (double)p * (double)p;
            atomicOr(bm + (int64_t)(s / kSmBlock) * nwords + (leaf >> 5), 1u << (leaf & 31));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        pa += __shfl_xor(pa, off, 64);
        qa += __shfl_xor(qa, off, 64);
    }
    if (lane == 0) {
        stats[3 * s] = sa;
        stats[3 * s + 1] = pa;
        stats[3 * s + 2] = qa;
    }
}

template <int MODE>
__device__ __forceinline__ float leaf_finish(const double *st, int64_t a, int64_t b, double acc) {
    const double sa = st[3 * a], pa = st[3 * a + 1], qa = st[3 * a + 2];
    const double sb = st[3 * b], pb = st[3 * b + 1], qb = st[3 * b + 2];
    double sim;
    if constexpr (MODE == CMS_METRIC_COS) {
        sim = (qa * qb == 0.0) ? 0.0 : acc / sqrt(qa * qb);
        sim = sim > 1.0 ? 1.0 : sim;
    } else if constexpr (MODE == CMS_METRIC_EU) {
        sim = (sa <= 0.0 || sb <= 0.0) ? 0.0 : 1.0 - sqrt(fmax(qa + qb - 2.0 * acc, 0.0));
    } else if constexpr (MODE == CMS_METRIC_BC) {
        sim = (sa <= 0.0 || sb <= 0.0) ? 0.0 : 2.0 * acc / (pa + pb);
        sim = sim > 1.0 ? 1.0 : sim;
    } else {
        if (sa == 0.0 && sb == 0.0) {
            sim = 1.0;
        } else if (sa == 0.0 || sb == 0.0) {
            return __uint_as_float(0xffc00000u);  // the NaN of 0/0 on x86 (sign set), as the reference's
        } else {
            const double S = 0.69314718055994530942 * (pa + pb) + acc;
            sim = 1.0 - sqrt(fmax(S, 0.0) / 2.0);
        }
    }
    return (float)sim;
}

// All pairs, persistent over upper-triangle block tiles (as
// sm_persistent_kernel), packed upper triangle in the caller's order.
template <int MODE, class Op, int CHUNK, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void leaf_persistent_kernel(LeafData d, int64_t nb, int64_t t_count,
                                                                     unsigned long long *__restrict__ work,
                                                                     float *__restrict__ out) {
    __shared__ uint16_t s_list[WAVES][64 * 32];
    __shared__ uint32_t s_stage[WAVES][CHUNK * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ty = lane >> 3, tx = lane & 7;
    for (;;) {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(work, 1ull);
        t = __shfl(t, 0, 64);
        if (t >= (unsigned long long)t_count) break;
        // decode the upper-triangle block tile
        const double nn = (double)nb;
        int64_t ra = (int64_t)(((2.0 * nn + 1.0) - sqrt((2.0 * nn + 1.0) * (2.0 * nn + 1.0) - 8.0 * (double)t)) * 0.5);
        if (ra < 0) ra = 0;
        while (ra > 0 && ra * nb - ra * (ra - 1) / 2 > (int64_t)t) ra--;
        while ((ra + 1) * nb - (ra + 1) * ra / 2 <= (int64_t)t) ra++;
        const int64_t cb = ra + ((int64_t)t - (ra * nb - ra * (ra - 1) / 2));
        const int64_t abase = ra * kSmBlock, bbase = cb * kSmBlock;
        const int64_t my_s = (lane < 32) ? abase + lane : bbase + (lane - 32);
        SquareRows<false> rows;
        rows.Y = reinterpret_cast<const uint32_t *>(d.P);
        rows.stride = d.npad;
        rows.ok = true;
        rows.col4 = (uint32_t)my_s * 4u;
        double acc[4][4];
        sm_tile_accumulate<CHUNK, SquareRows<false>, Op>(d.bm + ra * d.nwords, nullptr, d.bm + cb * d.nwords,
                                                          nullptr, 0, d.nwords, rows, s_list[wave], s_stage[wave],
                                                          lane, acc);
#pragma unroll
        for (int ii = 0; ii < 4; ii++)
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
                const int64_t pa = abase + ty * 4 + ii, pb = bbase + tx * 4 + jj;
                if (pa >= pb || pb >= d.n) continue;
                out[pa * (int64_t)d.n + pb - (pa + 1) * (pa + 2) / 2] = leaf_finish<MODE>(d.stats, pa, pb, acc[ii][jj]);
            }
    }
}

}  // namespace

hipError_t launch_leaf_prepare(const int64_t *rowptr, const int32_t *col, const float *val, int n, const LeafData &d,
                               hipStream_t st) {
    if (n <= 0) return hipSuccess;
    leaf_prepare_kernel<<<n, 64, 0, st>>>(rowptr, col, val, n, d.leaf_n, d.npad, d.nwords, const_cast<float *>(d.P),
                                         const_cast<uint32_t *>(d.bm), const_cast<double *>(d.stats));
    return hipGetLastError();
}

hipError_t launch_leaf_all_pairs(const LeafData &d, int mode, unsigned long long *work, float *out, int64_t out_cap,
                                 int grid_waves, hipStream_t st) {
    const int64_t nb = ((int64_t)d.n + kSmBlock - 1) / kSmBlock;
    const int64_t ntiles = nb * (nb + 1) / 2;
    const int64_t n = d.n;
    if (n < 2 || ntiles <= 0) return hipSuccess;
    if (out_cap < n * (n - 1) / 2 || nb * kSmBlock > d.npad) return hipErrorInvalidValue;
    constexpr int CH = 16, WV = 4;
    const unsigned wg = (unsigned)std::max<int64_t>(1, std::min<int64_t>((ntiles + WV - 1) / WV, grid_waves / WV));
    switch (mode) {
    case CMS_METRIC_COS:
        leaf_persistent_kernel<CMS_METRIC_COS, DotOp, CH, WV><<<wg, 64 * WV, 0, st>>>(d, nb, ntiles, work, out);
        break;
    case CMS_METRIC_EU:
        leaf_persistent_kernel<CMS_METRIC_EU, DotOp, CH, WV><<<wg, 64 * WV, 0, st>>>(d, nb, ntiles, work, out);
        break;
    case CMS_METRIC_BC:
        leaf_persistent_kernel<CMS_METRIC_BC, MinOp, CH, WV><<<wg, 64 * WV, 0, st>>>(d, nb, ntiles, work, out);
        break;
    case CMS_METRIC_JSD:
        leaf_persistent_kernel<CMS_METRIC_JSD, JsdOp, CH, WV><<<wg, 64 * WV, 0, st>>>(d, nb, ntiles, work, out);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cms
