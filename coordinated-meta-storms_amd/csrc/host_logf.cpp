// The host C library's logf on (0, 4], as the leaf merge kernels reproduce it
// for JSD (dist.h:112-116 call log on a float, i.e. the C library's logf).
//
// The device computes log(x) in double (table_log below, the same IEEE
// operations on both sides) rounded to float. That equals the host logf
// except where the host's result is not the correctly rounded one, or where
// log(x) lies so close to a midpoint between two floats that two double
// evaluations could round to different floats. Both sets are found here by
// evaluating every float in (0, 4] once (about 1.1e9 inputs, a few seconds
// over the host's threads) and kept as a bitmap over the input bits plus a
// sorted (input, output) table, so the device answers exactly what this
// process's logf answers. A second bitmap with one bit per 256 consecutive
// inputs (528 KB, about 8 % of its bits set) stays in cache and spares most
// lookups the 135 MB one.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include <omp.h>

#include "cms_internal.h"

namespace cms {

namespace {

// Distance to the midpoint, in units of the float ulp, below which the
// rounding of a double evaluation is not trusted (2^-16 of a float ulp is
// about 8,000 double ulps). The device computes table_log with the same
// operations as here, so this margin only guards against a difference in
// how the two evaluate them.
constexpr double kMidpointMargin = 1.0 / 65536.0;

bool is_exception(uint32_t u, const LogfTable &t, uint32_t &host_bits) {
    float x;
    std::memcpy(&x, &u, 4);
    const float h = logf(x);
    const double d = table_log(x, t.invc, t.logc);
    const float c = (float)d;
    std::memcpy(&host_bits, &h, 4);
    if (std::memcmp(&h, &c, 4) != 0) return true;
    if ((double)c == d) return false;
    const float n = nextafterf(c, d > (double)c ? INFINITY : -INFINITY);
    const double ulp = std::fabs((double)n - (double)c);
    const double mid = 0.5 * ((double)c + (double)n);
    return std::fabs(d - mid) <= ulp * kMidpointMargin;
}

void make_log_pieces(LogfTable &t) {
    for (int i = 0; i < kLogPieces; i++) {
        const uint32_t lo = kLogOff + ((uint32_t)i << 17), hi = lo + (1u << 17);
        float flo, fhi;
        std::memcpy(&flo, &lo, 4);
        std::memcpy(&fhi, &hi, 4);
        if (flo <= 1.0f && 1.0f < fhi) {
            t.invc[i] = 1.0;
            t.logc[i] = 0.0;
        } else {
            const long double c = ((long double)flo + (long double)fhi) / 2;
            t.invc[i] = (double)(1.0L / c);
            t.logc[i] = (double)(-logl((long double)t.invc[i]));
        }
    }
}

}  // namespace

double table_log(float x, const double *invc, const double *logc) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    int32_t k0 = 0;
    if (u < 0x00800000u) {  // subnormal: scale into the normal range (exact)
        x *= 0x1p23f;
        std::memcpy(&u, &x, 4);
        k0 = -23;
    }
    const uint32_t tmp = u - kLogOff;
    const int i = (int)((tmp >> 17) & (kLogPieces - 1));
    const int32_t k = ((int32_t)tmp >> 23) + k0;
    const uint32_t iz = u - (tmp & 0xff800000u);
    float zf;
    std::memcpy(&zf, &iz, 4);
    const double r = std::fma((double)zf, invc[i], -1.0);
    double q = std::fma(r, -0.125, 1.0 / 7.0);
    q = std::fma(r, q, -1.0 / 6.0);
    q = std::fma(r, q, 0.2);
    q = std::fma(r, q, -0.25);
    q = std::fma(r, q, 1.0 / 3.0);
    q = std::fma(r, q, -0.5);
    const double lp = std::fma(r * r, q, r);
    return std::fma((double)k, 0x1.62e42fefa39efp-1, logc[i]) + lp;
}

const LogfTable &host_logf_table() {
    static LogfTable t;
    static std::once_flag once;
    std::call_once(once, [] {
        make_log_pieces(t);
        const uint32_t last = kLogfMaxBits;
        t.bitmap.assign((size_t)(last >> 5) + 1, 0u);
        const int nth = std::max(1, omp_get_max_threads());
        std::vector<std::vector<uint64_t>> found((size_t)nth);
#pragma omp parallel num_threads(nth)
        {
            std::vector<uint64_t> &mine = found[(size_t)omp_get_thread_num()];
#pragma omp for schedule(static, 1 << 20)
            for (int64_t v = 1; v <= (int64_t)last; v++) {
                uint32_t hb;
                if (is_exception((uint32_t)v, t, hb)) mine.push_back(((uint64_t)v << 32) | hb);
            }
        }
        // static schedule: thread k's chunks are ascending and interleaved
        // with the others', so sort the concatenation
        for (auto &f : found) t.table.insert(t.table.end(), f.begin(), f.end());
        std::sort(t.table.begin(), t.table.end());
        t.blocks.assign((size_t)(last >> 13) + 1, 0u);
        for (uint64_t e : t.table) {
            const uint32_t u = (uint32_t)(e >> 32);
            t.bitmap[u >> 5] |= 1u << (u & 31);
            t.blocks[u >> 13] |= 1u << ((u >> 8) & 31);
        }
        const float l2 = logf(2.0f);
        std::memcpy(&t.ln2_bits, &l2, 4);
    });
    return t;
}

}  // namespace cms
