// Check of the .dist value formatter's integer path against snprintf("%.6g")
// (the C library std::ostream formats through): every value 1 - (double)f for
// f a float in [0, 1] on a stride of 29 bit patterns, seeded random doubles log-uniform over
// [1e-15, 1e5), and dyadic rationals m / 2^j (whose decimal expansions end,
// so that exact ties at the seventh digit occur and round half to even).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "cms_internal.h"

int main() {
    long long bad = 0, n = 0;
    // floats in [0, 1]: every 29th bit pattern up to 1.0f
#pragma omp parallel for reduction(+ : bad, n) schedule(static, 1 << 16)
    for (long long w = 0; w <= 0x3f800000ll; w += 29) {
        const uint32_t u = (uint32_t)w;
        float f;
        memcpy(&f, &u, 4);
        const double v = 1.0 - (double)f;
        bad += cms::g6_fast_differs(v) + cms::g6_fast_differs((double)f);
        n += 2;
    }
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(std::log(1e-15), std::log(1e5));
    for (int i = 0; i < 4000000; i++) {
        const double v = std::exp(U(rng));
        bad += cms::g6_fast_differs(v);
        n++;
    }
    for (int j = 1; j <= 40; j++)
        for (uint64_t m = 1; m < 200000; m += (j < 20 ? 1 : 97)) {
            const double v = std::ldexp((double)m, -j);
            bad += cms::g6_fast_differs(v);
            n++;
        }
    std::printf("values %lld, mismatches %lld\n", n, bad);
    return bad == 0 ? 0 : 1;
}
