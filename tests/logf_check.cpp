// Exhaustive check of the rule the leaf merge kernels use for JSD's logf:
// for every float x in (0, 4], take the host's logf(2) when x == 2, the
// listed result when x is listed (both bitmaps set), and table_log(x) rounded
// to float otherwise; the result must be the C library's logf(x) bit for bit,
// and every listed input must have its block bit set. The device evaluates
// table_log with the same IEEE operations, so this covers the device rule.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "cms_internal.h"

using namespace cms;

int main() {
    const LogfTable &t = host_logf_table();
    long long bad = 0, listed = 0;
#pragma omp parallel for reduction(+ : bad, listed) schedule(static, 1 << 20)
    for (long long v = 1; v <= (long long)kLogfMaxBits; v++) {
        const uint32_t u = (uint32_t)v, b = u >> 8;
        float x, got;
        std::memcpy(&x, &u, 4);
        const bool in_bm = (t.bitmap[u >> 5] >> (u & 31)) & 1u;
        const bool in_blk = (t.blocks[b >> 5] >> (b & 31)) & 1u;
        if (in_bm && !in_blk) bad++;
        if (x == 2.0f) {
            std::memcpy(&got, &t.ln2_bits, 4);
        } else if (in_blk && in_bm) {
            listed++;
            auto it = std::lower_bound(t.table.begin(), t.table.end(), (uint64_t)u << 32);
            if (it == t.table.end() || (uint32_t)(*it >> 32) != u) {
                bad++;
                continue;
            }
            const uint32_t hb = (uint32_t)*it;
            std::memcpy(&got, &hb, 4);
        } else {
            got = (float)table_log(x, t.invc, t.logc);
        }
        const float ref = logf(x);
        if (std::memcmp(&got, &ref, 4) != 0) bad++;
    }
    // inputs outside (0, 4] that the rule answers before any table: +-0 ->
    // -inf, +inf (x / h when h rounds to 0) and NaN unchanged, negative ->
    // the x86 default NaN
    const uint32_t special[] = {0x7f800000u, 0x7fc00000u, 0x7fa00000u, 0x00000000u, 0x80000000u,
                                0xbf800000u, 0x80000001u, 0xff800000u, 0xffc00000u};
    for (uint32_t u : special) {
        float x, got;
        std::memcpy(&x, &u, 4);
        if ((u & 0x7fffffffu) == 0u) {
            got = -INFINITY;
        } else if (u <= 0x7fffffffu || u > 0xff800000u) {
            got = x;
        } else {
            const uint32_t dn = 0xffc00000u;
            std::memcpy(&got, &dn, 4);
        }
        const float ref = logf(x);
        const bool same = std::isnan(ref) ? std::isnan(got) : std::memcmp(&got, &ref, 4) == 0;
        if (!same) {
            bad++;
            uint32_t gb, rb;
            std::memcpy(&gb, &got, 4);
            std::memcpy(&rb, &ref, 4);
            std::printf("special input %08x: rule %08x, logf %08x\n", u, gb, rb);
        }
    }
    std::printf("listed %lld of %zu table entries, mismatches %lld\n", listed, t.table.size(), bad);
    return bad == 0 ? 0 : 1;
}
