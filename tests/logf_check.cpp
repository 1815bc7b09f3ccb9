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
    std::printf("listed %lld of %zu table entries, mismatches %lld\n", listed, t.table.size(), bad);
    return bad == 0 ? 0 : 1;
}
