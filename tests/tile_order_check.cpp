// Checks the tile order of the persistent all-pairs launches
// (csrc/tile_order.h) on the host: for every launch over whole block rows
// [r0, r1) of small grids, and a few of large ones, the grouped order visits
// every tile of those rows exactly once, stays inside them, and (where every
// band holds at least 256 tiles) a run of 256 consecutive tiles spans at most
// two bands of kTileGroup rows; launches
// that start or end inside a row keep the row-major order. Prints "ok".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tile_order.h"

using namespace cms;

static int fail(const char *what, int64_t nb, int64_t r0, int64_t r1, int64_t t) {
    std::fprintf(stderr, "FAIL %s: nb %lld rows [%lld, %lld) t %lld\n", what, (long long)nb, (long long)r0,
                 (long long)r1, (long long)t);
    return 1;
}

static int check_rows(int64_t nb, int64_t r0, int64_t r1) {
    const int64_t tb = upper_row_start(r0, nb), te = upper_row_start(r1, nb);
    const TileRows g = host_tile_rows(tb, te - tb, nb);
    if (g.r0 != r0 || g.r1 != r1) return fail("host_tile_rows", nb, r0, r1, -1);
    std::vector<unsigned char> seen((size_t)(nb * nb), 0);
    // the run bound holds where every band has at least a run of tiles
    bool runs = true;
    for (int64_t b0 = r0; b0 < r1; b0 += kTileGroup) {
        const int64_t b1 = b0 + kTileGroup < r1 ? b0 + kTileGroup : r1;
        if (upper_row_start(b1, nb) - upper_row_start(b0, nb) < 256) runs = false;
    }
    int64_t run_lo = 0, run_hi = 0;
    for (int64_t t = tb; t < te; t++) {
        int64_t rb, c, ra, cb;
        host_decode_upper(t, nb, rb, c);
        grouped_tile(t, nb, g, rb, ra, cb);
        if (ra < r0 || ra >= r1 || cb < ra || cb >= nb) return fail("tile outside the rows", nb, r0, r1, t);
        if (seen[(size_t)(ra * nb + cb)]++) return fail("tile visited twice", nb, r0, r1, t);
        const int64_t k = (t - tb) % 256;
        if (k == 0 || t == tb) run_lo = run_hi = ra;
        run_lo = ra < run_lo ? ra : run_lo;
        run_hi = ra > run_hi ? ra : run_hi;
        if (runs && run_hi - run_lo >= 2 * kTileGroup) return fail("run spans more than two bands", nb, r0, r1, t);
    }
    for (int64_t a = r0; a < r1; a++)
        for (int64_t b = a; b < nb; b++)
            if (seen[(size_t)(a * nb + b)] != 1) return fail("tile missed", nb, r0, r1, a * nb + b);
    return 0;
}

int main() {
    for (int64_t nb = 1; nb <= 40; nb++)
        for (int64_t r0 = 0; r0 < nb; r0++)
            for (int64_t r1 = r0 + 1; r1 <= nb; r1++)
                if (check_rows(nb, r0, r1)) return 1;
    const int64_t big[][3] = {{313, 0, 313}, {313, 17, 200}, {1000, 0, 1000}, {1000, 993, 1000}, {3125, 40, 99}};
    for (const auto &b : big)
        if (check_rows(b[0], b[1], b[2])) return 1;
    // ranges that start or end inside a row: row-major order
    for (int64_t nb = 2; nb <= 24; nb++) {
        const int64_t total = nb * (nb + 1) / 2;
        for (int64_t tb = 0; tb < total; tb++)
            for (int64_t te = tb + 1; te <= total; te++) {
                int64_t rb, cb, re, ce = 0;
                host_decode_upper(tb, nb, rb, cb);
                re = nb;
                if (te < total) host_decode_upper(te, nb, re, ce);
                const bool whole = cb == rb && (te == total || ce == re);
                const TileRows g = host_tile_rows(tb, te - tb, nb);
                if (whole != (g.r0 >= 0)) return fail("host_tile_rows alignment", nb, tb, te, -1);
                if (!whole) {
                    for (int64_t t = tb; t < te; t++) {
                        int64_t r, c, ra, cc;
                        host_decode_upper(t, nb, r, c);
                        grouped_tile(t, nb, g, r, ra, cc);
                        if (ra != r || cc != c) return fail("row-major fallback", nb, tb, te, t);
                    }
                }
            }
    }
    std::printf("ok\n");
    return 0;
}
