// Order of the upper-triangle block tiles in the persistent all-pairs
// launches. Plain integer code, shared by the kernels (constexpr functions
// are usable on the device) and the host.
#pragma once

#include <cstdint>

namespace cms {

// A persistent launch over whole block rows [r0, r1) takes its tiles column
// by column within bands of kTileGroup block rows, so one XCD's run of
// consecutive tiles (kXcdRun) spans kTileGroup row blocks and
// kXcdRun / kTileGroup column blocks, and both blocks of a tile are shared
// with other waves of the XCD (a row-major run is one row block against 256
// column blocks). Launches that start or end inside a row keep the row-major
// order (r0 = r1 = -1).
constexpr int64_t kTileGroup = 8;
struct TileRows {
    int64_t r0 = -1, r1 = -1;
};

// First tile of block row r in row-major order over an nb x nb grid.
constexpr int64_t upper_row_start(int64_t r, int64_t nb) { return r * nb - r * (r - 1) / 2; }

// Block row rb and column cb of upper-triangle tile t in row-major order, by
// binary search over the rows.
inline void host_decode_upper(int64_t t, int64_t nb, int64_t &rb, int64_t &cb) {
    int64_t lo = 0, hi = nb - 1;  // the largest r with upper_row_start(r) <= t
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (upper_row_start(mid, nb) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    rb = lo;
    cb = lo + (t - upper_row_start(lo, nb));
}

// Tile t (row-major index, whose row-major block row is rb) in the order of a
// launch over whole block rows g: the band of kTileGroup rows [b0, b0 + h)
// holding row rb keeps the band's tiles, taken column by column: first the
// triangle of columns b0 .. b0 + h - 1 (column c holds rows b0 .. c), then h
// rows per column. Row-major (ra = rb) when g.r0 < 0.
constexpr void grouped_tile(int64_t t, int64_t nb, const TileRows &g, int64_t rb, int64_t &ra, int64_t &cb) {
    if (g.r0 < 0) {
        ra = rb;
        cb = rb + (t - upper_row_start(rb, nb));
        return;
    }
    const int64_t b0 = g.r0 + ((rb - g.r0) / kTileGroup) * kTileGroup;
    const int64_t h = (b0 + kTileGroup < g.r1 ? b0 + kTileGroup : g.r1) - b0;
    const int64_t u = t - upper_row_start(b0, nb);  // index inside the band
    const int64_t tri = h * (h + 1) / 2;
    if (u < tri) {
        int64_t k = 0;  // column b0 + k: the largest k with k (k + 1) / 2 <= u (k < kTileGroup)
        while ((k + 1) * (k + 2) / 2 <= u) k++;
        cb = b0 + k;
        ra = b0 + (u - k * (k + 1) / 2);
    } else {
        const int64_t v = u - tri;
        cb = b0 + h + v / h;
        ra = b0 + v % h;
    }
}

// The block rows a launch over tiles [t_begin, t_begin + t_count) of an
// nb x nb grid covers when it holds whole rows (TileRows).
inline TileRows host_tile_rows(int64_t t_begin, int64_t t_count, int64_t nb) {
    TileRows g;
    const int64_t total = nb * (nb + 1) / 2, end = t_begin + t_count;
    if (t_count <= 0 || t_begin < 0 || end > total) return g;
    int64_t rb, cb;
    host_decode_upper(t_begin, nb, rb, cb);
    if (cb != rb) return g;
    int64_t re = nb;
    if (end < total) {
        int64_t ce;
        host_decode_upper(end, nb, re, ce);
        if (ce != re) return g;
    }
    g.r0 = rb;
    g.r1 = re;
    return g;
}

}  // namespace cms
