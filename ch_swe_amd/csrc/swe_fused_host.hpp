// Host-side tables of the fused stage kernel: the tiling of each panel, the
// edge segments (one per tile row / column along each panel side) and, for
// every tile, the segments whose band its output region feeds.  A segment's
// records are built from the stage output over the band tile's load region
// (ext_rg of the band [1, 4] or [n-3, n] widened by 6, minimum extent 14),
// so every tile whose output box meets that region is a contributor.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

struct FusedTables {
    int ntx = 0, nty = 0, ntiles = 0, nseg = 0;
    std::vector<int> tile_segs;  // [ntiles][F_MAXSEG], -1 terminated
    std::vector<int> expect;     // contributors per segment
    std::vector<int4> segs;      // (panel, side, lo, hi)
};

// host twin of ext_rg for region 0 (margin rule of the device code, m = 14)
inline void ext_rg_host(int n, int lo, int hi, int &a, int &b)
{
    for (int pass = 0; pass < 2; pass++) {
        lo = lo <= 3 ? 1 : lo;
        hi = hi >= n - 2 ? n : hi;
        if (hi >= n) lo = std::min(lo, n - 13);
        if (lo <= 1) hi = std::max(hi, 14);
    }
    a = std::max(lo, 1);
    b = std::min(hi, n);
}

inline int build_fused_tables(int N, int TY, FusedTables &ft, std::string &err)
{
    const int n = N + 1;
    ft.ntx = (n + F_TX - 1) / F_TX;
    ft.nty = (n + TY - 1) / TY;
    ft.ntiles = 6 * ft.ntx * ft.nty;
    struct Box {
        int x0, x1, y0, y1;
    };
    std::vector<Box> tiles(ft.ntiles);
    for (int p = 0; p < 6; p++)
        for (int ty = 0; ty < ft.nty; ty++)
            for (int tx = 0; tx < ft.ntx; tx++) {
                int i0 = 1 + tx * F_TX, j0 = 1 + ty * TY;
                tiles[(p * ft.nty + ty) * ft.ntx + tx] =
                    Box{i0, std::min(n, i0 + F_TX - 1), j0, std::min(n, j0 + TY - 1)};
            }
    std::vector<Box> bands;
    std::vector<int> segp;
    ft.segs.clear();
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++) {
            int cnt = side <= 1 ? ft.nty : ft.ntx;
            for (int k = 0; k < cnt; k++) {
                int lo = side <= 1 ? 1 + k * TY : 1 + k * F_TX;
                int hi = std::min(n, side <= 1 ? lo + TY - 1 : lo + F_TX - 1);
                ft.segs.push_back(make_int4(p, side, lo, hi));
                int bx0, bx1, by0, by1;
                if (side == 0) { bx0 = 1; bx1 = 4; by0 = lo; by1 = hi; }
                else if (side == 1) { bx0 = n - 3; bx1 = n; by0 = lo; by1 = hi; }
                else if (side == 2) { bx0 = lo; bx1 = hi; by0 = 1; by1 = 4; }
                else { bx0 = lo; bx1 = hi; by0 = n - 3; by1 = n; }
                Box b;
                ext_rg_host(n, bx0 - 6, bx1 + 6, b.x0, b.x1);
                ext_rg_host(n, by0 - 6, by1 + 6, b.y0, b.y1);
                bands.push_back(b);
                segp.push_back(p);
            }
        }
    ft.nseg = (int)ft.segs.size();
    ft.tile_segs.assign((size_t)ft.ntiles * F_MAXSEG, -1);
    ft.expect.assign(ft.nseg, 0);
    for (int t = 0; t < ft.ntiles; t++) {
        int p = t / (ft.ntx * ft.nty), cnt = 0;
        const Box &a = tiles[t];
        for (int k = 0; k < ft.nseg; k++) {
            if (segp[k] != p) continue;
            const Box &b = bands[k];
            if (a.x1 < b.x0 || b.x1 < a.x0 || a.y1 < b.y0 || b.y1 < a.y0) continue;
            if (cnt == F_MAXSEG) {
                err = "internal: a tile touches more than F_MAXSEG edge segments";
                return 1;
            }
            ft.tile_segs[(size_t)t * F_MAXSEG + cnt++] = k;
            ft.expect[k]++;
        }
    }
    return 0;
}
