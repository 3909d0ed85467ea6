// Host-side tables of the fused stage kernel: the tiling of each panel and
// the edge segments (one per tile row / column along each panel side) whose
// records one block of the records kernel builds.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

struct FusedTables {
    int ntx = 0, nty = 0, ntiles = 0, nseg = 0;
    int fwx = 0, fwy = 0;    // width of the first tile along x / y
    std::vector<int4> segs;  // (panel, side, lo, hi)
};

// Tiling of one axis of n points with tiles of width tw: k full tiles between
// two edge tiles of fw >= 8 and n - fw - k tw >= 9 points, so that the full
// tiles qualify for the interior path (9 points clear of both ends) and only
// the edge tiles take the general path.  Edge tiles are at most tw + 8 wide
// (the LDS tile holds tw + 17 points with the 6-point halo).  Axes too short
// for that are tiled uniformly from the start (all tiles general).
inline void tile_axis(int n, int tw, int &nt, int &fw)
{
    const int k = (n - 17) / tw;
    if (n < 17 + tw || k < 1) {
        nt = (n + tw - 1) / tw;
        fw = tw;
        return;
    }
    const int m = n - k * tw;  // 17 <= m < 17 + tw
    nt = k + 2;
    fw = m / 2;
}

inline int build_fused_tables(int N, int TY, FusedTables &ft, std::string &err)
{
    const int n = N + 1;
    tile_axis(n, F_TX, ft.ntx, ft.fwx);
    tile_axis(n, TY, ft.nty, ft.fwy);
    ft.ntiles = 6 * ft.ntx * ft.nty;
    ft.segs.clear();
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++) {
            // records segments: uniform pieces of at most TY (sides 0, 1) or
            // F_TX (sides 2, 3) points, the extents the band tiles fit in LDS
            int cnt = side <= 1 ? (n + TY - 1) / TY : (n + F_TX - 1) / F_TX;
            for (int k = 0; k < cnt; k++) {
                int lo = side <= 1 ? 1 + k * TY : 1 + k * F_TX;
                int hi = std::min(n, side <= 1 ? lo + TY - 1 : lo + F_TX - 1);
                ft.segs.push_back(make_int4(p, side, lo, hi));
            }
        }
    ft.nseg = (int)ft.segs.size();
    (void)err;
    return 0;
}
