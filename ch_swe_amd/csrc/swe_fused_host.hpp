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

// A requested first-tile width fw_req (0: the split above) of an axis with
// edge tiles: between 8 and m - 9, m = the points the full tiles leave.
inline bool tile_axis_req(int n, int tw, int fw_req, int &nt, int &fw, const char *axis, std::string &err)
{
    tile_axis(n, tw, nt, fw);
    if (fw_req == 0) return true;
    const int k = (n - 17) / tw;
    if (n < 17 + tw || k < 1) {
        err = std::string("tile_") + axis + "0: this axis is tiled uniformly (too few points for edge tiles)";
        return false;
    }
    const int m = n - k * tw;
    if (fw_req < 8 || fw_req > m - 9) {
        err = std::string("tile_") + axis + "0 = " + std::to_string(fw_req) + " out of range 8 .. " + std::to_string(m - 9);
        return false;
    }
    fw = fw_req;
    return true;
}

// Every tile of a tiling fits the stage kernel: at most F_THREADS points
// (phase D takes one per thread) and its region 0 within the LDS frame
// (F_TX + 17 by TY + 17 points).
inline bool tiling_fits(int n, int TY, const FusedTables &ft, std::string &err)
{
    for (int ty = 0; ty < ft.nty; ty++)
        for (int tx = 0; tx < ft.ntx; tx++) {
            const Rg sx = tile_span(tx, ft.ntx, ft.fwx, F_TX, n), sy = tile_span(ty, ft.nty, ft.fwy, TY, n);
            const int wx = sx.hi - sx.lo + 1, wy = sy.hi - sy.lo + 1;
            const Rg rx = ext_rg(sx.lo - 6, sx.hi + 6, n, 14), ry = ext_rg(sy.lo - 6, sy.hi + 6, n, 14);
            if (wx * wy > F_THREADS || rx.hi - rx.lo + 1 > F_TX + 17 || ry.hi - ry.lo + 1 > TY + 17) {
                err = "tile of " + std::to_string(wx) + " x " + std::to_string(wy) +
                      " points exceeds the stage kernel's limits (" + std::to_string(F_THREADS) + " points, LDS frame " +
                      std::to_string(F_TX + 17) + " x " + std::to_string(TY + 17) + ")";
                return false;
            }
        }
    return true;
}

inline int build_fused_tables(int N, int TY, FusedTables &ft, std::string &err, int req_x = 0, int req_y = 0)
{
    const int n = N + 1;
    if (!tile_axis_req(n, F_TX, req_x, ft.ntx, ft.fwx, "x", err) || !tile_axis_req(n, TY, req_y, ft.nty, ft.fwy, "y", err))
        return 1;
    ft.ntiles = 6 * ft.ntx * ft.nty;
    ft.segs.clear();
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++) {
            // records segments: uniform pieces of at most R_SEG points, the
            // extents a records frame holds
            const int cnt = (n + R_SEG - 1) / R_SEG;
            for (int k = 0; k < cnt; k++) {
                const int lo = 1 + k * R_SEG, hi = std::min(n, lo + R_SEG - 1);
                ft.segs.push_back(make_int4(p, side, lo, hi));
            }
        }
    ft.nseg = (int)ft.segs.size();
    return tiling_fits(n, TY, ft, err) ? 0 : 1;
}
