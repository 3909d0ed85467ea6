// Host-side tables of the fused stage kernel: the tiling of each panel and
// the edge segments (one per tile row / column along each panel side) whose
// records one block of the records kernel builds.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

struct FusedTables {
    int ntx = 0, nty = 0, ntiles = 0, nseg = 0;
    std::vector<int4> segs;  // (panel, side, lo, hi)
};

inline int build_fused_tables(int N, int TY, FusedTables &ft, std::string &err)
{
    const int n = N + 1;
    ft.ntx = (n + F_TX - 1) / F_TX;
    ft.nty = (n + TY - 1) / TY;
    ft.ntiles = 6 * ft.ntx * ft.nty;
    ft.segs.clear();
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++) {
            int cnt = side <= 1 ? ft.nty : ft.ntx;
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
