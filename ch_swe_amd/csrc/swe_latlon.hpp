// Plan-time tables of the lat-lon output of h (host side).
//
// The shallow-water output writes h on a lat-lon grid (h.dat;
// latlon_outputer_mod.f90 L33-63) by cubic interpolation of the xy field
// after an ECS halo procedure (ecs_halo_xy_mod.f90 ext_ecs_tile_halo L88-203,
// weights ecs_halo_factory_mod.f90 init_halo_xy_interp L230-269): the copies
// of each edge point are averaged, the corners' three copies too, and two
// virtual rows beyond every edge are interpolated along the neighbour's rows.
// That procedure is linear in the field and fixed for a grid, so it is
// evaluated here once, on linear forms, into a sparse map from the field to
// an extended per-panel array (indices -1 .. N+3 on both axes); the
// interpolation itself is described per lat-lon point by the extended-array
// index of its 4 x 4 stencil's first point and its fractional positions.
#pragma once

#include <string>
#include <vector>

namespace swe {

struct LatlonTables {
    int N = 0, nlat = 0, nlon = 0;
    int ew = 0;                  // extended width N + 5 (halo 2 on each side)
    // extended array = map(h): row r (panel-major, (j+1) ew + (i+1)) =
    // sum over rp[r] <= k < rp[r+1] of w[k] h[src[k]], src packed as
    // (p (N+1) + j-1) (N+1) + i-1
    std::vector<int> rp, src;
    std::vector<double> w;
    // per lat-lon point, longitude fastest: extended index of stencil point
    // (x_ind-1, y_ind-1) and the fractional positions zdx - x_ind, zdy - y_ind
    std::vector<int> base;
    std::vector<double> fx, fy;
    // winds only: per point (q11, q12, q21, q22), zonal = q11 u + q12 v,
    // meridional = q21 u + q22 v of the interpolated covariant (u, v)
    std::vector<double> q;
};

// h: returns 0, or 1 with a message (sizes out of range)
int build_latlon_tables(int N, int nlat, int nlon, LatlonTables &t, std::string &err);

// the winds (u.dat, v.dat): src indexes the colocated covariant components
// at the xy points, U then V (6 (N+1)^2 each, packed); the rows hold the U
// component's extended arrays, then the V component's
int build_latlon_vec_tables(int N, int nlat, int nlon, LatlonTables &t, std::string &err);

}  // namespace swe
