// Plan-time cube topology and edge maps; see swe_topology.hpp.
#include "swe_topology.hpp"

#include <stdexcept>

namespace swe {
namespace {

// topology/cubed_sphere_topology_mod.f90 L29-55: corner r of point (1,1),
// directions ex / ey of increasing i / j, on the unit cube.
const int EX[6][3] = {{-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0}};
const int EY[6][3] = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
const int R0[6][3] = {{1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 1, 1}};

struct P3 {
    int x, y, z;
    bool operator==(const P3 &o) const { return x == o.x && y == o.y && z == o.z; }
};

P3 corner(int p, int a, int b)
{
    return {R0[p][0] + a * EX[p][0] + b * EY[p][0], R0[p][1] + a * EX[p][1] + b * EY[p][1],
            R0[p][2] + a * EX[p][2] + b * EY[p][2]};
}

// endpoints of side s in the order of increasing along-edge index
void side_ends(int p, int s, P3 &a, P3 &b)
{
    switch (s) {
    case LEFT: a = corner(p, 0, 0); b = corner(p, 0, 1); break;
    case RIGHT: a = corner(p, 1, 0); b = corner(p, 1, 1); break;
    case BOTTOM: a = corner(p, 0, 0); b = corner(p, 1, 0); break;
    default: a = corner(p, 0, 1); b = corner(p, 1, 1); break;
    }
}

}  // namespace

SideMatch match_side(int p, int s)
{
    P3 a, b;
    side_ends(p, s, a, b);
    for (int q = 0; q < 6; q++) {
        if (q == p) continue;
        for (int sq = 0; sq < 4; sq++) {
            P3 c, d;
            side_ends(q, sq, c, d);
            if (c == a && d == b) return {q, sq, false};
            if (c == b && d == a) return {q, sq, true};
        }
    }
    throw std::logic_error("cube side without neighbour");
}

// One halo point: target (p, side, depth d, along index t) on a target mesh
// type; returns the source point and sign for target component tc.
static HaloEntry halo_point(int N, int p, int side, int d, int t, MeshType mt, int tc)
{
    SideMatch m = match_side(p, side);
    bool t_lr = side == LEFT || side == RIGHT;   // target edge runs along y
    bool q_lr = m.q_side == LEFT || m.q_side == RIGHT;
    int along_type = t_lr ? mt.yt : mt.xt;
    int across_type = t_lr ? mt.xt : mt.yt;
    int n_along = axis_points(N, along_type);
    int n_across = axis_points(N, across_type);

    // source mesh: same axis types relative to the shared edge
    MeshType smt = q_lr ? MeshType{across_type, along_type} : MeshType{along_type, across_type};
    int ta = m.reversed ? n_along + 1 - t : t;  // along index on the neighbour
    int da;                                     // across index on the neighbour
    if (m.q_side == LEFT || m.q_side == BOTTOM) da = d;
    else da = n_across + 1 - d;

    HaloEntry e{};
    e.tp = p;
    e.sp = m.q;
    if (side == LEFT) { e.ti = 1 - d; e.tj = t; }
    else if (side == RIGHT) { e.ti = axis_points(N, mt.xt) + d; e.tj = t; }
    else if (side == BOTTOM) { e.ti = t; e.tj = 1 - d; }
    else { e.ti = t; e.tj = axis_points(N, mt.yt) + d; }
    if (q_lr) { e.si = da; e.sj = ta; }
    else { e.si = ta; e.sj = da; }
    (void)smt;

    // components: the target's along-edge component is the neighbour's
    // along-edge component (sign: direction of the along index); the
    // across component compares the target's unfolded axis (into the
    // neighbour on right/top sides) with the neighbour's axis (into the
    // neighbour on its left/bottom sides)
    int t_along_comp = t_lr ? 1 : 0;
    int q_along_comp = q_lr ? 1 : 0;
    e.tc = tc;
    if (tc == t_along_comp) {
        e.sc = q_along_comp;
        e.sign = m.reversed ? -1 : 1;
    } else {
        e.sc = 1 - q_along_comp;
        bool t_into = side == RIGHT || side == TOP;
        bool q_into = m.q_side == LEFT || m.q_side == BOTTOM;
        e.sign = t_into == q_into ? 1 : -1;
    }
    return e;
}

std::vector<HaloEntry> build_halo_map(int N, int kind, int depth)
{
    std::vector<HaloEntry> out;
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++) {
            bool lr = side == LEFT || side == RIGHT;
            if (kind == 3) {
                // tangential components at the duplicated edge points: v (mesh
                // x) across left/right, u (mesh y) across bottom/top, depth 1
                // (the neighbour's copy of the edge point): the f(0), f(n + 2)
                // that the i2c SAT of the curl reads (curl_sbp_sat_mod.f90
                // L26-44 with sbp_diff_i2c_*_mod.f90 add_SAT_correction)
                for (int t = 1; t <= N; t++) out.push_back(halo_point(N, p, side, 1, t, lr ? MESH_V : MESH_U, lr ? 1 : 0));
            } else if (kind == 1) {
                // Ch vector: u (mesh y) across left/right, v (mesh x) across
                // bottom/top; the across axis is the centre-type one
                MeshType mt = lr ? MESH_U : MESH_V;
                int tc = lr ? 0 : 1;
                for (int d = 1; d <= depth; d++)
                    for (int t = 1; t <= N + 1; t++) out.push_back(halo_point(N, p, side, d, t, mt, tc));
            } else {
                for (int d = 1; d <= depth; d++)
                    for (int t = 1; t <= N + 1; t++) {
                        if (kind == 0) {
                            out.push_back(halo_point(N, p, side, d, t, MESH_XY, 0));
                            out.back().tc = 0;
                            out.back().sc = 0;
                            out.back().sign = 1;
                        } else {
                            out.push_back(halo_point(N, p, side, d, t, MESH_XY, 0));
                            out.push_back(halo_point(N, p, side, d, t, MESH_XY, 1));
                        }
                    }
            }
        }
    return out;
}

}  // namespace swe
