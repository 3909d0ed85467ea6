// Cube topology and plan-time edge maps (host side).
//
// The six panels and their frames follow topology/cubed_sphere_topology_mod.f90
// L29-55 (the panel numbering and orientation define the state layout, so
// they must match).  Halo contents follow the MPI exchanges of
// parallel/exchange_factory_mod.f90 (xy builder L197-317, Ch-vector builders
// L369-672) and the packing rules of parallel/buffer_mod.f90 L171-475; here
// they are derived combinatorially: each panel side is matched with the
// neighbour side that has the same two cube corners, which fixes the
// neighbour, the along-edge direction and the component signs.
#pragma once

#include <cstdint>
#include <vector>

namespace swe {

enum Side { LEFT = 0, RIGHT = 1, BOTTOM = 2, TOP = 3 };

// Axis types: vertex (N+1 points) or centre (N points).
enum AxisType { VERTEX = 0, CENTRE = 1 };

struct MeshType {
    int xt, yt;
};
constexpr MeshType MESH_XY{VERTEX, VERTEX};
constexpr MeshType MESH_U{CENTRE, VERTEX};  // mesh 'y': covariant u
constexpr MeshType MESH_V{VERTEX, CENTRE};  // mesh 'x': covariant v

inline int axis_points(int N, int t) { return t == VERTEX ? N + 1 : N; }

struct SideMatch {
    int q;         // neighbour panel (0-based)
    int q_side;    // neighbour side touching this one
    bool reversed; // along-edge index runs the other way on the neighbour
};

// neighbour of panel p across side s
SideMatch match_side(int p, int s);

struct HaloEntry {
    int tp, ti, tj, tc;  // target panel, 1-based indices, component (0 u, 1 v)
    int sp, si, sj, sc;  // source
    int sign;
};

// kind 0: xy scalar depth 1; 1: Ch vector (u left/right, v bottom/top) of
// the given depth; 2: xy vector depth 1; 3: the tangential components at the
// duplicated edge points (v left/right, u bottom/top), depth 1.
std::vector<HaloEntry> build_halo_map(int N, int kind, int depth);

}  // namespace swe
