// Plan-time tables of the lat-lon output of h; see swe_latlon.hpp.
#include "swe_latlon.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

#include "swe_topology.hpp"

namespace swe {
namespace {

constexpr double kPi = 3.14159265358979323846;

// a linear form over the packed xy field
using Form = std::vector<std::pair<int, double>>;

Form lin(double a, const Form &x, double b, const Form &y)
{
    Form r;
    r.reserve(x.size() + y.size());
    for (auto &e : x) r.emplace_back(e.first, a * e.second);
    for (auto &e : y) r.emplace_back(e.first, b * e.second);
    return r;
}

void scale_add(Form &acc, double a, const Form &x)
{
    for (auto &e : x) acc.emplace_back(e.first, a * e.second);
}

// sorted, duplicates merged, exact zeros kept out
Form tidy(Form f)
{
    std::sort(f.begin(), f.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
        return a.first < b.first;
    });
    Form r;
    for (auto &e : f) {
        if (!r.empty() && r.back().first == e.first) r.back().second += e.second;
        else r.push_back(e);
    }
    r.erase(std::remove_if(r.begin(), r.end(), [](const std::pair<int, double> &e) { return e.second == 0.0; }),
            r.end());
    return r;
}

// cubic Lagrange weights at nodes -1, 0, 1, 2 for position z
void cubic(double z, double w[4])
{
    w[0] = -z * (z - 1.0) * (z - 2.0) / 6.0;
    w[1] = (z + 1.0) * (z - 1.0) * (z - 2.0) / 2.0;
    w[2] = -(z + 1.0) * z * (z - 2.0) / 2.0;
    w[3] = (z + 1.0) * z * (z - 1.0) / 6.0;
}

// one panel's working array, indices -HW1 .. n+HW1 (HW1 = 3: the exchange
// brings the neighbour's rows 0..2 inside the edge)
struct Panel {
    static constexpr int HW1 = 3;
    int n, w;
    std::vector<Form> f;
    explicit Panel(int n_) : n(n_), w(n_ + 2 * HW1), f((size_t)w * w) {}
    Form &at(int i, int j) { return f[(size_t)(j - 1 + HW1) * w + (i - 1 + HW1)]; }
    const Form &at(int i, int j) const { return f[(size_t)(j - 1 + HW1) * w + (i - 1 + HW1)]; }
};

// init_halo_xy_interp (ecs_halo_factory_mod.f90 L230-269): the virtual
// point at along index a, depth j beyond an edge, projected onto the
// neighbour's row j inside; cubic weights over ind-1 .. ind+2
void halo_weights(int n, double hx, int a, int j, double w[4], int &ind)
{
    const double zb = 0.25 * kPi + j * hx, za0 = -0.25 * kPi + (a - 1.0) * hx;
    const double za = std::atan(std::tan(za0) / std::tan(zb));
    ind = std::min(std::max(2, (int)std::floor((za + 0.25 * kPi) / hx) + 1), n - 2);
    cubic((za + 0.25 * kPi) / hx - (ind - 1.0), w);
}

// ext_ecs_tile_halo (ecs_halo_xy_mod.f90 L88-203) on one whole panel: the
// corners' three copies averaged, their diagonals and the cube-edge
// extensions beside them set, then edges bottom, top, left, right: the
// edge's copies averaged and hw virtual rows interpolated along the
// neighbour's rows
void ecs_halo(Panel &F, int n, double hx, int hw)
{
    auto avg3 = [&](int i, int j, int i1, int j1, int i2, int j2) {
        Form s = lin(1.0 / 3.0, F.at(i, j), 1.0 / 3.0, F.at(i1, j1));
        scale_add(s, 1.0 / 3.0, F.at(i2, j2));
        F.at(i, j) = tidy(s);
    };
    auto diag = [&](int di, int dj, int ai, int aj, int bi, int bj) {
        F.at(di, dj) = tidy(lin(0.5, F.at(ai, aj), 0.5, F.at(bi, bj)));
        F.at(ai, aj) = F.at(di, dj);
        F.at(bi, bj) = F.at(di, dj);
    };
    // corners (1,1), (n,1), (n,n), (1,n): copies averaged, then the
    // diagonal virtual points and the cube-edge extensions beside them
    avg3(1, 1, 1, 0, 0, 1);
    for (int j = 1; j <= hw; j++) diag(1 - j, 1 - j, -j, 1, 1, -j);
    avg3(n, 1, n, 0, n + 1, 1);
    for (int j = 1; j <= hw; j++) diag(n + j, 1 - j, n + j + 1, 1, n, -j);
    avg3(n, n, n, n + 1, n + 1, n);
    for (int j = 1; j <= hw; j++) diag(n + j, n + j, n + j + 1, n, n, n + j + 1);
    avg3(1, n, 0, n, 1, n + 1);
    for (int j = 1; j <= hw; j++) diag(1 - j, n + j, 1, n + j + 1, -j, n);
    // edges bottom, top, left, right: (i, j) of along index a at depth
    // k (k = 0: the edge itself, k < 0: inside, k > 0: beyond)
    for (int e = 0; e < 4; e++) {
        auto pt = [&](int a, int k) -> Form & {
            if (e == 0) return F.at(a, 1 - k);
            if (e == 1) return F.at(a, n + k);
            if (e == 2) return F.at(1 - k, a);
            return F.at(n + k, a);
        };
        for (int a = 2; a <= n - 1; a++) pt(a, 0) = tidy(lin(0.5, pt(a, 0), 0.5, pt(a, 1)));
        for (int j = 1; j <= hw; j++) {
            // the neighbour's row j inside sits at depth j + 1
            std::vector<Form> row(n + 1);
            for (int a = 1; a <= n; a++) row[a] = pt(a, j + 1);
            for (int a = 2 - j; a <= n + j - 1; a++) {
                double w[4];
                int ind;
                halo_weights(n, hx, a, j, w, ind);
                Form s;
                for (int m = 0; m < 4; m++) scale_add(s, w[m], row[ind - 1 + m]);
                pt(a, j) = tidy(s);
            }
        }
    }

}

const int EX[6][3] = {{-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0}};
const int EY[6][3] = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
const int NN[6][3] = {{0, 0, 1}, {-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, -1, 0}};

double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// the point of panel p at (alpha, beta) (also beyond the panel, |alpha| <
// pi/2), its covariant basis a1 = dr/dalpha, a2 = dr/dbeta and the dual
// (contravariant) basis b1, b2, all on the unit sphere in the cube's frame
// (metric_2d_ecs_mod.f90: r = (ex X + ey Y - n) / sigma, X = tan alpha)
void panel_frame(int p, double al, double be, double r[3], double a1[3], double a2[3], double b1[3], double b2[3])
{
    const double X = std::tan(al), Y = std::tan(be), sg = std::sqrt(1.0 + X * X + Y * Y);
    double P[3];
    for (int k = 0; k < 3; k++) {
        P[k] = EX[p][k] * X + EY[p][k] * Y - NN[p][k];
        r[k] = P[k] / sg;
    }
    for (int k = 0; k < 3; k++) {
        a1[k] = (1.0 + X * X) * (EX[p][k] / sg - P[k] * X / (sg * sg * sg));
        a2[k] = (1.0 + Y * Y) * (EY[p][k] / sg - P[k] * Y / (sg * sg * sg));
    }
    const double g11 = dot3(a1, a1), g12 = dot3(a1, a2), g22 = dot3(a2, a2), det = g11 * g22 - g12 * g12;
    for (int k = 0; k < 3; k++) {
        b1[k] = (g22 * a1[k] - g12 * a2[k]) / det;
        b2[k] = (-g12 * a1[k] + g11 * a2[k]) / det;
    }
}

int check_sizes(int N, int nlat, int nlon, std::string &err)
{
    if (N < 4 || nlat < 2 || nlon < 1 || (long)nlat * nlon > (1L << 28)) {
        err = "lat-lon grid: need N >= 4, nlat >= 2, nlon >= 1 and nlat * nlon <= 2^28";
        return 1;
    }
    return 0;
}

// extended arrays' rows of panels P appended to t (row r = (p ew + j-1+hw)
// ew + i-1+hw, after the rows already there)
void append_rows(const std::vector<Panel> &P, int n, int hw, LatlonTables &t)
{
    if (t.rp.empty()) t.rp.push_back(0);
    for (int p = 0; p < 6; p++)
        for (int j = 1 - hw; j <= n + hw; j++)
            for (int i = 1 - hw; i <= n + hw; i++) {
                for (auto &e : P[p].at(i, j)) {
                    t.src.push_back(e.first);
                    t.w.push_back(e.second);
                }
                t.rp.push_back((int)t.src.size());
            }
}

// the lat-lon points: panel of the first maximal -n.r, (alpha, beta) by the
// panel's frame, stencil at x_ind = int(zdx); with q, the transform of the
// native covariant components there to zonal / meridional ones
// (init_transform_to_latlon_vec_tile, regrid_factory_mod.f90 L338-382)
int point_table(int N, int nlat, int nlon, int hw, LatlonTables &t, bool with_q, std::string &err)
{
    const int n = N + 1, ew = t.ew;
    const double hx = 0.5 * kPi / N;
    const size_t np = (size_t)nlat * nlon;
    t.base.resize(np);
    t.fx.resize(np);
    t.fy.resize(np);
    if (with_q) t.q.resize(4 * np);
    for (int j = 1; j <= nlat; j++) {
        const double phi = -0.5 * kPi + (j - 1) * kPi / (nlat - 1.0);
        for (int i = 1; i <= nlon; i++) {
            const double lam = (i - 1) * 2.0 * kPi / nlon;
            const double r[3] = {std::cos(phi) * std::cos(lam), std::cos(phi) * std::sin(lam), std::sin(phi)};
            int pan = 0;
            double rn = 0.0;
            for (int q = 0; q < 6; q++) {
                const double d = -(r[0] * NN[q][0] + r[1] * NN[q][1] + r[2] * NN[q][2]);
                if (d > rn) {
                    pan = q;
                    rn = d;
                }
            }
            const double al = std::atan((EX[pan][0] * r[0] + EX[pan][1] * r[1] + EX[pan][2] * r[2]) / rn);
            const double be = std::atan((EY[pan][0] * r[0] + EY[pan][1] * r[1] + EY[pan][2] * r[2]) / rn);
            const double zx = (al + 0.25 * kPi) / hx + 1.0, zy = (be + 0.25 * kPi) / hx + 1.0;
            const int xi = (int)zx, yi = (int)zy;
            if (xi < 0 || yi < 0 || xi > n || yi > n) {
                err = "lat-lon grid: a point outside its panel's extended array";
                return 1;
            }
            const size_t k = (size_t)(j - 1) * nlon + (i - 1);
            t.base[k] = (int)(((size_t)pan * ew + (yi - 1 - 1 + hw)) * ew + (xi - 1 - 1 + hw));
            t.fx[k] = zx - xi;
            t.fy[k] = zy - yi;
            if (with_q) {
                double rr[3], a1[3], a2[3], b1[3], b2[3];
                panel_frame(pan, al, be, rr, a1, a2, b1, b2);
                const double iv[3] = {-std::sin(lam), std::cos(lam), 0.0};
                const double jv[3] = {-std::sin(phi) * std::cos(lam), -std::sin(phi) * std::sin(lam), std::cos(phi)};
                t.q[4 * k + 0] = dot3(iv, b1);
                t.q[4 * k + 1] = dot3(iv, b2);
                t.q[4 * k + 2] = dot3(jv, b1);
                t.q[4 * k + 3] = dot3(jv, b2);
            }
        }
    }
    return 0;
}

}  // namespace

int build_latlon_tables(int N, int nlat, int nlon, LatlonTables &t, std::string &err)
{
    if (check_sizes(N, nlat, nlon, err)) return 1;
    const int n = N + 1, hw = 2;
    const double hx = 0.5 * kPi / N;
    t = LatlonTables();
    t.N = N;
    t.nlat = nlat;
    t.nlon = nlon;
    t.ew = n + 2 * hw;
    std::vector<Panel> P(6, Panel(n));
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++) P[p].at(i, j) = Form{{(p * n + j - 1) * n + i - 1, 1.0}};
    // the xy exchange: halo depth d holds the neighbour's point d-1 inside
    // the shared edge (depth 1: its copy of the edge point)
    for (const HaloEntry &e : build_halo_map(N, 0, hw + 1)) P[e.tp].at(e.ti, e.tj) = P[e.sp].at(e.si, e.sj);
    for (int p = 0; p < 6; p++) ecs_halo(P[p], n, hx, hw);
    append_rows(P, n, hw, t);
    return point_table(N, nlat, nlon, hw, t, false, err);
}

int build_latlon_vec_tables(int N, int nlat, int nlon, LatlonTables &t, std::string &err)
{
    if (check_sizes(N, nlat, nlon, err)) return 1;
    const int n = N + 1, hw = 2;
    const long nxy = 6L * n * n;
    const double hx = 0.5 * kPi / N;
    t = LatlonTables();
    t.N = N;
    t.nlat = nlat;
    t.nlon = nlon;
    t.ew = n + 2 * hw;
    // forms over the colocated components: U at index k, V at nxy + k (k
    // the packed xy index)
    std::vector<Panel> PU(6, Panel(n)), PV(6, Panel(n));
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++) {
                const int k = (p * n + j - 1) * n + i - 1;
                PU[p].at(i, j) = Form{{k, 1.0}};
                PV[p].at(i, j) = Form{{(int)(nxy + k), 1.0}};
            }
    // the vector halo (ecs_halo_vec_xy_mod.f90 L60-130): at every point the
    // exchange brings in (the neighbour's point d-1 inside the edge), the
    // neighbour's native covariant components are re-expressed in this
    // panel's extended coordinates at that point: (u, v)_target =
    // [b1s.a1t, b2s.a1t; b1s.a2t, b2s.a2t] (u, v)_source, b the neighbour's
    // contravariant basis, a this panel's covariant basis there
    for (const HaloEntry &e : build_halo_map(N, 0, hw + 1)) {
        double r[3], a1[3], a2[3], b1s[3], b2s[3], a1t[3], a2t[3], bt1[3], bt2[3], rt[3];
        panel_frame(e.sp, -0.25 * kPi + (e.si - 1) * hx, -0.25 * kPi + (e.sj - 1) * hx, r, a1, a2, b1s, b2s);
        const int p = e.tp;
        const double nr = -(NN[p][0] * r[0] + NN[p][1] * r[1] + NN[p][2] * r[2]);
        const double at = std::atan((EX[p][0] * r[0] + EX[p][1] * r[1] + EX[p][2] * r[2]) / nr);
        const double bt = std::atan((EY[p][0] * r[0] + EY[p][1] * r[1] + EY[p][2] * r[2]) / nr);
        panel_frame(p, at, bt, rt, a1t, a2t, bt1, bt2);
        const double tm[4] = {dot3(b1s, a1t), dot3(b2s, a1t), dot3(b1s, a2t), dot3(b2s, a2t)};
        const Form &us = PU[e.sp].at(e.si, e.sj), &vs = PV[e.sp].at(e.si, e.sj);
        PU[p].at(e.ti, e.tj) = tidy(lin(tm[0], us, tm[1], vs));
        PV[p].at(e.ti, e.tj) = tidy(lin(tm[2], us, tm[3], vs));
    }
    for (int p = 0; p < 6; p++) {
        ecs_halo(PU[p], n, hx, hw);
        ecs_halo(PV[p], n, hx, hw);
    }
    append_rows(PU, n, hw, t);
    append_rows(PV, n, hw, t);
    return point_table(N, nlat, nlon, hw, t, true, err);
}

}  // namespace swe
