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
};

}  // namespace

int build_latlon_tables(int N, int nlat, int nlon, LatlonTables &t, std::string &err)
{
    if (N < 4 || nlat < 2 || nlon < 1 || (long)nlat * nlon > (1L << 28)) {
        err = "lat-lon grid: need N >= 4, nlat >= 2, nlon >= 1 and nlat * nlon <= 2^28";
        return 1;
    }
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
    // ECS weights per (along-edge index a, depth j): the virtual point's
    // position projected onto the neighbour's row j
    auto halo_w = [&](int a, int j, double w[4], int &ind) {
        const double zb = 0.25 * kPi + j * hx, za0 = -0.25 * kPi + (a - 1.0) * hx;
        const double za = std::atan(std::tan(za0) / std::tan(zb));
        ind = std::min(std::max(2, (int)std::floor((za + 0.25 * kPi) / hx) + 1), n - 2);
        cubic((za + 0.25 * kPi) / hx - (ind - 1.0), w);
    };
    for (int p = 0; p < 6; p++) {
        Panel &F = P[p];
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
                    halo_w(a, j, w, ind);
                    Form s;
                    for (int m = 0; m < 4; m++) scale_add(s, w[m], row[ind - 1 + m]);
                    pt(a, j) = tidy(s);
                }
            }
        }
    }
    // the extended arrays' rows
    const int ew = t.ew;
    t.rp.assign((size_t)6 * ew * ew + 1, 0);
    for (int p = 0; p < 6; p++)
        for (int j = 1 - hw; j <= n + hw; j++)
            for (int i = 1 - hw; i <= n + hw; i++) {
                const Form &f = P[p].at(i, j);
                const size_t r = ((size_t)p * ew + (j - 1 + hw)) * ew + (i - 1 + hw);
                for (auto &e : f) {
                    t.src.push_back(e.first);
                    t.w.push_back(e.second);
                }
                t.rp[r + 1] = (int)t.src.size();
            }
    // the lat-lon points: panel of the first maximal -n.r, (alpha, beta)
    // by the panel's frame, stencil at x_ind = int(zdx)
    static const int EX[6][3] = {{-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0}};
    static const int EY[6][3] = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
    static const int NN[6][3] = {{0, 0, 1}, {-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
    const size_t np = (size_t)nlat * nlon;
    t.base.resize(np);
    t.fx.resize(np);
    t.fy.resize(np);
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
        }
    }
    return 0;
}

}  // namespace swe
