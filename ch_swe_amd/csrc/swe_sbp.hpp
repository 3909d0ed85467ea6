// 1-D staggered SBP stencils evaluated at one output point (device side).
//
// f is a line accessor: f(k) returns the value at 1-based line index k
// (halo indices <= 0 and > n are requested only by SAT terms, where the
// caller provides the neighbour's values); n = number of cell centres on the
// line; k = 1-based output index.  Coefficients follow
//   differential_operators/sbp_operators/sbp_diff_i2c_{21,42}_mod.f90,
//   sbp_diff_c2i_{21,42}_mod.f90, sbp_interp_i2c_{21,42}_mod.f90,
//   sbp_interp_c2i_{21,42}_mod.f90,
// restated as closure tables mirrored at the right end (derivatives change
// sign under the mirror, interpolations do not).  SAT terms (derivatives
// only) read the neighbour's duplicate / normal flux from the halo.
#pragma once

#include <hip/hip_runtime.h>

namespace swe {

struct SBP42 {
    // sbp_interp_i2c_42_mod.f90 L18-19
    static constexpr double c13 = 0.1264926207034815;
    static constexpr double c14 = -0.029892648949074313;
};

#define SWE_HD __host__ __device__ __forceinline__

// strided pointer line: element k at p[k * s] (p points at index 0)
template <typename T>
struct Line {
    const T *p;
    long s;
    SWE_HD T operator()(int k) const { return p[k * s]; }
};
template <typename T>
SWE_HD Line<T> line(const T *p, long s) { return Line<T>{p, s}; }

template <typename A>
SWE_HD auto ld(const A &f, int, int k) -> decltype(f(k)) { return f(k); }

// ---- interior rows (no closure, no SAT); the full stencils below use the
//      same expressions for their interior rows
template <int ORDER, typename T, typename A>
SWE_HD T d_i2c_in(const A &f, int k)
{
    if (ORDER == 21) return f(k + 1) - f(k);
    return (T(27) * (f(k + 1) - f(k)) - (f(k + 2) - f(k - 1))) * T(1.0 / 24);
}
template <int ORDER, typename T, typename A>
SWE_HD T d_c2i_in(const A &f, int k)
{
    if (ORDER == 21) return f(k) - f(k - 1);
    return (T(27) * (f(k) - f(k - 1)) - (f(k + 1) - f(k - 2))) * T(1.0 / 24);
}
template <int ORDER, typename T, typename A>
SWE_HD T i_i2c_in(const A &f, int k)
{
    if (ORDER == 21) return T(0.5) * (f(k + 1) + f(k));
    return (T(9) * (f(k + 1) + f(k)) - (f(k + 2) + f(k - 1))) * T(1.0 / 16);
}
template <int ORDER, typename T, typename A>
SWE_HD T i_c2i_in(const A &f, int k)
{
    if (ORDER == 21) return T(0.5) * (f(k) + f(k - 1));
    return (T(9) * (f(k) + f(k - 1)) - (f(k + 1) + f(k - 2))) * T(1.0 / 16);
}

// ---- interfaces -> centres derivative (+ SAT with duplicates f[0], f[n+2])
template <int ORDER, typename T, typename A>
SWE_HD T d_i2c_pt(const A &f, int s, int n, int k)
{
    if (ORDER == 21) {
        T v = d_i2c_in<21, T>([&](int kk) { return ld(f, s, kk); }, k);
        // 0.5*R/h with R = (3/2, -1/2), h = (1, 1)
        if (k == 1) v += T(0.75) * (ld(f, s, 1) - ld(f, s, 0));
        if (k == 2) v += T(-0.25) * (ld(f, s, 1) - ld(f, s, 0));
        if (k == n) v += T(0.75) * (ld(f, s, n + 2) - ld(f, s, n + 1));
        if (k == n - 1) v += T(-0.25) * (ld(f, s, n + 2) - ld(f, s, n + 1));
        return v;
    } else {
        // closure rows (q11..q35) and SAT weights 0.5*R_r/h_r
        constexpr double Q[3][5] = {{-79.0 / 78, 27.0 / 26, -1.0 / 26, 1.0 / 78, 0},
                                    {2.0 / 21, -9.0 / 7, 9.0 / 7, -2.0 / 21, 0},
                                    {1.0 / 75, 0, -27.0 / 25, 83.0 / 75, -1.0 / 25}};
        constexpr double S[3] = {0.5 * (15.0 / 8) / (13.0 / 12), 0.5 * (-5.0 / 4) / (7.0 / 8),
                                 0.5 * (3.0 / 8) / (25.0 / 24)};
        if (k <= 3) {
            const int r = k - 1;
            T v = T(Q[r][0]) * ld(f, s, 1) + T(Q[r][1]) * ld(f, s, 2) + T(Q[r][2]) * ld(f, s, 3) +
                  T(Q[r][3]) * ld(f, s, 4) + T(Q[r][4]) * ld(f, s, 5);
            v += T(S[r]) * (ld(f, s, 1) - ld(f, s, 0));
            return v;
        }
        if (k >= n - 2) {
            const int r = n - k;
            T v = -(T(Q[r][0]) * ld(f, s, n + 1) + T(Q[r][1]) * ld(f, s, n) + T(Q[r][2]) * ld(f, s, n - 1) +
                    T(Q[r][3]) * ld(f, s, n - 2) + T(Q[r][4]) * ld(f, s, n - 3));
            v += T(S[r]) * (ld(f, s, n + 2) - ld(f, s, n + 1));
            return v;
        }
        return d_i2c_in<42, T>([&](int kk) { return ld(f, s, kk); }, k);
    }
}

// sbp42 centres -> interfaces closure: rows 1..4 (Q), the SAT weight C and
// the extrapolation R of the neighbour's normal flux
struct Sbp42C2i {
    static constexpr double Q[4][5] = {{-2.0, 3.0, -1.0, 0, 0},
                                       {-1.0, 1.0, 0, 0, 0},
                                       {1.0 / 24, -9.0 / 8, 9.0 / 8, -1.0 / 24, 0},
                                       {-1.0 / 71, 6.0 / 71, -83.0 / 71, 81.0 / 71, -3.0 / 71}};
    static constexpr double C = 0.5 / (7.0 / 18);
    static constexpr double R[3] = {15.0 / 8, -5.0 / 4, 3.0 / 8};
};

// ---- centres -> interfaces derivative (+ SAT with neighbour's normal flux
//      at depths 1..ORDER/21+1 stored in f[0], f[-1], ... and f[n+1], ...)
template <int ORDER, typename T, typename A>
SWE_HD T d_c2i_pt(const A &f, int s, int n, int k)
{
    const int nx = n + 1;
    if (ORDER == 21) {
        T v;
        if (k == 1) v = ld(f, s, 2) - ld(f, s, 1);
        else if (k == nx) v = ld(f, s, nx - 1) - ld(f, s, nx - 2);
        else v = d_c2i_in<21, T>([&](int kk) { return ld(f, s, kk); }, k);
        // 0.5/h11 = 1; R = (3/2, -1/2)
        if (k == 1)
            v += T(1.5) * (ld(f, s, 1) - ld(f, s, 0)) - T(0.5) * (ld(f, s, 2) - ld(f, s, -1));
        if (k == nx)
            v -= T(1.5) * (ld(f, s, nx - 1) - ld(f, s, nx)) - T(0.5) * (ld(f, s, nx - 2) - ld(f, s, nx + 1));
        return v;
    } else {
        const auto &Q = Sbp42C2i::Q;
        constexpr double C = Sbp42C2i::C;
        const auto &R = Sbp42C2i::R;
        if (k <= 4) {
            const int r = k - 1;
            T v = T(Q[r][0]) * ld(f, s, 1) + T(Q[r][1]) * ld(f, s, 2) + T(Q[r][2]) * ld(f, s, 3) +
                  T(Q[r][3]) * ld(f, s, 4) + T(Q[r][4]) * ld(f, s, 5);
            if (k == 1)
                v += T(C) * (T(R[0]) * (ld(f, s, 1) - ld(f, s, 0)) + T(R[1]) * (ld(f, s, 2) - ld(f, s, -1)) +
                             T(R[2]) * (ld(f, s, 3) - ld(f, s, -2)));
            return v;
        }
        if (k >= nx - 3) {
            const int r = nx - k;
            T v = -(T(Q[r][0]) * ld(f, s, nx - 1) + T(Q[r][1]) * ld(f, s, nx - 2) + T(Q[r][2]) * ld(f, s, nx - 3) +
                    T(Q[r][3]) * ld(f, s, nx - 4) + T(Q[r][4]) * ld(f, s, nx - 5));
            if (k == nx)
                v -= T(C) * (T(R[0]) * (ld(f, s, nx - 1) - ld(f, s, nx)) +
                             T(R[1]) * (ld(f, s, nx - 2) - ld(f, s, nx + 1)) +
                             T(R[2]) * (ld(f, s, nx - 3) - ld(f, s, nx + 2)));
            return v;
        }
        return d_c2i_in<42, T>([&](int kk) { return ld(f, s, kk); }, k);
    }
}

// d_c2i_pt at row k on a window of the line held in registers: w[j] = f(ws +
// j), ws = c2i_window_start<ORDER>(k, n).  sbp21: 4 values (rows 1 and n + 1
// read the two end points and two SAT values, inner rows k - 1, k); sbp42: 8
// values (closure rows read the first / last five points and three SAT
// values, inner rows k - 2 .. k + 1).  The same arithmetic as d_c2i_pt, with
// a constant window index in every branch.
template <int ORDER>
SWE_HD int c2i_window_start(int k, int n)
{
    const int nx = n + 1;
    if (ORDER == 21) return k == 1 ? -1 : (k == nx ? nx - 2 : k - 1);
    return k <= 4 ? -2 : (k >= nx - 3 ? nx - 5 : k - 2);
}
template <int ORDER, typename T>
SWE_HD T d_c2i_win(const T *w, int k, int n)
{
    const int nx = n + 1;
    if (ORDER == 21) {
        if (k == 1) {
            T v = w[3] - w[2];
            v += T(1.5) * (w[2] - w[1]) - T(0.5) * (w[3] - w[0]);
            return v;
        }
        if (k == nx) {
            T v = w[1] - w[0];
            v -= T(1.5) * (w[1] - w[2]) - T(0.5) * (w[0] - w[3]);
            return v;
        }
        return w[1] - w[0];
    } else {
        const auto &Q = Sbp42C2i::Q;
        constexpr double C = Sbp42C2i::C;
        const auto &R = Sbp42C2i::R;
        if (k <= 4) {  // w[j] = f(j - 2)
            const int r = k - 1;
            T v = T(Q[r][0]) * w[3] + T(Q[r][1]) * w[4] + T(Q[r][2]) * w[5] + T(Q[r][3]) * w[6] + T(Q[r][4]) * w[7];
            if (k == 1) v += T(C) * (T(R[0]) * (w[3] - w[2]) + T(R[1]) * (w[4] - w[1]) + T(R[2]) * (w[5] - w[0]));
            return v;
        }
        if (k >= nx - 3) {  // w[j] = f(nx - 5 + j)
            const int r = nx - k;
            T v = -(T(Q[r][0]) * w[4] + T(Q[r][1]) * w[3] + T(Q[r][2]) * w[2] + T(Q[r][3]) * w[1] +
                    T(Q[r][4]) * w[0]);
            if (k == nx) v -= T(C) * (T(R[0]) * (w[4] - w[5]) + T(R[1]) * (w[3] - w[6]) + T(R[2]) * (w[2] - w[7]));
            return v;
        }
        // w[j] = f(k - 2 + j)
        return (T(27) * (w[2] - w[1]) - (w[3] - w[0])) * T(1.0 / 24);
    }
}

// ---- interfaces -> centres interpolation (no halo reads)
template <int ORDER, typename T, typename A>
SWE_HD T i_i2c_pt(const A &f, int s, int n, int k)
{
    if (ORDER == 21) return i_i2c_in<21, T>([&](int kk) { return ld(f, s, kk); }, k);
    constexpr double c13 = SBP42::c13, c14 = SBP42::c14;
    constexpr double Q[3][5] = {
        {0.5 + c13 + 2 * c14, 0.5 - 2 * c13 - 3 * c14, c13, c14, 0},
        {-8.0 / 63 - 52.0 / 21 * c13 - 104.0 / 21 * c14, 29.0 / 42 + 104.0 / 21 * c13 + 52.0 / 7 * c14,
         -52.0 / 21 * c13 + 0.5, -52.0 / 21 * c14 - 4.0 / 63, 0},
        {26.0 / 25 * c13 + 52.0 / 25 * c14 - 1.0 / 25, -1.0 / 50 - 52.0 / 25 * c13 - 78.0 / 25 * c14,
         3.0 / 5 + 26.0 / 25 * c13, 13.0 / 25 + 26.0 / 25 * c14, -3.0 / 50}};
    if (k <= 3) {
        const int r = k - 1;
        return T(Q[r][0]) * ld(f, s, 1) + T(Q[r][1]) * ld(f, s, 2) + T(Q[r][2]) * ld(f, s, 3) +
               T(Q[r][3]) * ld(f, s, 4) + T(Q[r][4]) * ld(f, s, 5);
    }
    if (k >= n - 2) {
        const int r = n - k;
        return T(Q[r][0]) * ld(f, s, n + 1) + T(Q[r][1]) * ld(f, s, n) + T(Q[r][2]) * ld(f, s, n - 1) +
               T(Q[r][3]) * ld(f, s, n - 2) + T(Q[r][4]) * ld(f, s, n - 3);
    }
    return i_i2c_in<42, T>([&](int kk) { return ld(f, s, kk); }, k);
}

// ---- centres -> interfaces interpolation (no halo reads)
template <int ORDER, typename T, typename A>
SWE_HD T i_c2i_pt(const A &f, int s, int n, int k)
{
    const int nx = n + 1;
    if (ORDER == 21) {
        if (k == 1) return ld(f, s, 1);
        if (k == nx) return ld(f, s, nx - 1);
        return i_c2i_in<21, T>([&](int kk) { return ld(f, s, kk); }, k);
    }
    constexpr double c13 = SBP42::c13, c14 = SBP42::c14;
    constexpr double Q[4][5] = {
        {39.0 / 28 + 39.0 / 14 * c13 + 39.0 / 7 * c14, -2.0 / 7 - 39.0 / 7 * c13 - 78.0 / 7 * c14,
         39.0 / 14 * c13 + 39.0 / 7 * c14 - 3.0 / 28, 0, 0},
        {13.0 / 27 - 52.0 / 27 * c13 - 26.0 / 9 * c14, 29.0 / 54 + 104.0 / 27 * c13 + 52.0 / 9 * c14,
         -1.0 / 54 - 52.0 / 27 * c13 - 26.0 / 9 * c14, 0, 0},
        {13.0 / 12 * c13, -13.0 / 6 * c13 + 7.0 / 16, 5.0 / 8 + 13.0 / 12 * c13, -1.0 / 16, 0},
        {78.0 / 71 * c14, -156.0 / 71 * c14 - 4.0 / 71, 39.0 / 71 + 78.0 / 71 * c14, 81.0 / 142, -9.0 / 142}};
    if (k <= 4) {
        const int r = k - 1;
        return T(Q[r][0]) * ld(f, s, 1) + T(Q[r][1]) * ld(f, s, 2) + T(Q[r][2]) * ld(f, s, 3) +
               T(Q[r][3]) * ld(f, s, 4) + T(Q[r][4]) * ld(f, s, 5);
    }
    if (k >= nx - 3) {
        const int r = nx - k;
        return T(Q[r][0]) * ld(f, s, nx - 1) + T(Q[r][1]) * ld(f, s, nx - 2) + T(Q[r][2]) * ld(f, s, nx - 3) +
               T(Q[r][3]) * ld(f, s, nx - 4) + T(Q[r][4]) * ld(f, s, nx - 5);
    }
    return i_c2i_in<42, T>([&](int kk) { return ld(f, s, kk); }, k);
}

}  // namespace swe
