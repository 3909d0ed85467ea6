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

// Closure coefficients of the compile-time tables below.  Row r of a table
// is picked by selects over constants (the double constants rounded to T at
// compile time): a table indexed by a runtime row is read from memory, and
// converted, inside the stencil, which put dependent loads into every level
// of the edge tiles.  The tables are namespace-scope constexpr arrays so that
// they can be template arguments.
template <typename T, const auto &Q, int J>
SWE_HD T qpick(int r)
{
    constexpr int R = int(sizeof(Q) / sizeof(Q[0]));
    static_assert(R >= 2 && R <= 4, "closure tables have 2 to 4 rows");
    constexpr T q0 = T(Q[0][J]), q1 = T(Q[1][J]);
    constexpr T q2 = T(Q[R > 2 ? 2 : 1][J]), q3 = T(Q[R > 3 ? 3 : 1][J]);
    T v = r == 1 ? q1 : q0;
    if constexpr (R > 2) v = r == 2 ? q2 : v;
    if constexpr (R > 3) v = r == 3 ? q3 : v;
    return v;
}
template <typename T, const auto &S>
SWE_HD T spick(int r)
{
    constexpr int R = int(sizeof(S) / sizeof(S[0]));
    static_assert(R >= 2 && R <= 4, "closure tables have 2 to 4 rows");
    constexpr T s0 = T(S[0]), s1 = T(S[1]);
    constexpr T s2 = T(S[R > 2 ? 2 : 1]), s3 = T(S[R > 3 ? 3 : 1]);
    T v = r == 1 ? s1 : s0;
    if constexpr (R > 2) v = r == 2 ? s2 : v;
    if constexpr (R > 3) v = r == 3 ? s3 : v;
    return v;
}
template <typename T, const auto &S, int I>
SWE_HD constexpr T cval()
{
    return T(S[I]);
}

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

// sbp42 interfaces -> centres closure rows (q11..q35) and SAT weights
// 0.5 R_r / h_r
constexpr double kDi2c42Q[3][5] = {{-79.0 / 78, 27.0 / 26, -1.0 / 26, 1.0 / 78, 0},
                                   {2.0 / 21, -9.0 / 7, 9.0 / 7, -2.0 / 21, 0},
                                   {1.0 / 75, 0, -27.0 / 25, 83.0 / 75, -1.0 / 25}};
constexpr double kDi2c42S[3] = {0.5 * (15.0 / 8) / (13.0 / 12), 0.5 * (-5.0 / 4) / (7.0 / 8),
                                0.5 * (3.0 / 8) / (25.0 / 24)};
// sbp42 centres -> interfaces closure rows 1..4, the SAT weight C and the
// extrapolation R of the neighbour's normal flux
constexpr double kDc2i42Q[4][5] = {{-2.0, 3.0, -1.0, 0, 0},
                                   {-1.0, 1.0, 0, 0, 0},
                                   {1.0 / 24, -9.0 / 8, 9.0 / 8, -1.0 / 24, 0},
                                   {-1.0 / 71, 6.0 / 71, -83.0 / 71, 81.0 / 71, -3.0 / 71}};
constexpr double kDc2i42C = 0.5 / (7.0 / 18);
constexpr double kDc2i42R[3] = {15.0 / 8, -5.0 / 4, 3.0 / 8};
// sbp42 interpolation closures (interfaces -> centres, centres -> interfaces)
constexpr double kIi2c42Q[3][5] = {
    {0.5 + SBP42::c13 + 2 * SBP42::c14, 0.5 - 2 * SBP42::c13 - 3 * SBP42::c14, SBP42::c13, SBP42::c14, 0},
    {-8.0 / 63 - 52.0 / 21 * SBP42::c13 - 104.0 / 21 * SBP42::c14,
     29.0 / 42 + 104.0 / 21 * SBP42::c13 + 52.0 / 7 * SBP42::c14, -52.0 / 21 * SBP42::c13 + 0.5,
     -52.0 / 21 * SBP42::c14 - 4.0 / 63, 0},
    {26.0 / 25 * SBP42::c13 + 52.0 / 25 * SBP42::c14 - 1.0 / 25,
     -1.0 / 50 - 52.0 / 25 * SBP42::c13 - 78.0 / 25 * SBP42::c14, 3.0 / 5 + 26.0 / 25 * SBP42::c13,
     13.0 / 25 + 26.0 / 25 * SBP42::c14, -3.0 / 50}};
constexpr double kIc2i42Q[4][5] = {
    {39.0 / 28 + 39.0 / 14 * SBP42::c13 + 39.0 / 7 * SBP42::c14, -2.0 / 7 - 39.0 / 7 * SBP42::c13 - 78.0 / 7 * SBP42::c14,
     39.0 / 14 * SBP42::c13 + 39.0 / 7 * SBP42::c14 - 3.0 / 28, 0, 0},
    {13.0 / 27 - 52.0 / 27 * SBP42::c13 - 26.0 / 9 * SBP42::c14, 29.0 / 54 + 104.0 / 27 * SBP42::c13 + 52.0 / 9 * SBP42::c14,
     -1.0 / 54 - 52.0 / 27 * SBP42::c13 - 26.0 / 9 * SBP42::c14, 0, 0},
    {13.0 / 12 * SBP42::c13, -13.0 / 6 * SBP42::c13 + 7.0 / 16, 5.0 / 8 + 13.0 / 12 * SBP42::c13, -1.0 / 16, 0},
    {78.0 / 71 * SBP42::c14, -156.0 / 71 * SBP42::c14 - 4.0 / 71, 39.0 / 71 + 78.0 / 71 * SBP42::c14, 81.0 / 142,
     -9.0 / 142}};

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
        // closure rows (q11..q35) kDi2c42Q and SAT weights 0.5*R_r/h_r kDi2c42S
        if (k <= 3) {
            const int r = k - 1;
            T v = qpick<T, kDi2c42Q, 0>(r) * ld(f, s, 1) + qpick<T, kDi2c42Q, 1>(r) * ld(f, s, 2) + qpick<T, kDi2c42Q, 2>(r) * ld(f, s, 3) +
                  qpick<T, kDi2c42Q, 3>(r) * ld(f, s, 4) + qpick<T, kDi2c42Q, 4>(r) * ld(f, s, 5);
            v += spick<T, kDi2c42S>(r) * (ld(f, s, 1) - ld(f, s, 0));
            return v;
        }
        if (k >= n - 2) {
            const int r = n - k;
            T v = -(qpick<T, kDi2c42Q, 0>(r) * ld(f, s, n + 1) + qpick<T, kDi2c42Q, 1>(r) * ld(f, s, n) + qpick<T, kDi2c42Q, 2>(r) * ld(f, s, n - 1) +
                    qpick<T, kDi2c42Q, 3>(r) * ld(f, s, n - 2) + qpick<T, kDi2c42Q, 4>(r) * ld(f, s, n - 3));
            v += spick<T, kDi2c42S>(r) * (ld(f, s, n + 2) - ld(f, s, n + 1));
            return v;
        }
        return d_i2c_in<42, T>([&](int kk) { return ld(f, s, kk); }, k);
    }
}

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
        constexpr T C = T(kDc2i42C);
        if (k <= 4) {
            const int r = k - 1;
            T v = qpick<T, kDc2i42Q, 0>(r) * ld(f, s, 1) + qpick<T, kDc2i42Q, 1>(r) * ld(f, s, 2) + qpick<T, kDc2i42Q, 2>(r) * ld(f, s, 3) +
                  qpick<T, kDc2i42Q, 3>(r) * ld(f, s, 4) + qpick<T, kDc2i42Q, 4>(r) * ld(f, s, 5);
            if (k == 1)
                v += C * (cval<T, kDc2i42R, 0>() * (ld(f, s, 1) - ld(f, s, 0)) + cval<T, kDc2i42R, 1>() * (ld(f, s, 2) - ld(f, s, -1)) +
                             cval<T, kDc2i42R, 2>() * (ld(f, s, 3) - ld(f, s, -2)));
            return v;
        }
        if (k >= nx - 3) {
            const int r = nx - k;
            T v = -(qpick<T, kDc2i42Q, 0>(r) * ld(f, s, nx - 1) + qpick<T, kDc2i42Q, 1>(r) * ld(f, s, nx - 2) + qpick<T, kDc2i42Q, 2>(r) * ld(f, s, nx - 3) +
                    qpick<T, kDc2i42Q, 3>(r) * ld(f, s, nx - 4) + qpick<T, kDc2i42Q, 4>(r) * ld(f, s, nx - 5));
            if (k == nx)
                v -= C * (cval<T, kDc2i42R, 0>() * (ld(f, s, nx - 1) - ld(f, s, nx)) +
                             cval<T, kDc2i42R, 1>() * (ld(f, s, nx - 2) - ld(f, s, nx + 1)) +
                             cval<T, kDc2i42R, 2>() * (ld(f, s, nx - 3) - ld(f, s, nx + 2)));
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
        constexpr T C = T(kDc2i42C);
        if (k <= 4) {  // w[j] = f(j - 2)
            const int r = k - 1;
            T v = qpick<T, kDc2i42Q, 0>(r) * w[3] + qpick<T, kDc2i42Q, 1>(r) * w[4] + qpick<T, kDc2i42Q, 2>(r) * w[5] + qpick<T, kDc2i42Q, 3>(r) * w[6] + qpick<T, kDc2i42Q, 4>(r) * w[7];
            if (k == 1) v += C * (cval<T, kDc2i42R, 0>() * (w[3] - w[2]) + cval<T, kDc2i42R, 1>() * (w[4] - w[1]) + cval<T, kDc2i42R, 2>() * (w[5] - w[0]));
            return v;
        }
        if (k >= nx - 3) {  // w[j] = f(nx - 5 + j)
            const int r = nx - k;
            T v = -(qpick<T, kDc2i42Q, 0>(r) * w[4] + qpick<T, kDc2i42Q, 1>(r) * w[3] + qpick<T, kDc2i42Q, 2>(r) * w[2] + qpick<T, kDc2i42Q, 3>(r) * w[1] +
                    qpick<T, kDc2i42Q, 4>(r) * w[0]);
            if (k == nx) v -= C * (cval<T, kDc2i42R, 0>() * (w[4] - w[5]) + cval<T, kDc2i42R, 1>() * (w[3] - w[6]) + cval<T, kDc2i42R, 2>() * (w[2] - w[7]));
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
    if (k <= 3) {
        const int r = k - 1;
        return qpick<T, kIi2c42Q, 0>(r) * ld(f, s, 1) + qpick<T, kIi2c42Q, 1>(r) * ld(f, s, 2) + qpick<T, kIi2c42Q, 2>(r) * ld(f, s, 3) +
               qpick<T, kIi2c42Q, 3>(r) * ld(f, s, 4) + qpick<T, kIi2c42Q, 4>(r) * ld(f, s, 5);
    }
    if (k >= n - 2) {
        const int r = n - k;
        return qpick<T, kIi2c42Q, 0>(r) * ld(f, s, n + 1) + qpick<T, kIi2c42Q, 1>(r) * ld(f, s, n) + qpick<T, kIi2c42Q, 2>(r) * ld(f, s, n - 1) +
               qpick<T, kIi2c42Q, 3>(r) * ld(f, s, n - 2) + qpick<T, kIi2c42Q, 4>(r) * ld(f, s, n - 3);
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
    if (k <= 4) {
        const int r = k - 1;
        return qpick<T, kIc2i42Q, 0>(r) * ld(f, s, 1) + qpick<T, kIc2i42Q, 1>(r) * ld(f, s, 2) + qpick<T, kIc2i42Q, 2>(r) * ld(f, s, 3) +
               qpick<T, kIc2i42Q, 3>(r) * ld(f, s, 4) + qpick<T, kIc2i42Q, 4>(r) * ld(f, s, 5);
    }
    if (k >= nx - 3) {
        const int r = nx - k;
        return qpick<T, kIc2i42Q, 0>(r) * ld(f, s, nx - 1) + qpick<T, kIc2i42Q, 1>(r) * ld(f, s, nx - 2) + qpick<T, kIc2i42Q, 2>(r) * ld(f, s, nx - 3) +
               qpick<T, kIc2i42Q, 3>(r) * ld(f, s, nx - 4) + qpick<T, kIc2i42Q, 4>(r) * ld(f, s, nx - 5);
    }
    return i_c2i_in<42, T>([&](int kk) { return ld(f, s, kk); }, k);
}

}  // namespace swe
