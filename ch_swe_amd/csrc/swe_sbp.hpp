// 1-D staggered SBP stencils evaluated at one output point (device side).
//
// f is a line accessor: f(k) returns the value at 1-based line index k
// (halo indices <= 0 and > n are requested only by SAT terms, where the
// caller provides the neighbour's values); n = number of cell centres on the
// line; k = 1-based output index.  Coefficients follow
//   differential_operators/sbp_operators/sbp_diff_i2c_{21,42,63}_mod.f90,
//   sbp_diff_c2i_{21,42,63}_mod.f90, sbp_interp_i2c_{21,42,63}_mod.f90,
//   sbp_interp_c2i_{21,42,63}_mod.f90,
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

// sbp63: closure rows evaluated to doubles (row r reads points 1..9 of its
// end, mirrored at the right end), interior stencils, the centre / interface
// norms' boundary weights and the SAT projection.  Only the multi-kernel
// path instantiates these (the fused kernels' regions are sized for the
// shorter sbp21 / sbp42 closures).
constexpr double kDi2c63Q[6][9] = {{-0.9795214020218487, 0.9597522747540616, -0.0021284121310923276, 0.04308560808739484, -0.02118806868851538, 0.0, 0.0, 0.0, 0.0},
 {0.10094185622089052, -1.3621007582168956, 1.480651137325344, -0.2787674248835621, 0.059275189554223895, 0.0, 0.0, 0.0, 0.0},
 {0.1050522197237869, -0.3755418649660661, -0.5066880707069372, 0.7227932046793403, 0.05138416400746168, 0.0030003472624146313, 0.0, 0.0, 0.0},
 {-0.010281664659776757, 0.009016888173475901, 0.1158762991569453, -1.3063741185003286, 1.2879206487212211, -0.10361859147794651, 0.007460538586412149, 0.0, 0.0},
 {-0.08522272532774755, 0.2300959768599209, -0.06748505183821596, -0.28075281972041166, -0.77959568679678, 1.0363920760788554, -0.05757733755993641, 0.0041455683043154215, 0.0},
 {0.03958146265158008, -0.10105244572213268, 0.0269485128845091, 0.10551623486389128, -0.009407429883264403, -1.1945625283432082, 1.1945625283432082, -0.066364584907956, 0.004778250113372832}};
constexpr double kDc2i63Q[5][9] = {{-3.137061163091724, 6.472195962306311, -4.476505551641335, 0.9669525120033824, 0.2921335971506185, -0.11771535672725233, 0.0, 0.0, 0.0},
 {-0.8550973738466245, 0.5371097071174838, 0.4442015766629761, -0.00428923422758698, -0.1969782207705677, 0.07505354506431924, 0.0, 0.0, 0.0},
 {0.003267033496606009, -1.0058841165930086, 1.032532798072641, -0.09496402962593153, 0.0995309639229438, -0.03448264927325102, 0.0, 0.0, 0.0},
 {-0.0460358301286618, 0.1318266207921457, -1.0252815152152988, 0.7452431221796392, 0.2882309687613436, -0.093983366389168, 0.0, 0.0, 0.0},
 {0.025410671711714995, -0.031462633326690415, -0.08181244708002722, -0.8246706459806198, 0.8983528780132024, 0.009405106598725824, 0.004777070063694267, 0.0, 0.0}};
constexpr double kIi2c63Q[6][9] = {{0.4559762640735722, 0.5156609261411327, 0.09215963713516839, -0.055207108988023945, -0.008589718361849435, 0.0, 0.0, 0.0, 0.0},
 {-0.2329150785274282, 1.0886403314451474, 0.006569476829126231, 0.1526003661160169, -0.014895095862863417, 0.0, 0.0, 0.0, 0.0},
 {-0.13300622901536394, 0.24335188147388265, 0.43548086151449933, 0.4385084167668463, 0.008164201104099517, 0.007500868156036579, 0.0, 0.0, 0.0},
 {0.10717671074312715, -0.3332211159670528, 0.3310769312612241, 0.29067744812134255, 0.7410665665922483, -0.15542788721691977, 0.018651346466030373, 0.0, 0.0},
 {0.09305432569823871, -0.07099703081266313, -0.291616405335888, 0.3005623879587947, 0.5268027700312061, 0.5181960380394277, -0.08636600633990461, 0.010363920760788554, 0.0},
 {-0.05115521553501697, 0.05753938634775091, 0.09801308881771609, -0.1230378129758785, -0.08832072291927771, 0.5972812641716041, 0.5972812641716041, -0.09954687736193402, 0.011945625283432081}};
constexpr double kIc2i63Q[5][9] = {{1.6267303202606325, -0.3677632041998704, -0.6299565857805535, 0.2041457799345569, 0.3189794129688649, -0.15213572318363067, 0.0, 0.0, 0.0},
 {0.4594313713417274, 0.42927755971902565, 0.287843512318841, -0.15850960868795075, -0.060778415165360984, 0.04273558047371772, 0.0, 0.0, 0.0},
 {0.14146161692901638, 0.004462990795171873, 0.8874262064612957, 0.2713272665549153, -0.43009320031948844, 0.12541511957908888, 0.0, 0.0, 0.0},
 {-0.058987332524401276, 0.07216334765484655, 0.6220237974938234, 0.16582184683346107, 0.3085681858542267, -0.10958984531195667, 0.0, 0.0, 0.0},
 {-0.0103015766372022, -0.007906156741522508, 0.012998815562764554, 0.47451358497362967, 0.607051568675785, -0.0882989109926898, 0.01194267515923567, 0.0, 0.0}};
constexpr double kH63c[6] = {325363.0 / 276480, 144001.0 / 276480, 43195.0 / 27648,
                             86857.0 / 138240, 312623.0 / 276480, 271229.0 / 276480};
constexpr double kH63i[5] = {95.0 / 288, 317.0 / 240, 23.0 / 30, 793.0 / 720, 157.0 / 160};
constexpr double kR63[4] = {35.0 / 16, -35.0 / 16, 21.0 / 16, -5.0 / 16};
constexpr double kQd63[3] = {75.0 / 64, -25.0 / 384, 3.0 / 640};
constexpr double kQi63[3] = {75.0 / 128, -25.0 / 256, 3.0 / 256};
constexpr double kDi2c63S[4] = {0.5 * kR63[0] / kH63c[0], 0.5 * kR63[1] / kH63c[1], 0.5 * kR63[2] / kH63c[2],
                                0.5 * kR63[3] / kH63c[3]};

// row r of a 9-point closure table at one end, summed from m = 0:
// sum_m Q[r][m] f(a + sg m)
template <typename T, const auto &Q, typename A>
SWE_HD T closure63(const A &f, int s, int r, int a, int sg)
{
    T v = T(0);
#pragma unroll
    for (int m = 0; m < 9; m++) v += T(Q[r][m]) * ld(f, s, a + sg * m);
    return v;
}

// ---- interior rows (no closure, no SAT); the full stencils below use the
//      same expressions for their interior rows
template <int ORDER, typename T, typename A>
SWE_HD T d_i2c_in(const A &f, int k)
{
    if (ORDER == 63)
        return T(kQd63[0]) * (f(k + 1) - f(k)) + T(kQd63[1]) * (f(k + 2) - f(k - 1)) +
               T(kQd63[2]) * (f(k + 3) - f(k - 2));
    if (ORDER == 21) return f(k + 1) - f(k);
    return (T(27) * (f(k + 1) - f(k)) - (f(k + 2) - f(k - 1))) * T(1.0 / 24);
}
template <int ORDER, typename T, typename A>
SWE_HD T d_c2i_in(const A &f, int k)
{
    if (ORDER == 63)
        return T(kQd63[0]) * (f(k) - f(k - 1)) + T(kQd63[1]) * (f(k + 1) - f(k - 2)) +
               T(kQd63[2]) * (f(k + 2) - f(k - 3));
    if (ORDER == 21) return f(k) - f(k - 1);
    return (T(27) * (f(k) - f(k - 1)) - (f(k + 1) - f(k - 2))) * T(1.0 / 24);
}
template <int ORDER, typename T, typename A>
SWE_HD T i_i2c_in(const A &f, int k)
{
    if (ORDER == 63)
        return T(kQi63[0]) * (f(k + 1) + f(k)) + T(kQi63[1]) * (f(k + 2) + f(k - 1)) +
               T(kQi63[2]) * (f(k + 3) + f(k - 2));
    if (ORDER == 21) return T(0.5) * (f(k + 1) + f(k));
    return (T(9) * (f(k + 1) + f(k)) - (f(k + 2) + f(k - 1))) * T(1.0 / 16);
}
template <int ORDER, typename T, typename A>
SWE_HD T i_c2i_in(const A &f, int k)
{
    if (ORDER == 63)
        return T(kQi63[0]) * (f(k) + f(k - 1)) + T(kQi63[1]) * (f(k + 1) + f(k - 2)) +
               T(kQi63[2]) * (f(k + 2) + f(k - 3));
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
    if constexpr (ORDER == 63) {
        // closure rows 1..6 of each end, then the SAT 0.5 R_r / hc_r on rows 1..4
        if (k <= 6) {
            T v = closure63<T, kDi2c63Q>(f, s, k - 1, 1, 1);
            if (k <= 4) v += T(kDi2c63S[k - 1]) * (ld(f, s, 1) - ld(f, s, 0));
            return v;
        }
        if (k >= n - 5) {
            T v = -closure63<T, kDi2c63Q>(f, s, n - k, n + 1, -1);
            if (k >= n - 3) v += T(kDi2c63S[n - k]) * (ld(f, s, n + 2) - ld(f, s, n + 1));
            return v;
        }
        return d_i2c_in<63, T>([&](int kk) { return ld(f, s, kk); }, k);
    } else if (ORDER == 21) {
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
    if constexpr (ORDER == 63) {
        // closure rows 1..5 of each end; the SAT on the end interfaces reads
        // the neighbour's normal flux at depths 1..4
        if (k <= 5) {
            T v = closure63<T, kDc2i63Q>(f, s, k - 1, 1, 1);
            if (k == 1) {
                T sl = T(0);
#pragma unroll
                for (int r = 0; r < 4; r++) sl += T(kR63[r]) * (ld(f, s, 1 + r) - ld(f, s, -r));
                v += T(0.5 / kH63i[0]) * sl;
            }
            return v;
        }
        if (k >= nx - 4) {
            T v = -closure63<T, kDc2i63Q>(f, s, nx - k, nx - 1, -1);
            if (k == nx) {
                T sr = T(0);
#pragma unroll
                for (int r = 0; r < 4; r++) sr += T(kR63[r]) * (ld(f, s, nx - 1 - r) - ld(f, s, nx + r));
                v -= T(0.5 / kH63i[0]) * sr;
            }
            return v;
        }
        return d_c2i_in<63, T>([&](int kk) { return ld(f, s, kk); }, k);
    } else if (ORDER == 21) {
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
    if constexpr (ORDER == 63) {
        if (k <= 6) return closure63<T, kIi2c63Q>(f, s, k - 1, 1, 1);
        if (k >= n - 5) return closure63<T, kIi2c63Q>(f, s, n - k, n + 1, -1);
        return i_i2c_in<63, T>([&](int kk) { return ld(f, s, kk); }, k);
    }
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
    if constexpr (ORDER == 63) {
        if (k <= 5) return closure63<T, kIc2i63Q>(f, s, k - 1, 1, 1);
        if (k >= nx - 4) return closure63<T, kIc2i63Q>(f, s, nx - k, nx - 1, -1);
        return i_c2i_in<63, T>([&](int kk) { return ld(f, s, kk); }, k);
    }
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
