// libchswe: MI355X (gfx950) implementation of the linear SWE RK4 step on the
// staggered (Ch) equiangular cubed sphere, behind the C ABI of
// include/ch_swe.h.
//
// Structure of one operator application (operator_swm_lin_mod.f90 L66-108):
//   gather h duplicates          (grad SAT needs the neighbour's edge value)
//   k_xy_w      : W1 = J Qi12 Ic2i_y(v), W2 = J Qi12 Ic2i_x(u)        at xy
//   k_contra    : ut = Qi11 u + Ii2c_x(W1)/J, vt = Qi22 v + Ii2c_y(W2)/J
//                 (co2contra_Ch_mod.f90 L27-94), plus J ut, J vt
//   gather J ut / J vt across edges (div SAT normal flux, depth 2|3)
//   k_xy_cd     : raw Coriolis (up, vp) = fJ2 (Ic2i_y vt, -Ic2i_x ut) and
//                 raw div = (Dc2i_x(J ut) + Dc2i_y(J vt))/(hx J a)        at xy
//   gather (up, vp, div) duplicates
//   k_edge_sync : covariant vector sync of (up, vp) and scalar projection of
//                 div at the panel boundary (ecs_Ah_vec_sync_mod.f90 L34-125,
//                 halo_Ah_scalar_sync_mod.f90 L43-98)
//   k_tend      : du = -g Di2c_x(h)/(hx a) + Ii2c_x(up)/J, dv likewise,
//                 dh = -H0 div
// The MPI halo exchanges of the reference become gathers driven by
// plan-time maps (swe_topology.cpp).  Fields live in halo-padded per-panel
// arrays (HALO = 4 on each side, common pitch for all three meshes).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <deque>
#include <map>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ch_swe.h"
#include "swe_sbp.hpp"
#include "swe_error.hpp"
#include "swe_topology.hpp"
#include "swe_latlon.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(SWE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

#include "swe_geom.hpp"

// ------------------------------------------------------------------------
// operator kernels
// ------------------------------------------------------------------------
constexpr int TPB = 256;

inline unsigned blocks_for(long n) { return (unsigned)((n + TPB - 1) / TPB); }

// depth of the normal-flux halo the centres -> interfaces derivative's SAT
// reads: the length of its boundary projection R (2, 3, 4 for sbp21 / 42 / 63)
inline int ch_depth(int order) { return order == 21 ? 2 : order == 42 ? 3 : 4; }

template <typename T, int O>
__global__ void k_xy_w(Dims d, const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ JQ12,
                       T *__restrict__ W1, T *__restrict__ W2)
{
    const int N = d.N;
    long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= 6L * (N + 1) * (N + 1)) return;
    int p, i, j;
    unflat(n, N + 1, N + 1, p, i, j);
    long o = off(d, p, i, j);
    T c = JQ12[o];
    W1[o] = c * swe::i_c2i_pt<O, T>(swe::line(v + off(d, p, i, 0), d.P), 0, N, j);
    W2[o] = c * swe::i_c2i_pt<O, T>(swe::line(u + off(d, p, 0, j), 1), 0, N, i);
}

template <typename T, int O>
__global__ void k_contra(Dims d, const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ W1,
                         const T *__restrict__ W2, const T *__restrict__ Qu, const T *__restrict__ iJu,
                         const T *__restrict__ Ju, const T *__restrict__ Qv, const T *__restrict__ iJv,
                         const T *__restrict__ Jv, T *__restrict__ ut, T *__restrict__ vt, T *__restrict__ Jut,
                         T *__restrict__ Jvt)
{
    const int N = d.N;
    const long nu = 6L * N * (N + 1);
    long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= 2 * nu) return;
    int p, i, j;
    if (n < nu) {
        unflat(n, N, N + 1, p, i, j);
        long o = off(d, p, i, j);
        T val = Qu[o] * u[o] + swe::i_i2c_pt<O, T>(swe::line(W1 + off(d, p, 0, j), 1), 0, N, i) * iJu[o];
        ut[o] = val;
        Jut[o] = Ju[o] * val;
    } else {
        unflat(n - nu, N + 1, N, p, i, j);
        long o = off(d, p, i, j);
        T val = Qv[o] * v[o] + swe::i_i2c_pt<O, T>(swe::line(W2 + off(d, p, i, 0), d.P), 0, N, j) * iJv[o];
        vt[o] = val;
        Jvt[o] = Jv[o] * val;
    }
}

// halo gather: dst_field[dst] = sign * src_field[src]
struct GatherEntry {
    int dst, src;
    int fields;  // dst field index (bits 0-3), src field index (bits 4-7)
    int sign;
};
struct FieldSet {
    void *f[4];
};

template <typename T>
__global__ void k_gather(const GatherEntry *__restrict__ e, int n, FieldSet fs)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    GatherEntry g = e[k];
    T *dst = (T *)fs.f[g.fields & 15];
    const T *src = (const T *)fs.f[(g.fields >> 4) & 15];
    dst[g.dst] = T(g.sign) * src[g.src];
}

template <typename T, int O>
__global__ void k_xy_cd(Dims d, const T *__restrict__ ut, const T *__restrict__ vt, const T *__restrict__ Jut,
                        const T *__restrict__ Jvt, const T *__restrict__ fJ2, const T *__restrict__ iJa,
                        T *__restrict__ up, T *__restrict__ vp, T *__restrict__ div)
{
    const int N = d.N;
    long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= 6L * (N + 1) * (N + 1)) return;
    int p, i, j;
    unflat(n, N + 1, N + 1, p, i, j);
    long o = off(d, p, i, j);
    T f = fJ2[o];
    // calc_cori_at_p_tile (coriolis_Ch_mod.f90 L97-120): (up,vp) <- f J^2 (vp, -up)
    up[o] = f * swe::i_c2i_pt<O, T>(swe::line(vt + off(d, p, i, 0), d.P), 0, N, j);
    vp[o] = -f * swe::i_c2i_pt<O, T>(swe::line(ut + off(d, p, 0, j), 1), 0, N, i);
    // calc_div_tile (div_sbp_sat_mod.f90 L54-80), hx = hy
    T dx = swe::d_c2i_pt<O, T>(swe::line(Jut + off(d, p, 0, j), 1), 0, N, i);
    T dy = swe::d_c2i_pt<O, T>(swe::line(Jvt + off(d, p, i, 0), d.P), 0, N, j);
    div[o] = (dx + dy) * iJa[o];
}

// boundary points of each panel, each once (corners included)
__device__ inline void boundary_point(int N, int t, int &i, int &j)
{
    if (t < N) { i = t + 1; j = 1; }
    else if (t < 2 * N) { i = N + 1; j = t - N + 1; }
    else if (t < 3 * N) { i = N + 1 - (t - 2 * N); j = N + 1; }
    else { i = 1; j = N + 1 - (t - 3 * N); }
}

template <typename T>
__global__ void k_edge_sync(Dims d, T *__restrict__ up, T *__restrict__ vp, T *__restrict__ div,
                            const T *__restrict__ qb, const T *__restrict__ qt, const T *__restrict__ ql,
                            const T *__restrict__ qr)
{
    const int N = d.N, e = N + 1;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= 6 * 4 * N) return;
    int p = n / (4 * N), i, j;
    boundary_point(N, n - p * 4 * N, i, j);
    // covariant sync works on (U, V) = (vp, up)  (ecs_Ah_vec_sync_mod.f90 L51-53)
#define U(a, b) vp[off(d, p, a, b)]
#define V(a, b) up[off(d, p, a, b)]
#define D(a, b) div[off(d, p, a, b)]
    const T third = T(1.0 / 3.0), half = T(0.5);
    T nu, nv, nd;
    bool L = i == 1, R = i == e, B = j == 1, Tp = j == e;
    if (L && B) {
        nu = (U(1, 1) + U(0, 1) + U(1, 0) + qb[0] * V(1, 0)) * third;
        nv = (V(1, 1) + V(0, 1) + V(1, 0) + ql[0] * U(0, 1)) * third;
        nd = (D(1, 1) + D(0, 1) + D(1, 0)) * third;
    } else if (R && B) {
        nu = (U(e, 1) + U(e + 1, 1) + U(e, 0) + qb[e - 1] * V(e, 0)) * third;
        nv = (V(e, 1) + V(e + 1, 1) + V(e, 0) + qr[0] * U(e + 1, 1)) * third;
        nd = (D(e, 1) + D(e + 1, 1) + D(e, 0)) * third;
    } else if (L && Tp) {
        nu = (U(1, e) + U(0, e) + U(1, e + 1) + qt[0] * V(1, e + 1)) * third;
        nv = (V(1, e) + V(0, e) + V(1, e + 1) + ql[e - 1] * U(0, e)) * third;
        nd = (D(1, e) + D(0, e) + D(1, e + 1)) * third;
    } else if (R && Tp) {
        nu = (U(e, e) + U(e + 1, e) + U(e, e + 1) + qt[e - 1] * V(e, e + 1)) * third;
        nv = (V(e, e) + V(e + 1, e) + V(e, e + 1) + qr[e - 1] * U(e + 1, e)) * third;
        nd = (D(e, e) + D(e + 1, e) + D(e, e + 1)) * third;
    } else if (L) {
        nu = half * (U(0, j) + U(1, j));
        nv = half * (V(1, j) + V(0, j) + ql[j - 1] * U(0, j));
        nd = half * (D(0, j) + D(1, j));
    } else if (R) {
        nu = half * (U(e, j) + U(e + 1, j));
        nv = half * (V(e, j) + V(e + 1, j) + qr[j - 1] * U(e + 1, j));
        nd = half * (D(e, j) + D(e + 1, j));
    } else if (B) {
        nv = half * (V(i, 0) + V(i, 1));
        nu = half * (U(i, 0) + U(i, 1) + qb[i - 1] * V(i, 0));
        nd = half * (D(i, 1) + D(i, 0));
    } else {
        nv = half * (V(i, e) + V(i, e + 1));
        nu = half * (U(i, e) + U(i, e + 1) + qt[i - 1] * V(i, e + 1));
        nd = half * (D(i, e) + D(i, e + 1));
    }
    U(i, j) = nu;
    V(i, j) = nv;
    D(i, j) = nd;
#undef U
#undef V
#undef D
}

template <typename T, int O>
__global__ void k_tend(Dims d, const T *__restrict__ h, const T *__restrict__ up, const T *__restrict__ vp,
                       const T *__restrict__ div, const T *__restrict__ iJu, const T *__restrict__ iJv, T g_hxa,
                       T H0, T *__restrict__ du, T *__restrict__ dv, T *__restrict__ dh)
{
    const int N = d.N;
    const long nu = 6L * N * (N + 1), nh = 6L * (N + 1) * (N + 1);
    long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int p, i, j;
    if (n < nu) {
        unflat(n, N, N + 1, p, i, j);
        long o = off(d, p, i, j);
        du[o] = -g_hxa * swe::d_i2c_pt<O, T>(swe::line(h + off(d, p, 0, j), 1), 0, N, i) +
                swe::i_i2c_pt<O, T>(swe::line(up + off(d, p, 0, j), 1), 0, N, i) * iJu[o];
    } else if (n < 2 * nu) {
        unflat(n - nu, N + 1, N, p, i, j);
        long o = off(d, p, i, j);
        dv[o] = -g_hxa * swe::d_i2c_pt<O, T>(swe::line(h + off(d, p, i, 0), d.P), 0, N, j) +
                swe::i_i2c_pt<O, T>(swe::line(vp + off(d, p, i, 0), d.P), 0, N, j) * iJv[o];
    } else if (n < 2 * nu + nh) {
        unflat(n - 2 * nu, N + 1, N + 1, p, i, j);
        long o = off(d, p, i, j);
        dh[o] = -H0 * div[o];
    }
}

// out = s1*x + s2*y (+ out if update), over whole state buffers
template <typename T>
__global__ void k_axpy(long n, T *__restrict__ out, T s1, const T *__restrict__ x, T s2, const T *__restrict__ y,
                       int update)
{
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    T r = s1 * x[k];
    if (y) r += s2 * y[k];
    out[k] = update ? out[k] + r : r;
}

// ---- the implicit solve's conjugate gradients ---------------------------------
// cg_solver_mod.f90 L19-90 on the h field with the xy quadrature's inner
// product (grid_field_based_vector_mod.f90 dot / norm L169-191).

// sum over the interior xy points of w a b (w: the SBP_C quadrature weight),
// FP64 partials per block in a fixed order
// the xy quadrature weights of k_xy_dot, once per plan (FP64, state layout)
template <int O>
__global__ void k_xy_weights(Dims d, double hx, double *w)
{
    const int n = d.N + 1;
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= 6L * n * n) return;
    int p, i, j;
    unflat(k, n, n, p, i, j);
    w[off(d, p, i, j)] = quad_weight(d.N, O, hx, 0, i, j);
}

template <typename T>
__global__ __launch_bounds__(256) void k_xy_dot(Dims d, const double *w, const T *a, const T *b, double *part)
{
    __shared__ double ws[4];
    const int n = d.N + 1;
    const long np = 6L * n * n;
    double s = 0;
    for (long k = blockIdx.x * 256L + threadIdx.x; k < np; k += gridDim.x * 256L) {
        int p, i, j;
        unflat(k, n, n, p, i, j);
        const long o = off(d, p, i, j);
        s += w[o] * (double)a[o] * (double)b[o];
    }
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_down(s, m);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// one wave: lane l sums partials l, l + 64, ... in order, then a fixed
// shuffle tree (deterministic; one thread walking all partials costs the
// serial latency of nb loads)
__global__ void k_dot_final(const double *part, int nb, double *out)
{
    const int l = threadIdx.x & 63;
    double s = 0;
    for (int b = l; b < nb; b += 64) s += part[b];
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_down(s, m);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_fill(double *p, int n, double v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// The convergence state lives on the device, so that the host reads it
// only every few iterations: st[0] = 1 once converged, st[1] = the
// iteration that met the tolerance, st[2] = the latest relative residual.
// After convergence the update kernels do nothing, so x stays the iterate
// that met the tolerance whatever iterations follow in the same batch.

// x += alpha p, r -= alpha Ap, alpha = |r|^2 / (p, Ap) with |r| = sqrt(*rr)
template <typename T>
__global__ void k_cg_step(long n, T *x, T *r, const T *p, const T *Ap, const double *rr, const double *pap,
                          const double *st)
{
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= n || st[0] != 0.0) return;
    const double rn = sqrt(*rr), alpha = rn * rn / *pap;
    x[k] = x[k] + T(alpha) * p[k];
    r[k] = r[k] + T(-alpha) * Ap[k];
}

// the relative residual of iteration it against the tolerance (one thread)
__global__ void k_cg_check(const double *rr_new, const double *bb, double rel_tol, double *st)
{
    if (threadIdx.x != 0 || st[0] != 0.0) return;
    // st[3] counts the iterations, so that every batch of iterations is the
    // same launch sequence (replayed from one graph)
    const double it = st[3] + 1.0;
    st[3] = it;
    const double rel = sqrt(*rr_new) / sqrt(*bb);
    st[2] = rel;
    if (rel < rel_tol) {
        st[0] = 1.0;
        st[1] = it;
    }
}

// p = r + beta p, beta = (|r_new| / |r_old|)^2
template <typename T>
__global__ void k_cg_dir(long n, T *p, const T *r, const double *rr_new, const double *rr_old, const double *st)
{
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= n || st[0] != 0.0) return;
    const double q = sqrt(*rr_new) / sqrt(*rr_old), beta = q * q;
    p[k] = T(1.0) * r[k] + T(beta) * p[k];
}

// FP64 reduction over interior points: per-block partials in a fixed order
template <typename T>
__global__ void k_reduce(Dims d, const T *h, const T *u, const T *v, const T *ut, const T *vt, const double *wxy,
                         const double *wu, const double *wv, double *part)
{
    const int N = d.N;
    const long nu = 6L * N * (N + 1), nh = 6L * (N + 1) * (N + 1);
    __shared__ double sm[3][TPB];
    double a = 0, b = 0, c = 0;
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nh + 2 * nu; n += (long)gridDim.x * blockDim.x) {
        int p, i, j;
        if (n < nh) {
            unflat(n, N + 1, N + 1, p, i, j);
            long o = off(d, p, i, j);
            double hv = (double)h[o];
            a += hv * wxy[o];
            b += hv * hv * wxy[o];
        } else if (n < nh + nu) {
            unflat(n - nh, N, N + 1, p, i, j);
            long o = off(d, p, i, j);
            c += (double)u[o] * (double)ut[o] * wu[o];
        } else {
            unflat(n - nh - nu, N + 1, N, p, i, j);
            long o = off(d, p, i, j);
            c += (double)v[o] * (double)vt[o] * wv[o];
        }
    }
    sm[0][threadIdx.x] = a;
    sm[1][threadIdx.x] = b;
    sm[2][threadIdx.x] = c;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; q++) sm[q][threadIdx.x] += sm[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) part[3 * blockIdx.x + q] = sm[q][0];
}

// one wave, fixed order: lane l sums blocks l, l + 64, ..., then a shuffle
// tree (one thread walking nb blocks costs nb serial loads)
__global__ void k_reduce_final(const double *part, int nb, double *out)
{
    const int l = threadIdx.x & 63;
    double s[3] = {0, 0, 0};
    for (int k = l; k < nb; k += 64)
        for (int q = 0; q < 3; q++) s[q] += part[3 * k + q];
    for (int q = 0; q < 3; q++)
        for (int m = 32; m >= 1; m >>= 1) s[q] += __shfl_down(s[q], m);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) out[q] = s[q];
}

#include "swe_fused.inc"
#include "swe_fused_host.hpp"
#include "swe_latlon.inc"

// ---- diagnostics on the multi-kernel path (any SBP order) -------------------
// The same sums as k_diag / k_tends (print_integrals_diag, print_CFL_diag and
// print_integral_tends: swm_output_diag_mod.f90 L110-342), point by point in
// FP64 from the plan's fields: the interpolations to the u / v points and the
// centres are evaluated inside each point's sum (h -> hu -> hq as nested line
// stencils), the curl reads copies of u, v whose tangential duplicates were
// gathered from the neighbours.  Blocks write partials in k_diag's and
// k_tends' layouts, combined by the same final kernels.
constexpr int MK_BLOCKS = 1024;

// cell-centre (mesh o) J and f, FP64 (metric_2d_ecs_mod.f90 L338-411 at the o
// points; coriolis_factory_mod.f90 calc_coriolis_parameter L239-280)
__device__ inline void mk_centre(const GeomParams &g, int p, int i, int j, double &J, double &f)
{
    const double X = tan(-0.25 * PI + (i - 0.5) * g.hx), Y = tan(-0.25 * PI + (j - 0.5) * g.hx);
    const double s2 = 1.0 + X * X + Y * Y, s = sqrt(s2);
    J = (1.0 + X * X) * (1.0 + Y * Y) / (s2 * s);
    if (g.coriolis_const) {
        f = g.omega;
    } else {
        double r[3];
        for (int k = 0; k < 3; k++) r[k] = (c_EX[p][k] * X + c_EY[p][k] * Y - c_NN[p][k]) / s;
        const double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        f = 2.0 * g.omega * (g.axis[0] * r[0] + g.axis[1] * r[1] + g.axis[2] * r[2]) / rn;
    }
}

// h at u point (i, j): i2c along x of xy row j; at v point (i, j): along y
template <int O, typename T>
__device__ inline double mk_hu(const Dims &d, const T *h, int p, int i, int j)
{
    return swe::i_i2c_pt<O, double>([&](int k) { return double(h[off(d, p, k, j)]); }, 0, d.N, i);
}
template <int O, typename T>
__device__ inline double mk_hv(const Dims &d, const T *h, int p, int i, int j)
{
    return swe::i_i2c_pt<O, double>([&](int k) { return double(h[off(d, p, i, k)]); }, 0, d.N, j);
}
// at centre (i, j): i2c along y of hu (the first output of interp2d_uv2qvec_Ch)
// and i2c along x of hv (its second)
template <int O, typename T>
__device__ inline double mk_hq1(const Dims &d, const T *h, int p, int i, int j)
{
    return swe::i_i2c_pt<O, double>([&](int k) { return mk_hu<O>(d, h, p, i, k); }, 0, d.N, j);
}
template <int O, typename T>
__device__ inline double mk_hq2(const Dims &d, const T *h, int p, int i, int j)
{
    return swe::i_i2c_pt<O, double>([&](int k) { return mk_hv<O>(d, h, p, k, j); }, 0, d.N, i);
}
// curl_ch_sbp_sat at centre (i, j) (curl_sbp_sat_mod.f90 L26-74): i2c
// derivatives of v along x and u along y, SAT against the tangential
// duplicates held in ua / va at indices 0 and N + 2; J from mk_centre
template <int O, typename T>
__device__ inline double mk_curl(const Dims &d, const GeomParams &g, const T *ua, const T *va, int p, int i, int j,
                                 double J)
{
    const double Dx = swe::d_i2c_pt<O, double>([&](int k) { return double(va[off(d, p, k, j)]); }, 0, d.N, i);
    const double Dy = swe::d_i2c_pt<O, double>([&](int k) { return double(ua[off(d, p, i, k)]); }, 0, d.N, j);
    return (Dx / g.hx - Dy / g.hx) / (J * g.radius);
}

template <int K>
__device__ inline void mk_block_partials(double (&s)[K], const int (&kind)[K], double *part)
{
    // kind 0: sum, 1: min, 2: max
    __shared__ double ws[K][4];
    for (int q = 0; q < K; q++) {
        double x = s[q];
        for (int m = 32; m >= 1; m >>= 1) {
            const double y = __shfl_down(x, m);
            x = kind[q] == 0 ? x + y : kind[q] == 1 ? fmin(x, y) : fmax(x, y);
        }
        if ((threadIdx.x & 63) == 0) ws[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int q = threadIdx.x;
        double x = ws[q][0];
        for (int w = 1; w < 4; w++) x = kind[q] == 0 ? x + ws[q][w] : kind[q] == 1 ? fmin(x, ws[q][w]) : fmax(x, ws[q][w]);
        part[K * blockIdx.x + q] = x;
    }
}

// hmin, hmax, mass, KE, PE, enstrophy, max |c ut u|, max |c vt v| over the
// xy, u, v points and the centres (one grid-stride index space)
template <typename T, int O>
__global__ __launch_bounds__(256) void k_mk_diag(Dims d, GeomParams g, const T *u, const T *v, const T *h,
                                                 const T *ut, const T *vt, const T *ua, const T *va, double grav,
                                                 double cfl_c, double *part)
{
    const int N = d.N;
    const long nxy = 6L * (N + 1) * (N + 1), nuv = 6L * N * (N + 1), no = 6L * N * N;
    double s[DIAG_K] = {INFINITY, -INFINITY, 0, 0, 0, 0, 0, 0};
    for (long n = blockIdx.x * 256L + threadIdx.x; n < nxy + 2 * nuv + no; n += gridDim.x * 256L) {
        int p, i, j;
        if (n < nxy) {
            unflat(n, N + 1, N + 1, p, i, j);
            const double hv = h[off(d, p, i, j)], w = quad_weight(N, O, g.hx, 0, i, j);
            s[0] = fmin(s[0], hv);
            s[1] = fmax(s[1], hv);
            s[2] += hv * w;
            s[4] += 0.5 * grav * hv * hv * w;
        } else if (n < nxy + nuv) {
            unflat(n - nxy, N, N + 1, p, i, j);
            const long o = off(d, p, i, j);
            const double uu = u[o], uc = ut[o];
            s[3] += 0.5 * (mk_hu<O>(d, h, p, i, j) * uc) * uu * quad_weight(N, O, g.hx, 1, i, j);
            s[6] = fmax(s[6], fabs(cfl_c * uc * uu));
        } else if (n < nxy + 2 * nuv) {
            unflat(n - nxy - nuv, N + 1, N, p, i, j);
            const long o = off(d, p, i, j);
            const double vv = v[o], vc = vt[o];
            s[3] += 0.5 * (mk_hv<O>(d, h, p, i, j) * vc) * vv * quad_weight(N, O, g.hx, 2, i, j);
            s[7] = fmax(s[7], fabs(cfl_c * vc * vv));
        } else {
            unflat(n - nxy - 2 * nuv, N, N, p, i, j);
            double J, f;
            mk_centre(g, p, i, j, J, f);
            const double hq = 0.5 * mk_hq1<O>(d, h, p, i, j) + 0.5 * mk_hq2<O>(d, h, p, i, j);
            const double c = mk_curl<O>(d, g, ua, va, p, i, j, J) + f;
            s[5] += (0.5 * c * c) / hq * quad_weight(N, O, g.hx, 3, i, j);
        }
    }
    constexpr int kind[DIAG_K] = {1, 2, 0, 0, 0, 0, 2, 2};
    mk_block_partials<DIAG_K>(s, kind, part);
}

// ke1, ke2, pe, e1, e2 of print_integral_tends: state (u, v, h) with
// contravariant uc, vc and tangential copies ua, va; tendency (tu, tv, th)
// with tangential copies tua, tva
template <typename T, int O>
__global__ __launch_bounds__(256) void k_mk_tends(Dims d, GeomParams g, const T *u, const T *v, const T *h,
                                                  const T *uc, const T *vc, const T *ua, const T *va, const T *tu,
                                                  const T *tv, const T *th, const T *tua, const T *tva, double grav,
                                                  double *part)
{
    const int N = d.N;
    const long nxy = 6L * (N + 1) * (N + 1), nuv = 6L * N * (N + 1), no = 6L * N * N;
    double s[TENDS_K] = {0, 0, 0, 0, 0};
    for (long n = blockIdx.x * 256L + threadIdx.x; n < nxy + 2 * nuv + no; n += gridDim.x * 256L) {
        int p, i, j;
        if (n < nxy) {
            unflat(n, N + 1, N + 1, p, i, j);
            const long o = off(d, p, i, j);
            s[2] += grav * double(h[o]) * double(th[o]) * quad_weight(N, O, g.hx, 0, i, j);
        } else if (n < nxy + nuv) {
            unflat(n - nxy, N, N + 1, p, i, j);
            const long o = off(d, p, i, j);
            const double w = quad_weight(N, O, g.hx, 1, i, j), c = uc[o];
            s[0] += 0.5 * (mk_hu<O>(d, th, p, i, j) * c) * double(u[o]) * w;
            s[1] += (mk_hu<O>(d, h, p, i, j) * c) * double(tu[o]) * w;
        } else if (n < nxy + 2 * nuv) {
            unflat(n - nxy - nuv, N + 1, N, p, i, j);
            const long o = off(d, p, i, j);
            const double w = quad_weight(N, O, g.hx, 2, i, j), c = vc[o];
            s[0] += 0.5 * (mk_hv<O>(d, th, p, i, j) * c) * double(v[o]) * w;
            s[1] += (mk_hv<O>(d, h, p, i, j) * c) * double(tv[o]) * w;
        } else {
            unflat(n - nxy - 2 * nuv, N, N, p, i, j);
            double J, f;
            mk_centre(g, p, i, j, J, f);
            const double w = quad_weight(N, O, g.hx, 3, i, j);
            const double hq = mk_hq1<O>(d, h, p, i, j), hqt = mk_hq1<O>(d, th, p, i, j);
            const double c = mk_curl<O>(d, g, ua, va, p, i, j, J) + f;
            const double zt = mk_curl<O>(d, g, tua, tva, p, i, j, J);
            s[3] += (c * zt) / hq * w;
            s[4] += (((-0.5 * c * c) * hqt) / hq) / hq * w;
        }
    }
    constexpr int kind[TENDS_K] = {0, 0, 0, 0, 0};
    mk_block_partials<TENDS_K>(s, kind, part);
}

// ------------------------------------------------------------------------
// engine
// ------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Optional per-launch timing of the fused kernels (measurement aid): each
// stage and records launch goes out through hipExtLaunchKernelGGL with a
// start and a stop event, which the dispatch itself timestamps (no marker
// packets between the kernels, so the spans are the kernels' own run times).
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    struct Span {
        hipEvent_t a, b;
        int kind;  // 0 stage kernel, 1 records kernel
    };
    std::vector<Span> spans;
    ~KernelTimer()
    {
        for (auto &s : spans) {
            (void)hipEventDestroy(s.a);
            (void)hipEventDestroy(s.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    // launch kernel on grid x nthr threads, timed when on (kind 0: stage
    // kernel, 1: records kernel)
    template <class K, class... A>
    void launch(K kernel, int grid, int nthr, hipStream_t s, int kind, A... args)
    {
        if (!on) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(nthr), 0, s, args...);
            return;
        }
        hipEvent_t a = get(), b = get();
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(nthr), 0, s, a, b, 0, args...);
        spans.push_back(Span{a, b, kind});
    }
    int collect(double *ms, int *cnt)
    {
        ms[0] = ms[1] = 0;
        cnt[0] = cnt[1] = 0;
        for (auto &sp : spans) {
            HIP_TRY(hipEventSynchronize(sp.b));
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, sp.a, sp.b));
            ms[sp.kind] += t;
            cnt[sp.kind]++;
            pool.push_back(sp.a);
            pool.push_back(sp.b);
        }
        spans.clear();
        return SWE_OK;
    }
};

struct EngineBase {
    KernelTimer timer;
    virtual ~EngineBase() = default;
    virtual int apply(void *out, void *in) = 0;
    virtual int axpy(void *out, double s1, void *x, double s2, void *y, int update) = 0;
    // part: >= 3 INT_MAX_BLOCKS doubles of partial sums, red: 4 doubles (device)
    virtual int integrals(void *st, double *part, double *red, double *mass, double *energy) = 0;
    virtual int integral_buffers(double **part, double **red) = 0;
    // print_integrals_diag's values: hmin, hmax, mass, TE, KE, PE, enstrophy
    virtual int integrals_diag(void *st, double grav, double out[7]) = 0;
    // operator_swm_lin_t%solve: vout = (I - dt L0)^-1 rhs (L0: no Coriolis)
    virtual int solve(void *vout, void *rhs, double dt, int max_iter, double rel_tol, int *iters, double *relres) = 0;
    // print_CFL_diag's values: gravity-wave CFL, advective CFL along x and y
    virtual int cfl_diag(void *st, double dt, double grav, double out[3]) = 0;
    // print_integral_tends's values of the operator's tendency: TE_tend, Enstrophy_tend
    virtual int integral_tends(void *st, double grav, double out[2]) = 0;
    // h on the nlat x nlon lat-lon grid as the output writes h.dat (host out)
    virtual int latlon_h(void *st, int nlat, int nlon, float *out) = 0;
    // the zonal and meridional winds on it as the output writes u.dat, v.dat
    virtual int latlon_uv(void *st, int nlat, int nlon, float *uo, float *vo) = 0;
    virtual int rk4(void *v0, void *acc, void *ya, void *yb, double dt, int nsteps) = 0;
    virtual bool fused() const = 0;
    virtual size_t elem() const = 0;
    virtual size_t device_bytes() const = 0;
};

struct Maps {
    GatherEntry *xy = nullptr, *ch = nullptr, *xyv = nullptr, *tan = nullptr;
    int n_xy = 0, n_ch = 0, n_xyv = 0, n_tan = 0;
};

}  // namespace

struct swe_plan {
    swe_config cfg;
    Dims d;
    int device = 0;  // the HIP device current when the plan was made
    hipStream_t stream = nullptr;
    std::unique_ptr<EngineBase> eng;
    swe_state *rk[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // k1..k4, y
    swe_state *imp = nullptr;  // v_new of the implicit Euler and Crank-Nicolson steps
    // the multi-kernel RK4 loop's graph of MK_GRAPH_STEPS steps, for the last
    // (state, dt) it was captured for
    hipGraphExec_t mk_graph = nullptr;
    const void *mk_key = nullptr;
    double mk_dt = 0;
};

struct swe_state {
    swe_plan *plan;
    void *buf;  // u, v, h fields back to back (3 * F elements)
};

namespace {

template <typename T>
struct Engine : EngineBase {
    swe_plan *pl;
    Dims d;
    int order;
    GeomParams geo{};
    std::vector<void *> allocs;
    size_t bytes = 0;  // device bytes allocated so far
    // per-panel metric, work fields and halo maps of the unfused path (made
    // by unfused plans only)
    bool full_ready = false;
    T *JQ12 = nullptr, *fJ2 = nullptr, *iJa = nullptr, *Qu = nullptr, *iJu = nullptr, *Ju = nullptr,
      *Qv = nullptr, *iJv = nullptr, *Jv = nullptr;
    T *qb, *qt, *ql, *qr;
    double *wxy = nullptr, *wu = nullptr, *wv = nullptr;
    T *W1 = nullptr, *W2 = nullptr, *ut = nullptr, *vt = nullptr, *Jut = nullptr, *Jvt = nullptr,
      *up = nullptr, *vp = nullptr, *div = nullptr;
    Maps maps;
    double *part = nullptr, *red = nullptr;
    static constexpr int nblk_red = 512;  // grid of the unfused path's k_reduce
    static_assert(nblk_red <= INT_MAX_BLOCKS, "partial sums of k_reduce must fit the integral buffers");

    // fused path
    static constexpr int TY = sizeof(T) == 4 ? 16 : 8;
    bool use_fused = true;
    Coef<T> cf{};
    T *rec[2] = {nullptr, nullptr};
    int ntx = 0, nty = 0, ntiles = 0, nseg = 0, fwx = 0, fwy = 0;
    int nthr = F_THREADS;  // workgroup size of the fused kernels

    size_t elem() const override { return sizeof(T); }
    bool fused() const override { return use_fused; }
    size_t device_bytes() const override { return bytes; }

    ~Engine() override
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (cg_graph) (void)hipGraphExecDestroy(cg_graph);
        for (void *p : allocs) (void)hipFree(p);
    }

    // hipFree of a buffer from alloc
    void release(void *p)
    {
        for (size_t k = 0; k < allocs.size(); k++)
            if (allocs[k] == p) {
                allocs.erase(allocs.begin() + k);
                (void)hipFree(p);
                return;
            }
    }

    int alloc(void **p, size_t nbytes)
    {
        hipError_t e = hipMalloc(p, nbytes);
        if (e != hipSuccess) return fail(SWE_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(*p);
        bytes += nbytes;
        e = hipMemsetAsync(*p, 0, nbytes, pl->stream);
        if (e != hipSuccess) return fail(SWE_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
        return SWE_OK;
    }

    template <typename U>
    int field(U **p)
    {
        return alloc((void **)p, d.F * sizeof(U));
    }

    int upload_map(int kind, int depth, const std::vector<int> &fsel_dst, GatherEntry **out, int *count)
    {
        auto m = swe::build_halo_map(d.N, kind, depth);
        std::vector<GatherEntry> ge(m.size());
        for (size_t k = 0; k < m.size(); k++) {
            const auto &e = m[k];
            ge[k].dst = (int)off(d, e.tp, e.ti, e.tj);
            ge[k].src = (int)off(d, e.sp, e.si, e.sj);
            ge[k].fields = fsel_dst[e.tc] | (fsel_dst[e.sc] << 4);
            ge[k].sign = e.sign;
        }
        int rc = alloc((void **)out, ge.size() * sizeof(GatherEntry));
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(*out, ge.data(), ge.size() * sizeof(GatherEntry), hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        *count = (int)ge.size();
        return SWE_OK;
    }

    int init(swe_plan *p)
    {
        pl = p;
        d = p->d;
        order = p->cfg.order;
        int rc = 0;
        T **qs[] = {&qb, &qt, &ql, &qr};
        for (T **q : qs)
            if ((rc = alloc((void **)q, (d.N + 1) * sizeof(T)))) return rc;
        geo.hx = 0.5 * PI / d.N;
        geo.radius = p->cfg.radius;
        geo.omega = p->cfg.omega;
        for (int k = 0; k < 3; k++) geo.axis[k] = p->cfg.axis[k];
        geo.coriolis_const = p->cfg.coriolis_const;
        geo.order = order;
        k_sync_weights<T><<<blocks_for(d.N + 1), TPB, 0, pl->stream>>>(d.N, geo.hx, qb, qt, ql, qr);
        HIP_TRY(hipGetLastError());
        // sbp63 runs on the multi-kernel path (the fused kernels' regions are
        // sized for the sbp21 / sbp42 closures)
        use_fused = !(p->cfg.flags & SWE_FLAG_UNFUSED) && order != 63;
        nthr = p->cfg.block_threads ? p->cfg.block_threads : F_THREADS;
        if ((rc = alloc((void **)&part, 3 * INT_MAX_BLOCKS * sizeof(double))) ||
            (rc = alloc((void **)&red, 4 * sizeof(double))) ||
            (rc = alloc((void **)&dpart, (size_t)DIAG_K * INT_MAX_BLOCKS * sizeof(double))) ||
            (rc = alloc((void **)&dout, 10 * sizeof(double))))
            return rc;
        if (use_fused) rc = init_fused(geo);
        else rc = ensure_full();
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(pl->stream));
        return SWE_OK;
    }

    // The unfused path's per-panel metric fields, work fields and halo maps,
    // and its integral weights.
    int ensure_full()
    {
        if (full_ready) return SWE_OK;
        int rc = 0;
        T **cf[] = {&JQ12, &fJ2, &iJa, &Qu, &iJu, &Ju, &Qv, &iJv, &Jv,
                    &W1, &W2, &ut, &vt, &Jut, &Jvt, &up, &vp, &div};
        for (T **f : cf)
            if ((rc = field(f))) return rc;
        if ((rc = field(&wxy)) || (rc = field(&wu)) || (rc = field(&wv))) return rc;
        long nxy = 6L * (d.N + 1) * (d.N + 1), nuv = 6L * d.N * (d.N + 1);
        k_metric<T><<<blocks_for(nxy), TPB, 0, pl->stream>>>(d, geo, 0, JQ12, fJ2, iJa, wxy);
        k_metric<T><<<blocks_for(nuv), TPB, 0, pl->stream>>>(d, geo, 1, Qu, iJu, Ju, wu);
        k_metric<T><<<blocks_for(nuv), TPB, 0, pl->stream>>>(d, geo, 2, Qv, iJv, Jv, wv);
        HIP_TRY(hipGetLastError());
        // maps: field slots 0 and 1 are the (u-type, v-type) components
        if ((rc = upload_map(0, 1, {0, 0}, &maps.xy, &maps.n_xy))) return rc;
        if ((rc = upload_map(1, ch_depth(order), {0, 1}, &maps.ch, &maps.n_ch))) return rc;
        if ((rc = upload_map(2, 1, {0, 1}, &maps.xyv, &maps.n_xyv))) return rc;
        if ((rc = upload_map(3, 1, {0, 1}, &maps.tan, &maps.n_tan))) return rc;
        HIP_TRY(hipStreamSynchronize(pl->stream));
        full_ready = true;
        return SWE_OK;
    }

    // the plan's buffers for the partial sums and totals of swe_integrals
    int integral_buffers(double **pp, double **pr) override
    {
        *pp = part;
        *pr = red;
        return SWE_OK;
    }

    int init_fused(const GeomParams &g)
    {
        const int N = d.N, n = N + 1;
        int rc;
        // coefficient tables
        T *tabs[8];
        for (auto &t : tabs)
            if ((rc = alloc((void **)&t, (size_t)n * n * sizeof(T)))) return rc;
        T *tf;
        if ((rc = alloc((void **)&tf, 6 * (size_t)n * n * sizeof(T)))) return rc;
        k_coef<T><<<blocks_for(6L * n * n), TPB, 0, pl->stream>>>(N, g, tabs[0], tabs[1], tabs[2], tabs[3], tabs[4],
                                                                  tabs[5], tabs[6], tabs[7], tf);
        HIP_TRY(hipGetLastError());
        cf = Coef<T>{tabs[0], tabs[1], tabs[2], tabs[3], tabs[4], tabs[5], tabs[6], tabs[7], tf, qb, qt, ql, qr};
        for (auto &r : rec)
            if ((rc = alloc((void **)&r, (size_t)24 * REC_KINDS * n * sizeof(T)))) return rc;
        // the kernels' packed cube side table must match the topology
        for (int q = 0; q < 6; q++)
            for (int sdi = 0; sdi < 4; sdi++) {
                const swe::SideMatch m = swe::match_side(q, sdi);
                const SideDesc k = side_of(q, sdi);
                if (k.q != m.q || k.qs != m.q_side || k.rev != (m.reversed ? 1 : 0))
                    return fail(SWE_EINVAL, "internal: packed cube side table disagrees with the topology");
            }
        // tiles, edge segments and which tiles write into each segment's band
        FusedTables ft;
        std::string err;
        if (build_fused_tables(N, TY, ft, err, pl->cfg.tile_x0, pl->cfg.tile_y0)) return fail(SWE_EINVAL, err);
        ntx = ft.ntx;
        nty = ft.nty;
        fwx = ft.fwx;
        fwy = ft.fwy;
        ntiles = ft.ntiles;
        nseg = ft.nseg;
        const std::vector<int4> &segs = ft.segs;
        // the records kernels decode their segment from the block index
        for (int i = 0; i < nseg; i++) {
            const int4 a = segment_of<TY>(N, i), b = segs[i];
            if (a.x != b.x || a.y != b.y || a.z != b.z || a.w != b.w)
                return fail(SWE_EINVAL, "internal: edge segment decode disagrees with the segment table");
        }
        HIP_TRY(hipStreamSynchronize(pl->stream));
        return SWE_OK;
    }

    template <int O>
    StageArgs<T> stage_args(const T *y) const
    {
        StageArgs<T> a{};
        a.d = d;
        a.c = cf;
        a.y = y;
        const swe_config &c = pl->cfg;
        a.g_hxa = T(c.grav / (0.5 * PI / d.N * c.radius));
        a.H0 = T(c.h_mean);
        a.ntx = ntx;
        a.nty = nty;
        a.fwx = fwx;
        a.fwy = fwy;
        a.order = O;
        a.xcd = (pl->cfg.flags & SWE_FLAG_LINEAR_BLOCKS) ? 0 : 1;
        a.general = (pl->cfg.flags & SWE_FLAG_GENERAL_TILES) ? 1 : 0;
        return a;
    }

    template <int O>
    void launch_records(const T *y, T *r)
    {
        timer.launch(k_records<T, O, TY>, nseg, F_REC_THREADS, pl->stream, 1, d, cf, y, r);
    }

    template <int O>
    void fused_apply(T *out, const T *in)
    {
        launch_records<O>(in, rec[0]);
        StageArgs<T> a = stage_args<O>(in);
        a.out = out;
        a.rec_in = rec[0];
        timer.launch(k_stage<T, O, TY, 0>, ntiles, nthr, pl->stream, 0, a);
    }

    // nsteps RK4 steps, each: records of v0, then the four stages, each
    // followed (but the last) by the records of its output
    // compensation field of state v0 (compensated plans), else nullptr
    T *cmp_of(T *v0) const { return (pl->cfg.flags & SWE_FLAG_COMPENSATED) ? v0 + 3 * d.F : nullptr; }

    template <int O>
    void launch_steps(T *v0, T *acc, T *ya, T *yb, double dt, int nsteps)
    {
        T *cmp = cmp_of(v0);
        const T *ins[4] = {v0, ya, yb, ya};
        T *nexts[4] = {ya, yb, ya, nullptr};
        const double A[4] = {0.5 * dt, 0.5 * dt, dt, 0.0};
        const double B[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
        for (int it = 0; it < nsteps; it++) {
            launch_records<O>(v0, rec[0]);
            for (int st = 0; st < 4; st++) {
                StageArgs<T> a = stage_args<O>(ins[st]);
                a.v0 = v0;
                a.acc = acc;
                a.ynext = nexts[st];
                a.out = v0;
                a.cmp = cmp;
                a.a_dt = T(A[st]);
                a.b_dt = T(B[st]);
                a.rec_in = rec[st & 1];
                if (st == 0) timer.launch(k_stage<T, O, TY, 1>, ntiles, nthr, pl->stream, 0, a);
                else if (st == 3 && cmp) timer.launch(k_stage<T, O, TY, 4>, ntiles, nthr, pl->stream, 0, a);
                else if (st == 3) timer.launch(k_stage<T, O, TY, 3>, ntiles, nthr, pl->stream, 0, a);
                else timer.launch(k_stage<T, O, TY, 2>, ntiles, nthr, pl->stream, 0, a);
                if (st < 3) launch_records<O>(nexts[st], rec[(st + 1) & 1]);
            }
        }
    }

    // The step loop replayed from a captured HIP graph of GRAPH_STEPS steps
    // (the launches of a step cost more than their gaps otherwise); the graph
    // is kept for the last (state, registers, dt) it was made for.  Short runs
    // and timed runs (per-launch events) launch directly.
    static constexpr int GRAPH_STEPS = 64;
    hipGraphExec_t graph = nullptr;
    const void *graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
    double graph_dt = 0;

    template <int O>
    int fused_rk4(T *v0, T *acc, T *ya, T *yb, double dt, int nsteps)
    {
        if (timer.on || nsteps < GRAPH_STEPS) {
            launch_steps<O>(v0, acc, ya, yb, dt, nsteps);
            return SWE_OK;
        }
        const void *key[4] = {v0, acc, ya, yb};
        if (!graph || std::memcmp(key, graph_key, sizeof key) != 0 || graph_dt != dt) {
            if (graph) (void)hipGraphExecDestroy(graph);
            graph = nullptr;
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal));
            launch_steps<O>(v0, acc, ya, yb, dt, GRAPH_STEPS);
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(pl->stream, &g);
            if (le != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(SWE_EHIP, std::string("step graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
            }
            const hipError_t ie = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                graph = nullptr;
                return fail(SWE_EHIP, std::string("step graph instantiate: ") + hipGetErrorString(ie));
            }
            std::memcpy(graph_key, key, sizeof key);
            graph_dt = dt;
        }
        int done = 0;
        for (; done + GRAPH_STEPS <= nsteps; done += GRAPH_STEPS) HIP_TRY(hipGraphLaunch(graph, pl->stream));
        launch_steps<O>(v0, acc, ya, yb, dt, nsteps - done);
        return SWE_OK;
    }

    int rk4(void *v0, void *acc, void *ya, void *yb, double dt, int nsteps) override
    {
        const int rc = order == 21 ? fused_rk4<21>((T *)v0, (T *)acc, (T *)ya, (T *)yb, dt, nsteps)
                                   : fused_rk4<42>((T *)v0, (T *)acc, (T *)ya, (T *)yb, dt, nsteps);
        if (rc) return rc;
        HIP_TRY(hipGetLastError());
        return SWE_OK;
    }

    void gather(const GatherEntry *m, int n, void *f0, void *f1)
    {
        FieldSet fs{{f0, f1, nullptr, nullptr}};
        k_gather<T><<<blocks_for(n), TPB, 0, pl->stream>>>(m, n, fs);
    }

    template <int O>
    void apply_o(T *out, T *in)
    {
        const int N = d.N;
        T *u = in, *v = in + d.F, *h = in + 2 * d.F;
        T *du = out, *dv = out + d.F, *dh = out + 2 * d.F;
        long nxy = 6L * (N + 1) * (N + 1), nuv = 6L * N * (N + 1);
        hipStream_t s = pl->stream;
        gather(maps.xy, maps.n_xy, h, h);
        k_xy_w<T, O><<<blocks_for(nxy), TPB, 0, s>>>(d, u, v, JQ12, W1, W2);
        k_contra<T, O><<<blocks_for(2 * nuv), TPB, 0, s>>>(d, u, v, W1, W2, Qu, iJu, Ju, Qv, iJv, Jv, ut, vt, Jut,
                                                           Jvt);
        gather(maps.ch, maps.n_ch, Jut, Jvt);
        k_xy_cd<T, O><<<blocks_for(nxy), TPB, 0, s>>>(d, ut, vt, Jut, Jvt, fJ2, iJa, up, vp, div);
        gather(maps.xyv, maps.n_xyv, up, vp);
        gather(maps.xy, maps.n_xy, div, div);
        k_edge_sync<T><<<blocks_for(24L * N), TPB, 0, s>>>(d, up, vp, div, qb, qt, ql, qr);
        const swe_config &c = pl->cfg;
        T g_hxa = T(c.grav / (0.5 * PI / N * c.radius));
        k_tend<T, O><<<blocks_for(2 * nuv + nxy), TPB, 0, s>>>(d, h, up, vp, div, iJu, iJv, g_hxa, T(c.h_mean), du,
                                                               dv, dh);
    }

    int apply(void *out, void *in) override
    {
        if (use_fused) {
            if (order == 21) fused_apply<21>((T *)out, (const T *)in);
            else fused_apply<42>((T *)out, (const T *)in);
        } else if (order == 21) apply_o<21>((T *)out, (T *)in);
        else if (order == 42) apply_o<42>((T *)out, (T *)in);
        else apply_o<63>((T *)out, (T *)in);
        HIP_TRY(hipGetLastError());
        return SWE_OK;
    }

    int axpy(void *out, double s1, void *x, double s2, void *y, int update) override
    {
        long n = 3 * d.F;
        k_axpy<T><<<blocks_for(n), TPB, 0, pl->stream>>>(n, (T *)out, T(s1), (const T *)x, T(s2), (const T *)y,
                                                          update);
        HIP_TRY(hipGetLastError());
        return SWE_OK;
    }

    double *dpart = nullptr, *dout = nullptr;  // diagnostics: partial results, totals
    template <int O>
    void launch_diag(const T *st, double grav)
    {
        const int grid = std::min(ntiles, INT_MAX_BLOCKS);
        const double dx = geo.radius * geo.hx;
        k_diag<T, O, TY><<<grid, F_THREADS, 0, pl->stream>>>(d, cf, geo, st, ntx, nty, fwx, fwy, grav, 1.0 / (dx * dx),
                                                              dpart);
        k_diag_final<<<1, 64, 0, pl->stream>>>(dpart, grid, dout);
    }
    // u, v of a state copied into (ua, va) with their tangential duplicates
    // (the curl's SAT values) gathered from the neighbours
    void tan_copies(const T *st, T *ua, T *va)
    {
        k_axpy<T><<<blocks_for(d.F), TPB, 0, pl->stream>>>(d.F, ua, T(1), st, T(0), nullptr, 0);
        k_axpy<T><<<blocks_for(d.F), TPB, 0, pl->stream>>>(d.F, va, T(1), st + d.F, T(0), nullptr, 0);
        gather(maps.tan, maps.n_tan, ua, va);
    }
    // the multi-kernel path's contravariant components of a state into ut, vt
    template <int O>
    void contra_o(const T *u, const T *v)
    {
        const long nxy = 6L * (d.N + 1) * (d.N + 1), nuv = 6L * d.N * (d.N + 1);
        k_xy_w<T, O><<<blocks_for(nxy), TPB, 0, pl->stream>>>(d, u, v, JQ12, W1, W2);
        k_contra<T, O><<<blocks_for(2 * nuv), TPB, 0, pl->stream>>>(d, u, v, W1, W2, Qu, iJu, Ju, Qv, iJv, Jv, ut, vt,
                                                                     Jut, Jvt);
    }
    template <int O>
    void mk_diag(const T *st, double grav)
    {
        const double dx = geo.radius * geo.hx;
        contra_o<O>(st, st + d.F);
        tan_copies(st, up, vp);
        k_mk_diag<T, O><<<MK_BLOCKS, 256, 0, pl->stream>>>(d, geo, st, st + d.F, st + 2 * d.F, ut, vt, up, vp, grav,
                                                            1.0 / (dx * dx), dpart);
        k_diag_final<<<1, 64, 0, pl->stream>>>(dpart, MK_BLOCKS, dout);
    }
    // the diagnostics' sums of a state into dout (fused kernels for sbp21 /
    // sbp42 fused plans, the multi-kernel ones otherwise)
    int run_diag(const T *st, double grav)
    {
        if (use_fused) {
            if (order == 21) launch_diag<21>(st, grav);
            else launch_diag<42>(st, grav);
        } else if (order == 21) mk_diag<21>(st, grav);
        else if (order == 42) mk_diag<42>(st, grav);
        else mk_diag<63>(st, grav);
        HIP_TRY(hipGetLastError());
        return SWE_OK;
    }
    int integrals_diag(void *st, double grav, double out[7]) override
    {
        int rc = run_diag((const T *)st, grav);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(out, dout, 7 * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        return SWE_OK;
    }
    int cfl_diag(void *st, double dt, double grav, double out[3]) override
    {
        int rc = run_diag((const T *)st, grav);
        if (rc) return rc;
        double v[9];
        HIP_TRY(hipMemcpyAsync(v, dout, 9 * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        // sqrt of the largest g h / dx^2 (the largest h: the scaling is
        // monotonic), of the largest |ut u| / dx^2 and |vt v| / dy^2, times dt
        const double dx = geo.radius * geo.hx;
        out[0] = std::sqrt(grav / (dx * dx) * v[1]) * dt;
        out[1] = std::sqrt(v[7]) * dt;
        out[2] = std::sqrt(v[8]) * dt;
        return SWE_OK;
    }

    // lat-lon output: per kind (0 h, 1 winds) the tables of the last grid
    // asked for, the extended arrays and the outputs (built on the first
    // call for a grid)
    struct LL {
        int nlat = 0, nlon = 0, nrows = 0, np = 0, ew = 0;
        size_t nbytes = 0;
        int *rp = nullptr, *col = nullptr, *base = nullptr;
        double *w = nullptr, *fx = nullptr, *fy = nullptr, *q = nullptr, *ext = nullptr, *uv = nullptr;
        float *out = nullptr;
    } ll[2];
    int latlon_tables(int kind, int nlat, int nlon)
    {
        LL &L = ll[kind];
        if (L.nlat == nlat && L.nlon == nlon) return SWE_OK;
        swe::LatlonTables t;
        std::string err;
        if (kind == 0 ? swe::build_latlon_tables(d.N, nlat, nlon, t, err)
                      : swe::build_latlon_vec_tables(d.N, nlat, nlon, t, err))
            return fail(SWE_EINVAL, std::string(kind == 0 ? "swe_latlon_h: " : "swe_latlon_uv: ") + err);
        HIP_TRY(hipStreamSynchronize(pl->stream));
        void *old[] = {L.rp, L.col, L.base, L.w, L.fx, L.fy, L.q, L.ext, L.uv, L.out};
        for (void *q : old)
            if (q) release(q);
        bytes -= L.nbytes;
        L = LL();
        // h: sources are the state's h points (element offsets); winds:
        // the colocated components' packed indices as they are
        const int n = d.N + 1;
        std::vector<int> col(t.src.size());
        for (size_t k = 0; k < col.size(); k++) {
            const int s = t.src[k], p = s / (n * n), r = s - p * n * n, j = r / n + 1, i = r - (j - 1) * n + 1;
            col[k] = kind == 0 ? (int)off(d, p, i, j) : s;
        }
        const size_t nr = t.rp.size() - 1, np = t.base.size(), nnz = col.size(), nout = kind == 0 ? np : 2 * np;
        const size_t before = bytes;
        int rc;
        if ((rc = alloc((void **)&L.rp, (nr + 1) * sizeof(int))) || (rc = alloc((void **)&L.col, nnz * sizeof(int))) ||
            (rc = alloc((void **)&L.w, nnz * sizeof(double))) || (rc = alloc((void **)&L.base, np * sizeof(int))) ||
            (rc = alloc((void **)&L.fx, np * sizeof(double))) || (rc = alloc((void **)&L.fy, np * sizeof(double))) ||
            (rc = alloc((void **)&L.ext, nr * sizeof(double))) || (rc = alloc((void **)&L.out, nout * sizeof(float))))
            return rc;
        if (kind == 1 && ((rc = alloc((void **)&L.q, 4 * np * sizeof(double))) ||
                          (rc = alloc((void **)&L.uv, 2 * (size_t)6 * n * n * sizeof(double)))))
            return rc;
        HIP_TRY(hipMemcpyAsync(L.rp, t.rp.data(), (nr + 1) * sizeof(int), hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipMemcpyAsync(L.col, col.data(), nnz * sizeof(int), hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipMemcpyAsync(L.w, t.w.data(), nnz * sizeof(double), hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipMemcpyAsync(L.base, t.base.data(), np * sizeof(int), hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipMemcpyAsync(L.fx, t.fx.data(), np * sizeof(double), hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipMemcpyAsync(L.fy, t.fy.data(), np * sizeof(double), hipMemcpyHostToDevice, pl->stream));
        if (kind == 1)
            HIP_TRY(hipMemcpyAsync(L.q, t.q.data(), 4 * np * sizeof(double), hipMemcpyHostToDevice, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        L.nlat = nlat;
        L.nlon = nlon;
        L.nrows = (int)nr;
        L.np = (int)np;
        L.ew = t.ew;
        L.nbytes = bytes - before;
        return SWE_OK;
    }
    int latlon_h(void *st, int nlat, int nlon, float *out) override
    {
        int rc = latlon_tables(0, nlat, nlon);
        if (rc) return rc;
        LL &L = ll[0];
        const T *h = (const T *)st + 2 * d.F;
        k_ll_ext<T><<<(L.nrows + 255) / 256, 256, 0, pl->stream>>>(h, L.rp, L.col, L.w, L.nrows, L.ext);
        k_ll_interp<<<(L.np + 255) / 256, 256, 0, pl->stream>>>(L.ext, L.base, L.fx, L.fy, L.np, L.ew, L.out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, L.out, (size_t)L.np * sizeof(float), hipMemcpyDeviceToHost, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        return SWE_OK;
    }
    int latlon_uv(void *st, int nlat, int nlon, float *uo, float *vo) override
    {
        int rc = latlon_tables(1, nlat, nlon);
        if (rc) return rc;
        LL &L = ll[1];
        const T *u = (const T *)st, *v = u + d.F;
        const long nxy = 6L * (d.N + 1) * (d.N + 1);
        k_ll_colocate<T><<<(unsigned)((nxy + 255) / 256), 256, 0, pl->stream>>>(d, u, v, L.uv);
        k_ll_ext<double><<<(L.nrows + 255) / 256, 256, 0, pl->stream>>>(L.uv, L.rp, L.col, L.w, L.nrows, L.ext);
        k_ll_interp_uv<<<(L.np + 255) / 256, 256, 0, pl->stream>>>(L.ext, L.nrows / 2, L.base, L.fx, L.fy, L.q, L.np,
                                                                  L.ew, L.out, L.out + L.np);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(uo, L.out, (size_t)L.np * sizeof(float), hipMemcpyDeviceToHost, pl->stream));
        HIP_TRY(hipMemcpyAsync(vo, L.out + L.np, (size_t)L.np * sizeof(float), hipMemcpyDeviceToHost, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        return SWE_OK;
    }

    // the implicit solve: two apply states and their results, the CG
    // vectors r, p, Ap, b (h fields), dot partials and scalars (allocated on
    // the first call)
    T *sv_s = nullptr, *sv_t1 = nullptr, *sv_t2 = nullptr, *sv_r = nullptr, *sv_p = nullptr, *sv_ap = nullptr,
      *sv_b = nullptr;
    double *sv_part = nullptr, *sv_sc = nullptr, *sv_w = nullptr;  // sv_w: xy quadrature weights
    static constexpr int DOT_BLOCKS = 256;
    static constexpr int CG_BATCH = 8;  // CG iterations between convergence read-backs
    void xy_dot(const T *a, const T *b, double *out)
    {
        k_xy_dot<T><<<DOT_BLOCKS, 256, 0, pl->stream>>>(d, sv_w, a, b, sv_part);
        k_dot_final<<<1, 64, 0, pl->stream>>>(sv_part, DOT_BLOCKS, out);
    }
    // swm_helm_lin_operator_t%apply: az = z - gamma div co2contra grad z with
    // gamma = H g dt^2, through two applies of the linear operator: apply(0,
    // 0, z) has u, v = -g grad z, the next apply's h = -H div co2contra of
    // those = H g div co2contra grad z
    int helm(const T *z, T *az, double dt)
    {
        int rc;
        HIP_TRY(hipMemcpyAsync(sv_s + 2 * d.F, z, d.F * sizeof(T), hipMemcpyDeviceToDevice, pl->stream));
        if ((rc = apply(sv_t1, sv_s)) || (rc = apply(sv_t2, sv_t1))) return rc;
        k_axpy<T><<<blocks_for(d.F), TPB, 0, pl->stream>>>(d.F, az, T(1.0), z, T(-(dt * dt)), sv_t2 + 2 * d.F, 0);
        return SWE_OK;
    }
    // one CG iteration (cg_solver_mod.f90 L48-88); parity: which of the slots
    // 0 / 2 holds |r|^2 before the step (they alternate); every kernel
    // after convergence is a no-op (the state's flag)
    int cg_iter(T *x, double dt, double rel_tol, double *cst, int parity)
    {
        const long F = d.F;
        const int a = parity ? 2 : 0, b = parity ? 0 : 2;
        int rc;
        if ((rc = helm(sv_p, sv_ap, dt))) return rc;
        xy_dot(sv_p, sv_ap, sv_sc + 1);
        k_cg_step<T><<<blocks_for(F), TPB, 0, pl->stream>>>(F, x, sv_r, sv_p, sv_ap, sv_sc + a, sv_sc + 1, cst);
        xy_dot(sv_r, sv_r, sv_sc + b);
        k_cg_check<<<1, 64, 0, pl->stream>>>(sv_sc + b, sv_sc + 3, rel_tol, cst);
        k_cg_dir<T><<<blocks_for(F), TPB, 0, pl->stream>>>(F, sv_p, sv_r, sv_sc + b, sv_sc + a, cst);
        return SWE_OK;
    }
    // CG_BATCH iterations (an even number: the slots are back in place
    // after it) from a graph captured for the last (x, dt, rel_tol); the
    // iterations launch about a dozen small kernels each, whose launch gaps
    // bound the solve at small N
    hipGraphExec_t cg_graph = nullptr;
    const void *cg_key = nullptr;
    double cg_dt = 0, cg_tol = 0;
    static_assert(CG_BATCH % 2 == 0, "a CG batch must leave the |r|^2 slots in place");
    int cg_batch(T *x, double dt, double rel_tol, double *cst)
    {
        if (!cg_graph || cg_key != x || cg_dt != dt || cg_tol != rel_tol) {
            if (cg_graph) (void)hipGraphExecDestroy(cg_graph);
            cg_graph = nullptr;
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal));
            int rc = SWE_OK;
            for (int q = 0; q < CG_BATCH && rc == SWE_OK; q++) rc = cg_iter(x, dt, rel_tol, cst, q & 1);
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(pl->stream, &g);
            if (rc || le != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (rc) return rc;
                return fail(SWE_EHIP, std::string("CG graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
            }
            const hipError_t ie = hipGraphInstantiate(&cg_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                cg_graph = nullptr;
                return fail(SWE_EHIP, std::string("CG graph instantiate: ") + hipGetErrorString(ie));
            }
            cg_key = x;
            cg_dt = dt;
            cg_tol = rel_tol;
        }
        HIP_TRY(hipGraphLaunch(cg_graph, pl->stream));
        return SWE_OK;
    }
    int solve(void *vout, void *rhs, double dt, int max_iter, double rel_tol, int *iters, double *relres) override
    {
        int rc;
        if (!sv_s) {
            T **st3[] = {&sv_s, &sv_t1, &sv_t2};
            for (T **q : st3)
                if ((rc = alloc((void **)q, 3 * d.F * sizeof(T)))) return rc;
            T **f1[] = {&sv_r, &sv_p, &sv_ap, &sv_b};
            for (T **q : f1)
                if ((rc = field(q))) return rc;
            if ((rc = alloc((void **)&sv_part, DOT_BLOCKS * sizeof(double))) ||
                (rc = alloc((void **)&sv_sc, 8 * sizeof(double))) ||
                (rc = alloc((void **)&sv_w, d.F * sizeof(double))))
                return rc;
            const long nxy = 6L * (d.N + 1) * (d.N + 1);
            if (order == 21) k_xy_weights<21><<<blocks_for(nxy), TPB, 0, pl->stream>>>(d, geo.hx, sv_w);
            else if (order == 42) k_xy_weights<42><<<blocks_for(nxy), TPB, 0, pl->stream>>>(d, geo.hx, sv_w);
            else k_xy_weights<63><<<blocks_for(nxy), TPB, 0, pl->stream>>>(d, geo.hx, sv_w);
        }
        T *out = (T *)vout, *in = (T *)rhs, *x = out + 2 * d.F;
        const long F = d.F;
        // Helmholtz right-hand side b = rhs_h - H dt div co2contra(rhs_u, rhs_v)
        if ((rc = apply(sv_t1, in))) return rc;
        k_axpy<T><<<blocks_for(F), TPB, 0, pl->stream>>>(F, sv_b, T(1.0), in + 2 * F, T(dt), sv_t1 + 2 * F, 0);
        HIP_TRY(hipMemcpyAsync(x, sv_b, F * sizeof(T), hipMemcpyDeviceToDevice, pl->stream));
        if ((rc = helm(x, sv_ap, dt))) return rc;
        k_axpy<T><<<blocks_for(F), TPB, 0, pl->stream>>>(F, sv_r, T(-1.0), sv_ap, T(1.0), sv_b, 0);
        xy_dot(sv_r, sv_r, sv_sc + 0);
        xy_dot(sv_b, sv_b, sv_sc + 3);
        double h2[4];
        HIP_TRY(hipMemcpyAsync(h2, sv_sc, 4 * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        const double bn = std::sqrt(h2[3]);
        // a zero right-hand side has the solution zero (x = b = 0 already;
        // the reference's 0 / 0 would go on with NaN)
        double rel = bn == 0.0 ? 0.0 : std::sqrt(h2[0]) / bn;
        int it = 0;
        if (!(rel < rel_tol)) {
            HIP_TRY(hipMemcpyAsync(sv_p, sv_r, F * sizeof(T), hipMemcpyDeviceToDevice, pl->stream));
            double *cst = sv_sc + 4;  // convergence state (k_cg_check)
            k_fill<<<1, 64, 0, pl->stream>>>(cst, 4, 0.0);
            bool done = false;
            it = max_iter + 1;  // not converged (also when max_iter is 0)
            // batches of CG_BATCH iterations (replayed from a graph), the
            // rest one by one; the host looks at the convergence state after
            // each batch and after the last iteration
            for (int k = 0; k < max_iter && !done;) {
                const int nb = std::min(CG_BATCH, max_iter - k);
                if (nb == CG_BATCH && !timer.on) {
                    if ((rc = cg_batch(x, dt, rel_tol, cst))) return rc;
                } else {
                    for (int q = 0; q < nb; q++)
                        if ((rc = cg_iter(x, dt, rel_tol, cst, (k + q) & 1))) return rc;
                }
                k += nb;
                double cs[3];
                HIP_TRY(hipMemcpyAsync(cs, cst, 3 * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
                HIP_TRY(hipStreamSynchronize(pl->stream));
                done = cs[0] != 0.0;
                rel = cs[2];
                it = done ? (int)cs[1] : k + 1;
            }
        }
        if (iters) *iters = it;
        if (relres) *relres = rel;
        // u, v = rhs_u, rhs_v - g dt grad h
        HIP_TRY(hipMemcpyAsync(sv_s + 2 * F, x, F * sizeof(T), hipMemcpyDeviceToDevice, pl->stream));
        if ((rc = apply(sv_t1, sv_s))) return rc;
        k_axpy<T><<<blocks_for(2 * F), TPB, 0, pl->stream>>>(2 * F, out, T(1.0), in, T(dt), sv_t1, 0);
        HIP_TRY(hipGetLastError());
        return SWE_OK;
    }

    T *tbuf = nullptr;  // the tendency of swe_integral_tends (allocated on its first call)
    template <int O>
    void launch_tends(const T *st, double grav)
    {
        const int grid = std::min(ntiles, INT_MAX_BLOCKS);
        k_tends<T, O, TY><<<grid, F_THREADS, 0, pl->stream>>>(d, cf, geo, st, tbuf, ntx, nty, fwx, fwy, grav, dpart);
        k_tends_final<<<1, 64, 0, pl->stream>>>(dpart, grid, dout);
    }
    template <int O>
    void mk_tends(const T *st, double grav)
    {
        const T *t = tbuf;
        contra_o<O>(st, st + d.F);  // ut, vt; W1, W2 are free after it
        tan_copies(st, up, vp);
        tan_copies(t, W1, W2);
        k_mk_tends<T, O><<<MK_BLOCKS, 256, 0, pl->stream>>>(d, geo, st, st + d.F, st + 2 * d.F, ut, vt, up, vp, t,
                                                             t + d.F, t + 2 * d.F, W1, W2, grav, dpart);
        k_tends_final<<<1, 64, 0, pl->stream>>>(dpart, MK_BLOCKS, dout);
    }
    int integral_tends(void *st, double grav, double out[2]) override
    {
        int rc;
        if (!tbuf && (rc = alloc((void **)&tbuf, 3 * (size_t)d.F * sizeof(T)))) return rc;
        if ((rc = apply(tbuf, st))) return rc;
        if (!use_fused) {
            if (order == 21) mk_tends<21>((const T *)st, grav);
            else if (order == 42) mk_tends<42>((const T *)st, grav);
            else mk_tends<63>((const T *)st, grav);
        } else if (order == 21) launch_tends<21>((const T *)st, grav);
        else launch_tends<42>((const T *)st, grav);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, dout, 2 * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
        HIP_TRY(hipStreamSynchronize(pl->stream));
        return SWE_OK;
    }

    template <int O>
    void fused_integrals(const T *st, double *prt, double *rd)
    {
        const int grid = std::min(ntiles, INT_MAX_BLOCKS);
        k_integrals<T, O, TY><<<grid, F_THREADS, 0, pl->stream>>>(d, cf, st, ntx, nty, fwx, fwy, prt);
        k_reduce_final<<<1, 64, 0, pl->stream>>>(prt, grid, rd);
    }

    // FP64 integrals of state st; prt / rd: device buffers for the partial
    // sums (3 INT_MAX_BLOCKS doubles) and the totals (4 doubles).  The fused
    // path computes ut, vt tile by tile (no per-panel fields); the unfused
    // path runs its own co2contra kernels into its work fields.
    int integrals(void *st, double *prt, double *rd, double *mass, double *energy) override
    {
        const int N = d.N;
        T *u = (T *)st, *v = u + d.F, *h = u + 2 * d.F;
        hipStream_t s = pl->stream;
        if (use_fused) {
            if (order == 21) fused_integrals<21>(u, prt, rd);
            else fused_integrals<42>(u, prt, rd);
        } else {
            int rc = ensure_full();
            if (rc) return rc;
            long nxy = 6L * (N + 1) * (N + 1), nuv = 6L * N * (N + 1);
            if (order == 21) {
                k_xy_w<T, 21><<<blocks_for(nxy), TPB, 0, s>>>(d, u, v, JQ12, W1, W2);
                k_contra<T, 21><<<blocks_for(2 * nuv), TPB, 0, s>>>(d, u, v, W1, W2, Qu, iJu, Ju, Qv, iJv, Jv, ut,
                                                                    vt, Jut, Jvt);
            } else if (order == 42) {
                k_xy_w<T, 42><<<blocks_for(nxy), TPB, 0, s>>>(d, u, v, JQ12, W1, W2);
                k_contra<T, 42><<<blocks_for(2 * nuv), TPB, 0, s>>>(d, u, v, W1, W2, Qu, iJu, Ju, Qv, iJv, Jv, ut,
                                                                    vt, Jut, Jvt);
            } else {
                k_xy_w<T, 63><<<blocks_for(nxy), TPB, 0, s>>>(d, u, v, JQ12, W1, W2);
                k_contra<T, 63><<<blocks_for(2 * nuv), TPB, 0, s>>>(d, u, v, W1, W2, Qu, iJu, Ju, Qv, iJv, Jv, ut,
                                                                    vt, Jut, Jvt);
            }
            k_reduce<T><<<nblk_red, TPB, 0, s>>>(d, h, u, v, ut, vt, wxy, wu, wv, prt);
            k_reduce_final<<<1, 64, 0, s>>>(prt, nblk_red, rd);
        }
        HIP_TRY(hipGetLastError());
        double r[3];
        HIP_TRY(hipMemcpyAsync(r, rd, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const swe_config &c = pl->cfg;
        if (mass) *mass = r[0];
        if (energy) *energy = 0.5 * c.grav * r[1] + 0.5 * c.h_mean * r[2];
        return SWE_OK;
    }
};

int check_state(swe_plan *plan, swe_state *s, const char *what)
{
    if (!s) return fail(SWE_EINVAL, std::string(what) + " is NULL");
    if (s->plan != plan) return fail(SWE_EINVAL, std::string(what) + " belongs to another plan");
    return SWE_OK;
}

// The checks swe_plan_create applies to a configuration (also run by
// swe_server_submit, so that a request that could never be admitted is
// refused when it is submitted).
int validate_config(const swe_config *cfg, const char *who)
{
    const std::string w = std::string(who) + ": ";
    if (!cfg) return fail(SWE_EINVAL, w + "NULL configuration");
    if (cfg->order != 21 && cfg->order != 42 && cfg->order != 63)
        return fail(SWE_EINVAL, w + "order must be 21, 42 or 63");
    if (cfg->order == 21 && cfg->N < 4) return fail(SWE_EINVAL, w + "sbp21 needs N >= 4");
    if (cfg->order == 42 && cfg->N < 8) return fail(SWE_EINVAL, w + "sbp42 needs N >= 8");
    if (cfg->order == 63 && cfg->N < 12) return fail(SWE_EINVAL, w + "sbp63 needs N >= 12");
    if (cfg->order == 63 && (cfg->flags & SWE_FLAG_COMPENSATED))
        return fail(SWE_EINVAL, w + "compensated summation runs on the fused path (sbp21 / sbp42 only)");
    if (cfg->order == 63 && (cfg->tile_x0 || cfg->tile_y0))
        return fail(SWE_EINVAL, w + "tile_x0 / tile_y0 apply to the fused path (sbp21 / sbp42 only)");
    if (cfg->N > 4096) return fail(SWE_EINVAL, w + "N above 4096 is not supported");
    if (cfg->precision != SWE_FP64 && cfg->precision != SWE_FP32)
        return fail(SWE_EINVAL, w + "precision must be SWE_FP64 or SWE_FP32");
    if (cfg->block_threads != 0 && cfg->block_threads != F_THREADS)
        return fail(SWE_EINVAL, w + "block_threads must be 0 or 512 (the fused kernels' workgroup size)");
    if (cfg->flags & ~(SWE_FLAG_UNFUSED | SWE_FLAG_LINEAR_BLOCKS | SWE_FLAG_GENERAL_TILES | SWE_FLAG_COMPENSATED))
        return fail(SWE_EINVAL, w + "unknown bits in flags");
    if ((cfg->flags & SWE_FLAG_COMPENSATED) && (cfg->flags & SWE_FLAG_UNFUSED))
        return fail(SWE_EINVAL, w + "compensated summation runs on the fused path (not with SWE_FLAG_UNFUSED)");
    if (cfg->tile_x0 || cfg->tile_y0) {
        if (cfg->flags & SWE_FLAG_UNFUSED) return fail(SWE_EINVAL, w + "tile_x0 / tile_y0 apply to the fused path");
        FusedTables ft;
        std::string err;
        if (build_fused_tables(cfg->N, cfg->precision == SWE_FP32 ? 16 : 8, ft, err, cfg->tile_x0, cfg->tile_y0))
            return fail(SWE_EINVAL, w + err);
    }
    if (!(cfg->radius > 0) || !std::isfinite(cfg->radius) || !std::isfinite(cfg->grav) ||
        !std::isfinite(cfg->h_mean) || !std::isfinite(cfg->omega) || !std::isfinite(cfg->axis[0]) ||
        !std::isfinite(cfg->axis[1]) || !std::isfinite(cfg->axis[2]))
        return fail(SWE_EINVAL, w + "non-finite or non-positive physical constant");
    return SWE_OK;
}

}  // namespace

namespace {
template <typename T>
void pack_to_padded(const Dims &d, const double *src, int nx, int ny, T *dst)
{
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                dst[off(d, p, i, j)] = T(src[((size_t)p * ny + (j - 1)) * nx + (i - 1)]);
}
template <typename T>
void padded_to_pack(const Dims &d, const T *src, int nx, int ny, double *dst)
{
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                dst[((size_t)p * ny + (j - 1)) * nx + (i - 1)] = double(src[off(d, p, i, j)]);
}
template <typename T>
int upload_t(swe_state *st, const double *u, const double *v, const double *h)
{
    const Dims &d = st->plan->d;
    int N = d.N;
    std::vector<T> host(3 * d.F, T(0));
    pack_to_padded(d, u, N, N + 1, host.data());
    pack_to_padded(d, v, N + 1, N, host.data() + d.F);
    pack_to_padded(d, h, N + 1, N + 1, host.data() + 2 * d.F);
    HIP_TRY(hipMemcpyAsync(st->buf, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice,
                           st->plan->stream));
    // a new value carries no compensation
    if (st->plan->cfg.flags & SWE_FLAG_COMPENSATED)
        HIP_TRY(hipMemsetAsync((T *)st->buf + 3 * d.F, 0, 3 * d.F * sizeof(T), st->plan->stream));
    HIP_TRY(hipStreamSynchronize(st->plan->stream));
    return SWE_OK;
}
template <typename T>
int download_t(swe_state *st, double *u, double *v, double *h)
{
    const Dims &d = st->plan->d;
    int N = d.N;
    std::vector<T> host(3 * d.F);
    HIP_TRY(hipMemcpyAsync(host.data(), st->buf, host.size() * sizeof(T), hipMemcpyDeviceToHost,
                           st->plan->stream));
    HIP_TRY(hipStreamSynchronize(st->plan->stream));
    if (u) padded_to_pack(d, host.data(), N, N + 1, u);
    if (v) padded_to_pack(d, host.data() + d.F, N + 1, N, v);
    if (h) padded_to_pack(d, host.data() + 2 * d.F, N + 1, N + 1, h);
    return SWE_OK;
}

// the reference's tile arrays with a halo of width hw (grid_field_t tile
// p(1-hw:nx+hw, 1-hw:ny+hw), one tile per panel, panels in order) <-> packed
void halo_to_pack(const double *src, int nx, int ny, int hw, double *dst)
{
    const long sx = nx + 2 * hw, sy = ny + 2 * hw;
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                dst[((long)p * ny + j - 1) * nx + i - 1] = src[(p * sy + j - 1 + hw) * sx + i - 1 + hw];
}
void pack_to_halo(const double *src, int nx, int ny, int hw, double *dst)
{
    const long sx = nx + 2 * hw, sy = ny + 2 * hw;
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                dst[(p * sy + j - 1 + hw) * sx + i - 1 + hw] = src[((long)p * ny + j - 1) * nx + i - 1];
}
}  // namespace

namespace swe {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace swe

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

const char *swe_last_error(void) { return g_err.c_str(); }
const char *swe_version(void) { return "ch_swe_amd 0.1 (gfx950)"; }

int swe_plan_create(const swe_config *cfg, swe_plan **out)
{
    if (!cfg || !out) return fail(SWE_EINVAL, "swe_plan_create: NULL argument");
    *out = nullptr;
    int rc0 = validate_config(cfg, "swe_plan_create");
    if (rc0) return rc0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SWE_EHIP, "no HIP device available");
    auto *p = new swe_plan();
    p->cfg = *cfg;
    p->d = make_dims(cfg->N);
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete p;
        return fail(SWE_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    int rc;
    if (cfg->precision == SWE_FP64) {
        auto *en = new Engine<double>();
        p->eng.reset(en);
        rc = en->init(p);
    } else {
        auto *en = new Engine<float>();
        p->eng.reset(en);
        rc = en->init(p);
    }
    if (rc != SWE_OK) {
        std::string msg = g_err;
        swe_plan_destroy(p);
        return fail(rc, msg);
    }
    *out = p;
    return SWE_OK;
}

int swe_plan_destroy(swe_plan *plan)
{
    if (!plan) return SWE_OK;
    if (plan->stream) (void)hipStreamSynchronize(plan->stream);
    for (auto *&s : plan->rk) {
        swe_state_destroy(s);
        s = nullptr;
    }
    swe_state_destroy(plan->imp);
    plan->imp = nullptr;
    if (plan->mk_graph) (void)hipGraphExecDestroy(plan->mk_graph);
    plan->eng.reset();
    if (plan->stream) (void)hipStreamDestroy(plan->stream);
    delete plan;
    return SWE_OK;
}

int swe_plan_npts(const swe_plan *plan, int *n_u, int *n_v, int *n_h)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int N = plan->d.N;
    if (n_u) *n_u = 6 * N * (N + 1);
    if (n_v) *n_v = 6 * N * (N + 1);
    if (n_h) *n_h = 6 * (N + 1) * (N + 1);
    return SWE_OK;
}

void *swe_plan_stream(swe_plan *plan) { return plan ? (void *)plan->stream : nullptr; }

int swe_synchronize(swe_plan *plan)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    HIP_TRY(hipStreamSynchronize(plan->stream));
    return SWE_OK;
}

int swe_state_create(swe_plan *plan, swe_state **out)
{
    if (!plan || !out) return fail(SWE_EINVAL, "swe_state_create: NULL argument");
    // u, v, h, and with compensated summation their compensation fields
    const size_t nfields = (plan->cfg.flags & SWE_FLAG_COMPENSATED) ? 6 : 3;
    size_t bytes = nfields * plan->d.F * plan->eng->elem();
    void *buf = nullptr;
    hipError_t e = hipMalloc(&buf, bytes);
    if (e != hipSuccess) return fail(SWE_ENOMEM, std::string("hipMalloc(state): ") + hipGetErrorString(e));
    HIP_TRY(hipMemsetAsync(buf, 0, bytes, plan->stream));
    *out = new swe_state{plan, buf};
    return SWE_OK;
}

int swe_state_destroy(swe_state *st)
{
    if (!st) return SWE_OK;
    if (st->plan && st->plan->stream) (void)hipStreamSynchronize(st->plan->stream);
    if (st->buf) (void)hipFree(st->buf);
    delete st;
    return SWE_OK;
}

int swe_state_upload(swe_state *st, const double *u, const double *v, const double *h)
{
    if (!st || !u || !v || !h) return fail(SWE_EINVAL, "swe_state_upload: NULL argument");
    return st->plan->cfg.precision == SWE_FP64 ? upload_t<double>(st, u, v, h) : upload_t<float>(st, u, v, h);
}

int swe_state_download(swe_state *st, double *u, double *v, double *h)
{
    if (!st) return fail(SWE_EINVAL, "swe_state_download: NULL state");
    return st->plan->cfg.precision == SWE_FP64 ? download_t<double>(st, u, v, h) : download_t<float>(st, u, v, h);
}

int swe_state_upload_halo(swe_state *st, int hw, const double *u, const double *v, const double *h)
{
    if (!st || !u || !v || !h) return fail(SWE_EINVAL, "swe_state_upload_halo: NULL argument");
    if (hw < 0 || hw > 64) return fail(SWE_EINVAL, "swe_state_upload_halo: halo width out of range 0..64");
    const int N = st->plan->d.N;
    std::vector<double> pu((size_t)6 * N * (N + 1)), pv(pu.size()), ph((size_t)6 * (N + 1) * (N + 1));
    halo_to_pack(u, N, N + 1, hw, pu.data());
    halo_to_pack(v, N + 1, N, hw, pv.data());
    halo_to_pack(h, N + 1, N + 1, hw, ph.data());
    return swe_state_upload(st, pu.data(), pv.data(), ph.data());
}

int swe_state_download_halo(swe_state *st, int hw, double *u, double *v, double *h)
{
    if (!st) return fail(SWE_EINVAL, "swe_state_download_halo: NULL state");
    if (hw < 0 || hw > 64) return fail(SWE_EINVAL, "swe_state_download_halo: halo width out of range 0..64");
    const int N = st->plan->d.N;
    std::vector<double> pu((size_t)6 * N * (N + 1)), pv(pu.size()), ph((size_t)6 * (N + 1) * (N + 1));
    int rc = swe_state_download(st, pu.data(), pv.data(), ph.data());
    if (rc) return rc;
    if (u) pack_to_halo(pu.data(), N, N + 1, hw, u);
    if (v) pack_to_halo(pv.data(), N + 1, N, hw, v);
    if (h) pack_to_halo(ph.data(), N + 1, N + 1, hw, h);
    return SWE_OK;
}

int swe_apply(swe_plan *plan, swe_state *vout, swe_state *vin)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int rc;
    if ((rc = check_state(plan, vout, "vout")) || (rc = check_state(plan, vin, "vin"))) return rc;
    if (vout == vin) return fail(SWE_EINVAL, "swe_apply: vout and vin must differ");
    return plan->eng->apply(vout->buf, vin->buf);
}

static int axpy_common(swe_plan *plan, swe_state *out, double s1, swe_state *v1, double s2, swe_state *v2,
                       int update)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int rc;
    if ((rc = check_state(plan, out, "out")) || (rc = check_state(plan, v1, "v1"))) return rc;
    if (v2 && (rc = check_state(plan, v2, "v2"))) return rc;
    // assign gives out a new value, which carries no compensation; update
    // adds to the value v0 + cmp stands for, so its compensation stays
    if (!update && (plan->cfg.flags & SWE_FLAG_COMPENSATED)) {
        const size_t n = 3 * plan->d.F * plan->eng->elem();
        HIP_TRY(hipMemsetAsync((char *)out->buf + n, 0, n, plan->stream));
    }
    return plan->eng->axpy(out->buf, s1, v1->buf, s2, v2 ? v2->buf : nullptr, update);
}

int swe_assign(swe_plan *plan, swe_state *out, double s1, swe_state *v1, double s2, swe_state *v2)
{
    return axpy_common(plan, out, s1, v1, s2, v2, 0);
}

int swe_update(swe_plan *plan, swe_state *out, double s1, swe_state *v1, double s2, swe_state *v2)
{
    return axpy_common(plan, out, s1, v1, s2, v2, 1);
}

int swe_rk4_steps(swe_plan *plan, swe_state *v0, double dt, int nsteps)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int rc = check_state(plan, v0, "v0");
    if (rc) return rc;
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    if (!std::isfinite(dt)) return fail(SWE_EINVAL, "dt is not finite");
    if (nsteps == 0) return SWE_OK;
    // the RK4 registers are made on first use: the fused path keeps three
    // (the increment accumulator and two stage states), the unfused one the
    // reference's five (k1..k4, y)
    const int nreg = plan->eng->fused() ? 3 : 5;
    for (int k = 0; k < nreg; k++)
        if (!plan->rk[k] && (rc = swe_state_create(plan, &plan->rk[k]))) return rc;
    swe_state *k1 = plan->rk[0], *k2 = plan->rk[1], *k3 = plan->rk[2], *k4 = plan->rk[3], *y = plan->rk[4];
    if (plan->eng->fused()) return plan->eng->rk4(v0->buf, k1->buf, k2->buf, k3->buf, dt, nsteps);
    // rk4_mod.f90 step_rk4 L28-40: about 45 launches a step on the
    // multi-kernel path, so chunks of MK_GRAPH_STEPS steps are replayed from a
    // captured graph (no memset nodes); the rest and timed runs launch directly
    auto step = [&]() -> int {
        int r;
        if ((r = swe_apply(plan, k1, v0))) return r;
        if ((r = swe_assign(plan, y, 1.0, v0, 0.5 * dt, k1))) return r;
        if ((r = swe_apply(plan, k2, y))) return r;
        if ((r = swe_assign(plan, y, 1.0, v0, 0.5 * dt, k2))) return r;
        if ((r = swe_apply(plan, k3, y))) return r;
        if ((r = swe_assign(plan, y, 1.0, v0, dt, k3))) return r;
        if ((r = swe_apply(plan, k4, y))) return r;
        if ((r = swe_update(plan, v0, dt / 6.0, k1, dt / 3.0, k2))) return r;
        return swe_update(plan, v0, dt / 3.0, k3, dt / 6.0, k4);
    };
    constexpr int MK_GRAPH_STEPS = 16;
    int done = 0;
    if (!plan->eng->timer.on && nsteps >= MK_GRAPH_STEPS) {
        if (!plan->mk_graph || plan->mk_key != v0->buf || plan->mk_dt != dt) {
            if (plan->mk_graph) (void)hipGraphExecDestroy(plan->mk_graph);
            plan->mk_graph = nullptr;
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(plan->stream, hipStreamCaptureModeThreadLocal));
            for (int it = 0; it < MK_GRAPH_STEPS && rc == SWE_OK; it++) rc = step();
            const hipError_t ce = hipStreamEndCapture(plan->stream, &g);
            if (rc || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : fail(SWE_EHIP, std::string("RK4 graph capture: ") + hipGetErrorString(ce));
            }
            const hipError_t ie = hipGraphInstantiate(&plan->mk_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                plan->mk_graph = nullptr;
                return fail(SWE_EHIP, std::string("RK4 graph instantiate: ") + hipGetErrorString(ie));
            }
            plan->mk_key = v0->buf;
            plan->mk_dt = dt;
        }
        for (; done + MK_GRAPH_STEPS <= nsteps; done += MK_GRAPH_STEPS)
            HIP_TRY(hipGraphLaunch(plan->mk_graph, plan->stream));
    }
    for (; done < nsteps; done++)
        if ((rc = step())) return rc;
    return SWE_OK;
}

int swe_plan_timing(swe_plan *plan, int enable)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    double ms[2];
    int cnt[2];
    int rc = plan->eng->timer.collect(ms, cnt);  // drop what was recorded before
    plan->eng->timer.on = enable != 0;
    return rc;
}

int swe_plan_kernel_times(swe_plan *plan, double *stage_ms, int *n_stage, double *records_ms, int *n_records)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    double ms[2];
    int cnt[2];
    int rc = plan->eng->timer.collect(ms, cnt);
    if (rc) return rc;
    if (stage_ms) *stage_ms = ms[0];
    if (n_stage) *n_stage = cnt[0];
    if (records_ms) *records_ms = ms[1];
    if (n_records) *n_records = cnt[1];
    return SWE_OK;
}

int swe_integrals(swe_plan *plan, swe_state *st, double *mass, double *energy)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int rc = check_state(plan, st, "state");
    if (rc) return rc;
    double *part = nullptr, *red = nullptr;
    if ((rc = plan->eng->integral_buffers(&part, &red))) return rc;
    return plan->eng->integrals(st->buf, part, red, mass, energy);
}

int swe_integrals_diag(swe_plan *plan, swe_state *st, double grav, double out[7])
{
    if (!plan || !out) return fail(SWE_EINVAL, "swe_integrals_diag: NULL argument");
    if (!std::isfinite(grav)) return fail(SWE_EINVAL, "swe_integrals_diag: non-finite gravity");
    int rc = check_state(plan, st, "state");
    if (rc) return rc;
    return plan->eng->integrals_diag(st->buf, grav, out);
}

int swe_solve(swe_plan *plan, swe_state *vout, swe_state *rhs, double dt, int max_iter, double rel_tol, int *iters,
              double *relres)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int rc;
    if ((rc = check_state(plan, vout, "vout")) || (rc = check_state(plan, rhs, "rhs"))) return rc;
    if (vout == rhs) return fail(SWE_EINVAL, "swe_solve: vout and rhs must differ");
    if (!std::isfinite(dt) || !(rel_tol > 0) || max_iter < 0)
        return fail(SWE_EINVAL, "swe_solve: need a finite dt, rel_tol > 0 and max_iter >= 0");
    return plan->eng->solve(vout->buf, rhs->buf, dt, max_iter, rel_tol, iters, relres);
}

// implicit_Eul_mod.f90 step_implicit_Eul L20-30 / crank_nikolson_mod.f90
// step_crank_nikolson L20-32, nsteps times; *iters, *relres: the largest
// CG iteration count and final relative residual over the steps
static int implicit_steps(swe_plan *plan, swe_state *v0, double dt, int nsteps, int max_iter, double rel_tol,
                          int *iters, double *relres, bool cn)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int rc = check_state(plan, v0, "v0");
    if (rc) return rc;
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    if (!plan->imp && (rc = swe_state_create(plan, &plan->imp))) return rc;
    swe_state *vn = plan->imp;
    int itmax = 0;
    double rmax = 0;
    for (int k = 0; k < nsteps; k++) {
        int it = 0;
        double rr = 0;
        if (cn) {
            if ((rc = swe_apply(plan, vn, v0)) || (rc = swe_assign(plan, vn, 0.5 * dt, vn, 1.0, v0)) ||
                (rc = swe_solve(plan, v0, vn, 0.5 * dt, max_iter, rel_tol, &it, &rr)))
                return rc;
        } else {
            if ((rc = swe_solve(plan, vn, v0, dt, max_iter, rel_tol, &it, &rr)) ||
                (rc = swe_assign(plan, v0, 1.0, vn, 0.0, nullptr)))
                return rc;
        }
        itmax = std::max(itmax, it);
        rmax = std::max(rmax, rr);
    }
    if (iters) *iters = itmax;
    if (relres) *relres = rmax;
    return SWE_OK;
}

int swe_implicit_euler_steps(swe_plan *plan, swe_state *v0, double dt, int nsteps, int max_iter, double rel_tol,
                             int *iters, double *relres)
{
    return implicit_steps(plan, v0, dt, nsteps, max_iter, rel_tol, iters, relres, false);
}

int swe_crank_nicolson_steps(swe_plan *plan, swe_state *v0, double dt, int nsteps, int max_iter, double rel_tol,
                             int *iters, double *relres)
{
    return implicit_steps(plan, v0, dt, nsteps, max_iter, rel_tol, iters, relres, true);
}

int swe_cfl_diag(swe_plan *plan, swe_state *st, double dt, double grav, double out[3])
{
    if (!plan || !out) return fail(SWE_EINVAL, "swe_cfl_diag: NULL argument");
    if (!std::isfinite(grav) || !std::isfinite(dt)) return fail(SWE_EINVAL, "swe_cfl_diag: non-finite dt or gravity");
    int rc = check_state(plan, st, "state");
    if (rc) return rc;
    return plan->eng->cfl_diag(st->buf, dt, grav, out);
}

int swe_latlon_uv(swe_plan *plan, swe_state *st, int nlat, int nlon, float *u_out, float *v_out)
{
    if (!plan || !u_out || !v_out) return fail(SWE_EINVAL, "swe_latlon_uv: NULL argument");
    int rc = check_state(plan, st, "state");
    if (rc) return rc;
    return plan->eng->latlon_uv(st->buf, nlat, nlon, u_out, v_out);
}

int swe_latlon_h(swe_plan *plan, swe_state *st, int nlat, int nlon, float *out)
{
    if (!plan || !out) return fail(SWE_EINVAL, "swe_latlon_h: NULL argument");
    int rc = check_state(plan, st, "state");
    if (rc) return rc;
    return plan->eng->latlon_h(st->buf, nlat, nlon, out);
}

int swe_integral_tends(swe_plan *plan, swe_state *st, double grav, double out[2])
{
    if (!plan || !out) return fail(SWE_EINVAL, "swe_integral_tends: NULL argument");
    if (!std::isfinite(grav)) return fail(SWE_EINVAL, "swe_integral_tends: non-finite gravity");
    int rc = check_state(plan, st, "state");
    if (rc) return rc;
    return plan->eng->integral_tends(st->buf, grav, out);
}

}  // extern "C"

// ---- swe_integrals_noalloc: the partial sums in a static device array ----------
// The array holds the partial sums of one swe_integrals call (3 per block of
// the integral kernels, at most INT_MAX_BLOCKS blocks) and its 3 totals:
// 24.6 KB, independent of N, on every device the library runs on until the
// process ends.  Every block writes its partial sums before the final sum
// reads them, so nothing needs clearing between calls.  Calls that overlap in
// time on one device share the array and race.
static __device__ __align__(256) double scratch[3 * INT_MAX_BLOCKS + 4];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[SCRATCH_MAX_DEVICES] = {};  // device address of scratch, per device

// the device address of scratch on device dev (looked up once per device,
// with that device current)
int scratch_on(int dev, double **out)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(SWE_EINVAL, "swe_integrals_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        int prev = 0;
        HIP_TRY(hipGetDevice(&prev));
        if (prev != dev) HIP_TRY(hipSetDevice(dev));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (prev != dev) HIP_TRY(hipSetDevice(prev));
        if (e != hipSuccess) return fail(SWE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = (double *)ptr;
    }
    *out = g_scratch_dev[dev];
    return SWE_OK;
}
}  // namespace

extern "C" {

int swe_integrals_noalloc(swe_plan *plan, swe_state *st, double *mass, double *energy)
{
    if (!plan) return fail(SWE_EINVAL, "NULL plan");
    int rc = check_state(plan, st, "state");
    if (rc) return rc;
    double *buf = nullptr;
    if ((rc = scratch_on(plan->device, &buf))) return rc;
    return plan->eng->integrals(st->buf, buf, buf + 3 * INT_MAX_BLOCKS, mass, energy);
}

int swe_halo_map(int N, int order, int kind, int *count, int *tp, int *ti, int *tj, int *tc, int *sp, int *si,
                 int *sj, int *sc, int *sign)
{
    if (!count) return fail(SWE_EINVAL, "NULL count");
    if (N < 1 || (order != 21 && order != 42 && order != 63) || kind < 0 || kind > 3)
        return fail(SWE_EINVAL, "swe_halo_map: bad arguments");
    int depth = kind == 1 ? ch_depth(order) : 1;
    auto m = swe::build_halo_map(N, kind, depth);
    *count = (int)m.size();
    if (tp) {
        for (size_t k = 0; k < m.size(); k++) {
            tp[k] = m[k].tp; ti[k] = m[k].ti; tj[k] = m[k].tj; tc[k] = m[k].tc;
            sp[k] = m[k].sp; si[k] = m[k].si; sj[k] = m[k].sj; sc[k] = m[k].sc;
            sign[k] = m[k].sign;
        }
    }
    return SWE_OK;
}

}  // extern "C"

namespace {
template <int TY>
int check_tables_ty(int N)
{
    for (int q = 0; q < 6; q++)
        for (int sd = 0; sd < 4; sd++) {
            const swe::SideMatch m = swe::match_side(q, sd);
            const SideDesc k = side_of(q, sd);
            if (k.q != m.q || k.qs != m.q_side || k.rev != (m.reversed ? 1 : 0))
                return fail(SWE_EINVAL, "packed cube side table disagrees with the topology");
        }
    FusedTables ft;
    std::string err;
    if (build_fused_tables(N, TY, ft, err)) return fail(SWE_EINVAL, err);
    // every first-tile width a plan may request gives a tiling that fits,
    // and one more or less is refused
    for (int ax = 0; ax < 2; ax++) {
        const int tw = ax == 0 ? F_TX : TY, n = N + 1, k = (n - 17) / tw;
        if (n < 17 + tw || k < 1) continue;
        const int m = n - k * tw;
        for (int req = 7; req <= m - 8; req++) {
            FusedTables f2;
            const bool ok = build_fused_tables(N, TY, f2, err, ax == 0 ? req : 0, ax == 1 ? req : 0) == 0;
            const bool in_range = req >= 8 && req <= m - 9;
            if (!in_range && ok) return fail(SWE_EINVAL, "a first-tile width out of range was accepted");
            if (ok && (ax == 0 ? f2.fwx : f2.fwy) != req) return fail(SWE_EINVAL, "a requested first-tile width was not applied");
        }
    }
    for (int i = 0; i < ft.nseg; i++) {
        const int4 a = segment_of<TY>(N, i), b = ft.segs[i];
        if (a.x != b.x || a.y != b.y || a.z != b.z || a.w != b.w)
            return fail(SWE_EINVAL, "edge segment decode disagrees with the segment table");
    }
    // every segment's band fits the records kernels' frame
    for (int i = 0; i < ft.nseg; i++) {
        const int4 g = ft.segs[i];
        RTile<float> rt;
        rt.setup(N, g.x, g.y, g.z, g.w, nullptr);
        if (rt.frame_points() > RSmem<float>::CAP || rt.frame_points() > RTile<float>::CH * F_REC_THREADS)
            return fail(SWE_EINVAL, "an edge segment's band exceeds the records frame");
    }
    // the launch order is a bijection of the tiles, side tiles first
    const int nt = ft.ntiles, e = side_tiles(ft.ntx, ft.nty);
    std::vector<char> seen(nt, 0);
    for (int i = 0; i < nt; i++) {
        const int t = tile_at(i, ft.ntx, ft.nty);
        if (t < 0 || t >= nt || seen[t]) return fail(SWE_EINVAL, "tile order is not a bijection");
        seen[t] = 1;
        const int r = t / 6, ty = r / ft.ntx, tx = r - ty * ft.ntx;
        const bool side = tx == 0 || tx == ft.ntx - 1 || ty == 0 || ty == ft.nty - 1;
        if ((i < e) != side) return fail(SWE_EINVAL, "tile order does not put the side tiles first");
    }
    return SWE_OK;
}
}  // namespace

extern "C" {

int swe_check_tables(int N, int precision)
{
    if (N < 4 || N > 4096) return fail(SWE_EINVAL, "swe_check_tables: N out of range");
    if (precision == SWE_FP32) return check_tables_ty<16>(N);
    if (precision == SWE_FP64) return check_tables_ty<8>(N);
    return fail(SWE_EINVAL, "precision must be SWE_FP64 or SWE_FP32");
}

}  // extern "C"
#include "swe_ensemble.inc"
#include "swe_server.inc"
