// Layout and FP64 geometry kernels of libchswe (device and host code; one
// translation unit includes this inside its anonymous namespace).
#pragma once

constexpr int HALO = 4;
constexpr double PI = 3.14159265358979323846;

// ------------------------------------------------------------------------
// layout
// ------------------------------------------------------------------------
struct Dims {
    int N;     // cells per panel edge
    int P;     // row pitch (elements)
    long PS;   // panel stride (elements)
    long F;    // field stride = 6 * PS
};

inline Dims make_dims(int N)
{
    Dims d;
    d.N = N;
    d.P = ((N + 1 + 2 * HALO) + 15) / 16 * 16;
    d.PS = (long)d.P * (N + 1 + 2 * HALO);
    d.F = 6 * d.PS;
    return d;
}

__host__ __device__ inline long off(const Dims &d, int p, int i, int j)
{
    return p * d.PS + (long)(j - 1 + HALO) * d.P + (i - 1 + HALO);
}

// Device-side addressing of the kernels: 32-bit element offsets (a field
// holds fewer than 2^32 / 8 elements for N <= 4096) scaled to 32-bit byte
// offsets from a wave-uniform base pointer, so that a load is one
// base-plus-offset instruction and arrays indexed alike share the offset
// register (no 64-bit address arithmetic per access).  Row products use the
// 24-bit multiplier (indices and pitches stay below 2^24).
__device__ __forceinline__ unsigned poff(const Dims &d, int p, int i, int j)
{
    return (unsigned)p * (unsigned)d.PS + __umul24((unsigned)(j - 1 + HALO), (unsigned)d.P) + (unsigned)(i - 1 + HALO);
}
template <typename T>
__device__ __forceinline__ const T &at(const T *base, unsigned idx)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T &at(T *base, unsigned idx)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (unsigned)sizeof(T));
}

// flat interior index -> (p, i, j) for a mesh with nx x ny points per panel
__device__ inline void unflat(long n, int nx, int ny, int &p, int &i, int &j)
{
    long per = (long)nx * ny;
    p = (int)(n / per);
    long r = n - p * per;
    j = (int)(r / nx) + 1;
    i = (int)(r - (long)(j - 1) * nx) + 1;
}

// ------------------------------------------------------------------------
// geometry (FP64 on the device): metric/metric_2d_ecs_mod.f90 L338-411,
// mesh/mesh_factory_mod.f90 L36-37, coriolis_factory_mod.f90 L239-280,
// quadrature_factory_mod.f90 L89-188
// ------------------------------------------------------------------------
__constant__ int c_EX[6][3] = {{-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0}};
__constant__ int c_EY[6][3] = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
__constant__ int c_NN[6][3] = {{0, 0, 1}, {-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, -1, 0}};

struct GeomParams {
    double hx, radius, omega, axis[3];
    int coriolis_const, order;
};

__device__ inline double sbp_weight(int idx, int n, bool centre, int order)
{
    // D21_A_interfaces/centers, D42_A_interfaces/centers, D63_A_interfaces/
    // centers (sbp_operators_collection_mod.f90 L169-170, L196-200)
    const double i21[2] = {0.5, 1.0}, c21[2] = {1.0, 1.0};
    const double i42[5] = {7.0 / 18, 9.0 / 8, 1.0, 71.0 / 72, 1.0};
    const double c42[4] = {13.0 / 12, 7.0 / 8, 25.0 / 24, 1.0};
    const double i63[5] = {95.0 / 288, 317.0 / 240, 23.0 / 30, 793.0 / 720, 157.0 / 160};
    const double c63[6] = {325363.0 / 276480, 144001.0 / 276480, 43195.0 / 27648,
                           86857.0 / 138240,  312623.0 / 276480, 271229.0 / 276480};
    const double *A = order == 21 ? (centre ? c21 : i21) : order == 42 ? (centre ? c42 : i42) : (centre ? c63 : i63);
    int ne = order == 21 ? 2 : order == 42 ? (centre ? 4 : 5) : (centre ? 6 : 5);
    double w = 1.0;
    if (idx <= ne) w = A[idx - 1];
    if (n - idx + 1 <= ne) w = A[n - idx];
    return w;
}

// SBP quadrature weight J hx hy Ax(i) Ay(j) of point (i, j) on mesh m (0 xy,
// 1 u at x centres, 2 v at y centres), FP64 (quadrature_factory_mod.f90
// L89-188: create_C_sbp_quadrature)
// mesh m: 0 xy, 1 u (x centres), 2 v (y centres), 3 o (cell centres)
__device__ inline double quad_weight(int N, int order, double hx, int m, int i, int j)
{
    const int xt = m == 1 || m == 3, yt = m == 2 || m == 3;
    const int nx = xt ? N : N + 1, ny = yt ? N : N + 1;
    const double al = -0.25 * PI + (i - 1 + 0.5 * xt) * hx;
    const double be = -0.25 * PI + (j - 1 + 0.5 * yt) * hx;
    const double X = tan(al), Y = tan(be);
    const double s2 = 1.0 + X * X + Y * Y, s = sqrt(s2);
    const double J = (1.0 + X * X) * (1.0 + Y * Y) / (s2 * s);
    return J * hx * hx * sbp_weight(i, nx, xt, order) * sbp_weight(j, ny, yt, order);
}

// mesh m: 0 xy, 1 u (x centres), 2 v (y centres).  Writes FP64 J, Qi, f.
template <typename T>
__global__ void k_metric(Dims d, GeomParams g, int m, T *c0, T *c1, T *c2, double *w)
{
    const int N = d.N;
    const int xt = m == 1, yt = m == 2;
    const int nx = xt ? N : N + 1, ny = yt ? N : N + 1;
    long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= 6L * nx * ny) return;
    int p, i, j;
    unflat(n, nx, ny, p, i, j);
    double al = -0.25 * PI + (i - 1 + 0.5 * xt) * g.hx;
    double be = -0.25 * PI + (j - 1 + 0.5 * yt) * g.hx;
    double X = tan(al), Y = tan(be);
    double s2 = 1.0 + X * X + Y * Y, s = sqrt(s2);
    double J = (1.0 + X * X) * (1.0 + Y * Y) / (s2 * s);
    double q11 = s2 / (1.0 + X * X), q12 = s2 * X * Y / ((1.0 + X * X) * (1.0 + Y * Y)),
           q22 = s2 / (1.0 + Y * Y);
    long o = off(d, p, i, j);
    w[o] = quad_weight(N, g.order, g.hx, m, i, j);
    if (m == 0) {
        // r = ex X/s + ey Y/s - n 1/s (ecs_proto2realface with identity rotation)
        double r[3];
        for (int k = 0; k < 3; k++) r[k] = (c_EX[p][k] * X + c_EY[p][k] * Y - c_NN[p][k]) / s;
        double f;
        if (g.coriolis_const) {
            f = g.omega;
        } else {
            double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            f = 2.0 * g.omega * (g.axis[0] * r[0] + g.axis[1] * r[1] + g.axis[2] * r[2]) / rn;
        }
        c0[o] = T(J * q12);              // J Qi12 (co2contra cross term)
        c1[o] = T(f * J * J);            // f J^2 (Coriolis)
        c2[o] = T(1.0 / (J * g.radius * g.hx));  // 1/(hx J a) (div)
    } else {
        c0[o] = T(m == 1 ? q11 : q22);   // Qi11 at u, Qi22 at v
        c1[o] = T(1.0 / J);
        c2[o] = T(J);
    }
}

// covariant edge-sync weights (ecs_Ah_vec_sync_factory_mod.f90 L97-152):
// own across-edge covariant vector . neighbour's along-edge contravariant
// vector, the neighbour's prototype face folded onto this one.
__device__ inline void ecs_basis(double al, double be, double a1[3], double a2[3], double b1[3], double b2[3])
{
    double X = tan(al), Y = tan(be);
    double s2 = 1 + X * X + Y * Y, s3 = s2 * sqrt(s2);
    double cx = 1 + X * X, cy = 1 + Y * Y;
    a1[0] = cx * cy / s3; a1[1] = -X * Y * cx / s3; a1[2] = -X * cx / s3;
    a2[0] = -X * Y * cy / s3; a2[1] = cx * cy / s3; a2[2] = -Y * cy / s3;
    double q11 = s2 / cx, q12 = s2 * X * Y / (cx * cy), q22 = s2 / cy;
    for (int k = 0; k < 3; k++) {
        b1[k] = q11 * a1[k] + q12 * a2[k];
        b2[k] = q12 * a1[k] + q22 * a2[k];
    }
}

template <typename T>
__global__ void k_sync_weights(int N, double hx, T *qb, T *qt, T *ql, T *qr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > N + 1) return;
    double c = -0.25 * PI + (i - 1) * hx, e = 0.25 * PI;
    double a1[3], a2[3], b1[3], b2[3], A1[3], A2[3], B1[3], B2[3];
    ecs_basis(c, -e, a1, a2, b1, b2);
    ecs_basis(c, e, A1, A2, B1, B2);
    qb[i - 1] = T(a2[0] * B1[0] - a2[1] * B1[2] + a2[2] * B1[1]);
    ecs_basis(c, e, a1, a2, b1, b2);
    ecs_basis(c, -e, A1, A2, B1, B2);
    qt[i - 1] = T(a2[0] * B1[0] + a2[1] * B1[2] - a2[2] * B1[1]);
    ecs_basis(-e, c, a1, a2, b1, b2);
    ecs_basis(e, c, A1, A2, B1, B2);
    ql[i - 1] = T(-a1[0] * B2[2] + a1[1] * B2[1] + a1[2] * B2[0]);
    ecs_basis(e, c, a1, a2, b1, b2);
    ecs_basis(-e, c, A1, A2, B1, B2);
    qr[i - 1] = T(a1[0] * B2[2] + a1[1] * B2[1] - a1[2] * B2[0]);
}

// panel-independent coefficient tables for the fused path (FP64 arithmetic,
// stored in T) on the (N+1) x (N+1) index grid, plus f J^2 per panel
template <typename T>
__global__ void k_coef(int N, GeomParams g, T *JQ12, T *iJa, T *Q11, T *iJu, T *Ju, T *Q22, T *iJv, T *Jv, T *fJ2)
{
    const int n = N + 1;
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= 6L * n * n) return;
    int p = (int)(k / ((long)n * n));
    int r = (int)(k - (long)p * n * n);
    int j = r / n + 1, i = r % n + 1;
    auto metric = [&](double xs, double ys, double &J, double &q11, double &q12, double &q22, double &X,
                      double &Y, double &s) {
        double al = -0.25 * PI + (i - 1 + xs) * g.hx, be = -0.25 * PI + (j - 1 + ys) * g.hx;
        X = tan(al);
        Y = tan(be);
        double s2 = 1.0 + X * X + Y * Y;
        s = sqrt(s2);
        J = (1.0 + X * X) * (1.0 + Y * Y) / (s2 * s);
        q11 = s2 / (1.0 + X * X);
        q12 = s2 * X * Y / ((1.0 + X * X) * (1.0 + Y * Y));
        q22 = s2 / (1.0 + Y * Y);
    };
    double J, q11, q12, q22, X, Y, s;
    metric(0.0, 0.0, J, q11, q12, q22, X, Y, s);
    double f;
    if (g.coriolis_const) {
        f = g.omega;
    } else {
        double rr[3];
        for (int m = 0; m < 3; m++) rr[m] = (c_EX[p][m] * X + c_EY[p][m] * Y - c_NN[p][m]) / s;
        double rn = sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
        f = 2.0 * g.omega * (g.axis[0] * rr[0] + g.axis[1] * rr[1] + g.axis[2] * rr[2]) / rn;
    }
    fJ2[k] = T(f * J * J);
    if (p != 0) return;
    JQ12[r] = T(J * q12);
    iJa[r] = T(1.0 / (J * g.radius * g.hx));
    if (i <= N) {
        metric(0.5, 0.0, J, q11, q12, q22, X, Y, s);
        Q11[r] = T(q11);
        iJu[r] = T(1.0 / J);
        Ju[r] = T(J);
    }
    if (j <= N) {
        metric(0.0, 0.5, J, q11, q12, q22, X, Y, s);
        Q22[r] = T(q22);
        iJv[r] = T(1.0 / J);
        Jv[r] = T(J);
    }
}

