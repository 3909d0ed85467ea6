/*
 * CPU oracle for the linear SWE RK4 step (Ch staggering, SBP 21 / 42).
 * TEST INFRASTRUCTURE ONLY -- see swe_oracle.h for scope and citations.
 *
 * The restatement keeps the reference's structure on purpose: halo-padded
 * per-panel arrays, an explicit halo fill before every operator that reads
 * across a panel edge, then 1-D SBP sweeps along rows and columns.  It is
 * deliberately simple and scalar; speed is not its job.
 */
#include "swe_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define HALO 4 /* widest read across an edge: the sbp63 c2i SAT, depth 4 */

/* Row loops are split over OpenMP threads when built with -fopenmp (used for
 * the timed CPU baseline).  Every parallel loop writes disjoint points and
 * holds no reduction, so results do not depend on the thread count. */
#ifdef _OPENMP
#include <omp.h>
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#define OMP_FOR_COLLAPSE2 _Pragma("omp parallel for collapse(2) schedule(static)")
int swo_num_threads(void) { return omp_get_max_threads(); }
void swo_set_num_threads(int n) { omp_set_num_threads(n); }
#else
#define OMP_FOR
#define OMP_FOR_COLLAPSE2
int swo_num_threads(void) { return 1; }
void swo_set_num_threads(int n) { (void)n; }
#endif

/* Planted mistakes for tests that prove the checks have power (0 = none):
 * 1 flip the sign of rotated Ch-vector halo entries, 2 drop the metric term
 * of the covariant edge sync, 3 read the depth-2 xy halo instead of the
 * duplicate in the gradient SAT, 4 perturb one 4th-order c2i closure
 * coefficient, 5 skip the div projection. */
static int g_mutation = 0;
void swo_set_mutation(int m) { g_mutation = m; }
static int g_no_sat = 0; /* set only inside swo_line_op */

/* ------------------------------------------------------------------ */
/* cube topology: topology/cubed_sphere_topology_mod.f90 L29-55.       */
/* Panel frames on the unit cube: r = corner of point (1,1), ex / ey = */
/* directions of increasing i / j, n = inward normal.                   */
/* ------------------------------------------------------------------ */
static const int T_EX[6][3] = {{-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0}};
static const int T_EY[6][3] = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
static const int T_N[6][3] = {{0, 0, 1}, {-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
static const int T_R[6][3] = {{1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 1, 1}};

static int idot(const int *a, const int *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ------------------------------------------------------------------ */
/* halo-padded field: index (p, i, j), i in [1-HALO, nx+HALO]           */
/* ------------------------------------------------------------------ */
typedef struct {
    double *d;
    int nx, ny, sx, sy;
} fld;

static void fld_init(fld *f, int nx, int ny)
{
    f->nx = nx;
    f->ny = ny;
    f->sx = nx + 2 * HALO;
    f->sy = ny + 2 * HALO;
    f->d = (double *)calloc((size_t)6 * f->sx * f->sy, sizeof(double));
}
static void fld_free(fld *f) { free(f->d); }
static inline double *AT(const fld *f, int p, int i, int j)
{
    return f->d + (size_t)p * f->sx * f->sy + (size_t)(j - 1 + HALO) * f->sx + (i - 1 + HALO);
}
static void fld_load(fld *f, const double *src)
{
    OMP_FOR_COLLAPSE2
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= f->ny; j++)
            for (int i = 1; i <= f->nx; i++)
                *AT(f, p, i, j) = src[((size_t)p * f->ny + (j - 1)) * f->nx + (i - 1)];
}
static void fld_store(const fld *f, double *dst)
{
    OMP_FOR_COLLAPSE2
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= f->ny; j++)
            for (int i = 1; i <= f->nx; i++)
                dst[((size_t)p * f->ny + (j - 1)) * f->nx + (i - 1)] = *AT(f, p, i, j);
}

/* ------------------------------------------------------------------ */
/* halo maps                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    int tp, ti, tj, tc; /* target panel, index, component (0 u-type, 1 v-type) */
    int sp, si, sj, sc; /* source */
    int sign;
} hentry;

typedef struct {
    hentry *e;
    int n;
} hmap;

/* axis types: 0 = vertex (N+1 points), 1 = centre (N points) */
typedef struct {
    int xt, yt;
} mtype;
static const mtype MT_XY = {0, 0}, MT_U = {1, 0}, MT_V = {0, 1};

static int npts_axis(int N, int t) { return t == 0 ? N + 1 : N; }

/*
 * Unfold a target halo point across one panel edge (the index rule of
 * topology/cubed_sphere_topology_mod.f90 transform_index L80-148 and the
 * component rule of parallel/buffer_mod.f90 pack_to_buf_vec L171-277):
 * the neighbour face is folded flat onto the target's plane; halo depth d
 * counts points of the target's axis type into the neighbour from the
 * shared edge, so for vertex-type axes d = 1 is the neighbour's duplicate
 * of the edge point.  Components follow the folding isometry:
 * target_comp = M * source_comp with M_ab = (unfolded target axis a) .
 * (source axis b).
 * side: 0 left (i<1), 1 right (i>nx), 2 bottom (j<1), 3 top (j>ny).
 * t = 1-based along-edge index in the target's indexing.
 */
static void unfold(int N, int p, int side, int d, int t, mtype mt, int *q_out,
                   int *si, int *sj, int M[2][2], mtype *smt)
{
    const int L = 2 * N; /* cube edge in half-steps */
    int X[3], ex_u[3], ey_u[3];
    int tang_type = (side < 2) ? mt.yt : mt.xt;
    int norm_type = (side < 2) ? mt.xt : mt.yt;
    int s = tang_type == 0 ? 2 * (t - 1) : 2 * t - 1;    /* along-edge position */
    int dist = norm_type == 0 ? 2 * (d - 1) : 2 * d - 1; /* depth into neighbour */
    int nq[3];
    for (int k = 0; k < 3; k++) {
        int R = L * T_R[p][k];
        switch (side) {
        case 0:
            X[k] = R + s * T_EY[p][k] + dist * T_N[p][k];
            ex_u[k] = -T_N[p][k];
            ey_u[k] = T_EY[p][k];
            nq[k] = T_EX[p][k];
            break;
        case 1:
            X[k] = R + L * T_EX[p][k] + s * T_EY[p][k] + dist * T_N[p][k];
            ex_u[k] = T_N[p][k];
            ey_u[k] = T_EY[p][k];
            nq[k] = -T_EX[p][k];
            break;
        case 2:
            X[k] = R + s * T_EX[p][k] + dist * T_N[p][k];
            ex_u[k] = T_EX[p][k];
            ey_u[k] = -T_N[p][k];
            nq[k] = T_EY[p][k];
            break;
        default:
            X[k] = R + L * T_EY[p][k] + s * T_EX[p][k] + dist * T_N[p][k];
            ex_u[k] = T_EX[p][k];
            ey_u[k] = T_N[p][k];
            nq[k] = -T_EY[p][k];
            break;
        }
    }
    int q = -1;
    for (int k = 0; k < 6; k++)
        if (T_N[k][0] == nq[0] && T_N[k][1] == nq[1] && T_N[k][2] == nq[2]) q = k;
    int rel[3];
    for (int k = 0; k < 3; k++) rel[k] = X[k] - L * T_R[q][k];
    int xq = idot(rel, T_EX[q]), yq = idot(rel, T_EY[q]);
    M[0][0] = idot(ex_u, T_EX[q]);
    M[0][1] = idot(ex_u, T_EY[q]);
    M[1][0] = idot(ey_u, T_EX[q]);
    M[1][1] = idot(ey_u, T_EY[q]);
    if (M[0][0] != 0) {
        *smt = mt;
    } else {
        smt->xt = mt.yt;
        smt->yt = mt.xt;
    }
    *si = smt->xt == 0 ? xq / 2 + 1 : (xq + 1) / 2;
    *sj = smt->yt == 0 ? yq / 2 + 1 : (yq + 1) / 2;
    *q_out = q;
}

static void hmap_push(hmap *m, hentry e)
{
    m->e = (hentry *)realloc(m->e, sizeof(hentry) * (m->n + 1));
    m->e[m->n++] = e;
}

/* xy points, sides in {0..3}, depth 1..depth, both components recorded */
static void build_xy_map(hmap *m, int N, int depth, int vector)
{
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++)
            for (int d = 1; d <= depth; d++)
                for (int t = 1; t <= N + 1; t++) {
                    int q, si, sj, M[2][2];
                    mtype smt;
                    unfold(N, p, side, d, t, MT_XY, &q, &si, &sj, M, &smt);
                    int ti, tj;
                    if (side == 0) { ti = 1 - d; tj = t; }
                    else if (side == 1) { ti = N + 1 + d; tj = t; }
                    else if (side == 2) { ti = t; tj = 1 - d; }
                    else { ti = t; tj = N + 1 + d; }
                    if (!vector) {
                        hentry e = {p, ti, tj, 0, q, si, sj, 0, 1};
                        hmap_push(m, e);
                    } else {
                        for (int c = 0; c < 2; c++) {
                            int sc = M[c][0] != 0 ? 0 : 1;
                            hentry e = {p, ti, tj, c, q, si, sj, sc, M[c][sc]};
                            hmap_push(m, e);
                        }
                    }
                }
}

/* Ch vector: u (mesh y) across left/right, v (mesh x) across bottom/top;
 * parallel/exchange_factory_mod.f90 L369-672 with halo width hw+1 along the
 * vertex-type axis, restricted to the rows the div SAT reads. */
static void build_ch_map(hmap *m, int N, int depth)
{
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++) {
            mtype mt = side < 2 ? MT_U : MT_V;
            int c = side < 2 ? 0 : 1;
            for (int d = 1; d <= depth; d++)
                for (int t = 1; t <= N + 1; t++) {
                    int q, si, sj, M[2][2];
                    mtype smt;
                    unfold(N, p, side, d, t, mt, &q, &si, &sj, M, &smt);
                    int ti, tj;
                    if (side == 0) { ti = 1 - d; tj = t; }
                    else if (side == 1) { ti = N + d; tj = t; }
                    else if (side == 2) { ti = t; tj = 1 - d; }
                    else { ti = t; tj = N + d; }
                    int sc = M[c][0] != 0 ? 0 : 1;
                    int sg = M[c][sc];
                    if (g_mutation == 1 && M[0][0] == 0) sg = -sg;
                    hentry e = {p, ti, tj, c, q, si, sj, sc, sg};
                    hmap_push(m, e);
                }
        }
}

/* tangential covariant components at the duplicated edge points: u (mesh y)
 * across bottom/top, v (mesh x) across left/right, depth 1 (the neighbour's
 * duplicate of the edge point) -- what the i2c SAT of curl_ch_sbp_sat reads
 * from its Ch vector exchange (curl_sbp_sat_mod.f90 L26-44,
 * sbp_diff_i2c_42_mod.f90 add_SAT_correction L162-274: f(0), f(n+2)). */
static void build_tan_map(hmap *m, int N)
{
    for (int p = 0; p < 6; p++)
        for (int side = 0; side < 4; side++) {
            mtype mt = side < 2 ? MT_V : MT_U;
            int c = side < 2 ? 1 : 0;
            for (int t = 1; t <= N; t++) {
                int q, si, sj, M[2][2];
                mtype smt;
                unfold(N, p, side, 1, t, mt, &q, &si, &sj, M, &smt);
                int ti, tj;
                if (side == 0) { ti = 0; tj = t; }
                else if (side == 1) { ti = N + 2; tj = t; }
                else if (side == 2) { ti = t; tj = 0; }
                else { ti = t; tj = N + 2; }
                int sc = M[c][0] != 0 ? 0 : 1;
                hentry e = {p, ti, tj, c, q, si, sj, sc, M[c][sc]};
                hmap_push(m, e);
            }
        }
}

static void apply_map(const hmap *m, fld *a, fld *b)
{
    fld *F[2] = {a, b};
    /* copy sources first: targets are halo points, sources interior */
    OMP_FOR
    for (int k = 0; k < m->n; k++) {
        const hentry *e = &m->e[k];
        *AT(F[e->tc], e->tp, e->ti, e->tj) = e->sign * *AT(F[e->sc], e->sp, e->si, e->sj);
    }
}

/* ------------------------------------------------------------------ */
/* 1-D SBP operators on strided lines.  f points at index 0.           */
/* ------------------------------------------------------------------ */
#define L1(f, st, k) ((f)[(ptrdiff_t)(k) * (st)])

/* free parameters of the 4th-order staggered interpolation,
 * sbp_interp_i2c_42_mod.f90 L18-19 */
static const double C13 = 0.1264926207034815, C14 = -0.029892648949074313;

/* sbp63 (sbp_diff_i2c_63_mod.f90 L11-75, sbp_diff_c2i_63_mod.f90 L13-60,
 * sbp_interp_i2c_63_mod.f90 L14-90, sbp_interp_c2i_63_mod.f90 L11-62): the
 * closure rows evaluated to doubles (row r reads points 1..9 of its end;
 * the right end mirrors them), the interior stencils, the centre and
 * interface norms' boundary weights and the SAT projection. */
static const double D63_I2C[6][9] = {{-0.9795214020218487, 0.9597522747540616, -0.0021284121310923276, 0.04308560808739484, -0.02118806868851538, 0.0, 0.0, 0.0, 0.0},
 {0.10094185622089052, -1.3621007582168956, 1.480651137325344, -0.2787674248835621, 0.059275189554223895, 0.0, 0.0, 0.0, 0.0},
 {0.1050522197237869, -0.3755418649660661, -0.5066880707069372, 0.7227932046793403, 0.05138416400746168, 0.0030003472624146313, 0.0, 0.0, 0.0},
 {-0.010281664659776757, 0.009016888173475901, 0.1158762991569453, -1.3063741185003286, 1.2879206487212211, -0.10361859147794651, 0.007460538586412149, 0.0, 0.0},
 {-0.08522272532774755, 0.2300959768599209, -0.06748505183821596, -0.28075281972041166, -0.77959568679678, 1.0363920760788554, -0.05757733755993641, 0.0041455683043154215, 0.0},
 {0.03958146265158008, -0.10105244572213268, 0.0269485128845091, 0.10551623486389128, -0.009407429883264403, -1.1945625283432082, 1.1945625283432082, -0.066364584907956, 0.004778250113372832}};
static const double D63_C2I[5][9] = {{-3.137061163091724, 6.472195962306311, -4.476505551641335, 0.9669525120033824, 0.2921335971506185, -0.11771535672725233, 0.0, 0.0, 0.0},
 {-0.8550973738466245, 0.5371097071174838, 0.4442015766629761, -0.00428923422758698, -0.1969782207705677, 0.07505354506431924, 0.0, 0.0, 0.0},
 {0.003267033496606009, -1.0058841165930086, 1.032532798072641, -0.09496402962593153, 0.0995309639229438, -0.03448264927325102, 0.0, 0.0, 0.0},
 {-0.0460358301286618, 0.1318266207921457, -1.0252815152152988, 0.7452431221796392, 0.2882309687613436, -0.093983366389168, 0.0, 0.0, 0.0},
 {0.025410671711714995, -0.031462633326690415, -0.08181244708002722, -0.8246706459806198, 0.8983528780132024, 0.009405106598725824, 0.004777070063694267, 0.0, 0.0}};
static const double I63_I2C[6][9] = {{0.4559762640735722, 0.5156609261411327, 0.09215963713516839, -0.055207108988023945, -0.008589718361849435, 0.0, 0.0, 0.0, 0.0},
 {-0.2329150785274282, 1.0886403314451474, 0.006569476829126231, 0.1526003661160169, -0.014895095862863417, 0.0, 0.0, 0.0, 0.0},
 {-0.13300622901536394, 0.24335188147388265, 0.43548086151449933, 0.4385084167668463, 0.008164201104099517, 0.007500868156036579, 0.0, 0.0, 0.0},
 {0.10717671074312715, -0.3332211159670528, 0.3310769312612241, 0.29067744812134255, 0.7410665665922483, -0.15542788721691977, 0.018651346466030373, 0.0, 0.0},
 {0.09305432569823871, -0.07099703081266313, -0.291616405335888, 0.3005623879587947, 0.5268027700312061, 0.5181960380394277, -0.08636600633990461, 0.010363920760788554, 0.0},
 {-0.05115521553501697, 0.05753938634775091, 0.09801308881771609, -0.1230378129758785, -0.08832072291927771, 0.5972812641716041, 0.5972812641716041, -0.09954687736193402, 0.011945625283432081}};
static const double I63_C2I[5][9] = {{1.6267303202606325, -0.3677632041998704, -0.6299565857805535, 0.2041457799345569, 0.3189794129688649, -0.15213572318363067, 0.0, 0.0, 0.0},
 {0.4594313713417274, 0.42927755971902565, 0.287843512318841, -0.15850960868795075, -0.060778415165360984, 0.04273558047371772, 0.0, 0.0, 0.0},
 {0.14146161692901638, 0.004462990795171873, 0.8874262064612957, 0.2713272665549153, -0.43009320031948844, 0.12541511957908888, 0.0, 0.0, 0.0},
 {-0.058987332524401276, 0.07216334765484655, 0.6220237974938234, 0.16582184683346107, 0.3085681858542267, -0.10958984531195667, 0.0, 0.0, 0.0},
 {-0.0103015766372022, -0.007906156741522508, 0.012998815562764554, 0.47451358497362967, 0.607051568675785, -0.0882989109926898, 0.01194267515923567, 0.0, 0.0}};
static const double H63_C[6] = {325363.0 / 276480, 144001.0 / 276480, 43195.0 / 27648,
                                86857.0 / 138240, 312623.0 / 276480, 271229.0 / 276480};
static const double H63_I[5] = {95.0 / 288, 317.0 / 240, 23.0 / 30, 793.0 / 720, 157.0 / 160};
static const double R63[4] = {35.0 / 16, -35.0 / 16, 21.0 / 16, -5.0 / 16};
static const double QD63[3] = {75.0 / 64, -25.0 / 384, 3.0 / 640};
static const double QI63[3] = {75.0 / 128, -25.0 / 256, 3.0 / 256};

/* interfaces -> centres derivative, n centres out, reads f[0..n+2]
 * (f[0], f[n+2] = neighbour's duplicates of the edge points)
 * sbp_diff_i2c_21_mod.f90 L24-144, sbp_diff_i2c_42_mod.f90 L29-274 */
static void d_i2c(int order, const double *f, ptrdiff_t fs, int n, double *o, ptrdiff_t os)
{
    if (order == 63) {
        for (int r = 0; r < 6; r++) {
            double sl = 0, sr = 0;
            for (int k = 0; k < 9; k++) {
                sl += D63_I2C[r][k] * L1(f, fs, 1 + k);
                sr += D63_I2C[r][k] * L1(f, fs, n + 1 - k);
            }
            L1(o, os, 1 + r) = sl;
            L1(o, os, n - r) = -sr;
        }
        for (int k = 7; k <= n - 6; k++)
            L1(o, os, k) = QD63[0] * (L1(f, fs, k + 1) - L1(f, fs, k)) + QD63[1] * (L1(f, fs, k + 2) - L1(f, fs, k - 1)) +
                           QD63[2] * (L1(f, fs, k + 3) - L1(f, fs, k - 2));
        if (g_no_sat) return;
        const double jl = L1(f, fs, 1) - L1(f, fs, 0), jr = L1(f, fs, n + 2) - L1(f, fs, n + 1);
        for (int r = 0; r < 4; r++) {
            L1(o, os, 1 + r) += 0.5 * R63[r] / H63_C[r] * jl;
            L1(o, os, n - r) += 0.5 * R63[r] / H63_C[r] * jr;
        }
        return;
    }
    if (order == 21) {
        for (int k = 1; k <= n; k++) L1(o, os, k) = L1(f, fs, k + 1) - L1(f, fs, k);
        if (g_no_sat) return;
        double jl = L1(f, fs, 1) - L1(f, fs, 0);
        double jr = L1(f, fs, n + 2) - L1(f, fs, n + 1);
        /* 0.5*R/h with R = (3/2,-1/2), h = (1,1) */
        L1(o, os, 1) += 0.75 * jl;
        L1(o, os, 2) += -0.25 * jl;
        L1(o, os, n) += 0.75 * jr;
        L1(o, os, n - 1) += -0.25 * jr;
        return;
    }
    static const double Q[3][5] = {{-79.0 / 78, 27.0 / 26, -1.0 / 26, 1.0 / 78, 0},
                                   {2.0 / 21, -9.0 / 7, 9.0 / 7, -2.0 / 21, 0},
                                   {1.0 / 75, 0, -27.0 / 25, 83.0 / 75, -1.0 / 25}};
    for (int r = 0; r < 3; r++) {
        double sl = 0, sr = 0;
        for (int k = 0; k < 5; k++) {
            sl += Q[r][k] * L1(f, fs, 1 + k);
            sr += Q[r][k] * L1(f, fs, n + 1 - k);
        }
        L1(o, os, 1 + r) = sl;
        L1(o, os, n - r) = -sr;
    }
    for (int k = 4; k <= n - 3; k++)
        L1(o, os, k) = (27.0 * (L1(f, fs, k + 1) - L1(f, fs, k)) - (L1(f, fs, k + 2) - L1(f, fs, k - 1))) / 24.0;
    if (g_no_sat) return;
    static const double H[3] = {13.0 / 12, 7.0 / 8, 25.0 / 24};
    static const double R[3] = {15.0 / 8, -5.0 / 4, 3.0 / 8};
    double jl = L1(f, fs, 1) - L1(f, fs, 0);
    double jr = L1(f, fs, n + 2) - L1(f, fs, n + 1);
    for (int r = 0; r < 3; r++) {
        L1(o, os, 1 + r) += 0.5 * R[r] / H[r] * jl;
        L1(o, os, n - r) += 0.5 * R[r] / H[r] * jr;
    }
}

/* centres -> interfaces derivative, n centres in, n+1 interfaces out,
 * SAT reads f[1-depth..0] and f[n+1..n+depth] (neighbour's normal flux)
 * sbp_diff_c2i_21_mod.f90 L24-146, sbp_diff_c2i_42_mod.f90 L32-251 */
static void d_c2i(int order, const double *f, ptrdiff_t fs, int n, double *o, ptrdiff_t os)
{
    int nx = n + 1;
    if (order == 63) {
        for (int r = 0; r < 5; r++) {
            double sl = 0, sr = 0;
            for (int k = 0; k < 9; k++) {
                sl += D63_C2I[r][k] * L1(f, fs, 1 + k);
                sr += D63_C2I[r][k] * L1(f, fs, nx - 1 - k);
            }
            L1(o, os, 1 + r) = sl;
            L1(o, os, nx - r) = -sr;
        }
        for (int k = 6; k <= nx - 5; k++)
            L1(o, os, k) = QD63[0] * (L1(f, fs, k) - L1(f, fs, k - 1)) + QD63[1] * (L1(f, fs, k + 1) - L1(f, fs, k - 2)) +
                           QD63[2] * (L1(f, fs, k + 2) - L1(f, fs, k - 3));
        if (g_no_sat) return;
        const double c = 0.5 / H63_I[0];
        double sl = 0, sr = 0;
        for (int r = 0; r < 4; r++) {
            sl += R63[r] * (L1(f, fs, 1 + r) - L1(f, fs, -r));
            sr += R63[r] * (L1(f, fs, nx - 1 - r) - L1(f, fs, nx + r));
        }
        L1(o, os, 1) += c * sl;
        L1(o, os, nx) -= c * sr;
        return;
    }
    if (order == 21) {
        L1(o, os, 1) = L1(f, fs, 2) - L1(f, fs, 1);
        for (int k = 2; k <= nx - 1; k++) L1(o, os, k) = L1(f, fs, k) - L1(f, fs, k - 1);
        L1(o, os, nx) = L1(f, fs, nx - 1) - L1(f, fs, nx - 2);
        if (g_no_sat) return;
        /* 0.5/h11 with h11 = 1/2; R = (3/2, -1/2) */
        L1(o, os, 1) += 1.0 * (1.5 * (L1(f, fs, 1) - L1(f, fs, 0)) - 0.5 * (L1(f, fs, 2) - L1(f, fs, -1)));
        L1(o, os, nx) -= 1.0 * (1.5 * (L1(f, fs, nx - 1) - L1(f, fs, nx)) -
                                0.5 * (L1(f, fs, nx - 2) - L1(f, fs, nx + 1)));
        return;
    }
    double Q[4][5] = {{-2.0, 3.0, -1.0, 0, 0},
                      {-1.0, 1.0, 0, 0, 0},
                      {1.0 / 24, -9.0 / 8, 9.0 / 8, -1.0 / 24, 0},
                      {-1.0 / 71, 6.0 / 71, -83.0 / 71, 81.0 / 71, -3.0 / 71}};
    if (g_mutation == 4) Q[3][2] = -82.0 / 71;
    for (int r = 0; r < 4; r++) {
        double sl = 0, sr = 0;
        for (int k = 0; k < 5; k++) {
            sl += Q[r][k] * L1(f, fs, 1 + k);
            sr += Q[r][k] * L1(f, fs, nx - 1 - k);
        }
        L1(o, os, 1 + r) = sl;
        L1(o, os, nx - r) = -sr;
    }
    for (int k = 5; k <= nx - 4; k++)
        L1(o, os, k) = (27.0 * (L1(f, fs, k) - L1(f, fs, k - 1)) - (L1(f, fs, k + 1) - L1(f, fs, k - 2))) / 24.0;
    if (g_no_sat) return;
    static const double R[3] = {15.0 / 8, -5.0 / 4, 3.0 / 8};
    const double c = 0.5 / (7.0 / 18);
    double sl = 0, sr = 0;
    for (int r = 0; r < 3; r++) {
        sl += R[r] * (L1(f, fs, 1 + r) - L1(f, fs, -r));
        sr += R[r] * (L1(f, fs, nx - 1 - r) - L1(f, fs, nx + r));
    }
    L1(o, os, 1) += c * sl;
    L1(o, os, nx) -= c * sr;
}

/* interfaces -> centres interpolation, n centres out, reads f[1..n+1]
 * sbp_interp_i2c_21_mod.f90 L19-57, sbp_interp_i2c_42_mod.f90 L18-198 */
static void i_i2c(int order, const double *f, ptrdiff_t fs, int n, double *o, ptrdiff_t os)
{
    if (order == 63) {
        for (int r = 0; r < 6; r++) {
            double sl = 0, sr = 0;
            for (int k = 0; k < 9; k++) {
                sl += I63_I2C[r][k] * L1(f, fs, 1 + k);
                sr += I63_I2C[r][k] * L1(f, fs, n + 1 - k);
            }
            L1(o, os, 1 + r) = sl;
            L1(o, os, n - r) = sr;
        }
        for (int k = 7; k <= n - 6; k++)
            L1(o, os, k) = QI63[0] * (L1(f, fs, k + 1) + L1(f, fs, k)) + QI63[1] * (L1(f, fs, k + 2) + L1(f, fs, k - 1)) +
                           QI63[2] * (L1(f, fs, k + 3) + L1(f, fs, k - 2));
        return;
    }
    if (order == 21) {
        for (int k = 1; k <= n; k++) L1(o, os, k) = 0.5 * (L1(f, fs, k + 1) + L1(f, fs, k));
        return;
    }
    const double c13 = C13, c14 = C14;
    const double Q[3][5] = {
        {0.5 + c13 + 2 * c14, 0.5 - 2 * c13 - 3 * c14, c13, c14, 0},
        {-8.0 / 63 - 52.0 / 21 * c13 - 104.0 / 21 * c14, 29.0 / 42 + 104.0 / 21 * c13 + 52.0 / 7 * c14,
         -52.0 / 21 * c13 + 0.5, -52.0 / 21 * c14 - 4.0 / 63, 0},
        {26.0 / 25 * c13 + 52.0 / 25 * c14 - 1.0 / 25, -1.0 / 50 - 52.0 / 25 * c13 - 78.0 / 25 * c14,
         3.0 / 5 + 26.0 / 25 * c13, 13.0 / 25 + 26.0 / 25 * c14, -3.0 / 50}};
    for (int r = 0; r < 3; r++) {
        double sl = 0, sr = 0;
        for (int k = 0; k < 5; k++) {
            sl += Q[r][k] * L1(f, fs, 1 + k);
            sr += Q[r][k] * L1(f, fs, n + 1 - k);
        }
        L1(o, os, 1 + r) = sl;
        L1(o, os, n - r) = sr;
    }
    for (int k = 4; k <= n - 3; k++)
        L1(o, os, k) = (9.0 * (L1(f, fs, k + 1) + L1(f, fs, k)) - (L1(f, fs, k + 2) + L1(f, fs, k - 1))) / 16.0;
}

/* centres -> interfaces interpolation, n centres in, n+1 out
 * sbp_interp_c2i_21_mod.f90 L19-87, sbp_interp_c2i_42_mod.f90 L13-218 */
static void i_c2i(int order, const double *f, ptrdiff_t fs, int n, double *o, ptrdiff_t os)
{
    int nx = n + 1;
    if (order == 63) {
        for (int r = 0; r < 5; r++) {
            double sl = 0, sr = 0;
            for (int k = 0; k < 9; k++) {
                sl += I63_C2I[r][k] * L1(f, fs, 1 + k);
                sr += I63_C2I[r][k] * L1(f, fs, nx - 1 - k);
            }
            L1(o, os, 1 + r) = sl;
            L1(o, os, nx - r) = sr;
        }
        for (int k = 6; k <= nx - 5; k++)
            L1(o, os, k) = QI63[0] * (L1(f, fs, k) + L1(f, fs, k - 1)) + QI63[1] * (L1(f, fs, k + 1) + L1(f, fs, k - 2)) +
                           QI63[2] * (L1(f, fs, k + 2) + L1(f, fs, k - 3));
        return;
    }
    if (order == 21) {
        L1(o, os, 1) = L1(f, fs, 1);
        for (int k = 2; k <= nx - 1; k++) L1(o, os, k) = 0.5 * (L1(f, fs, k) + L1(f, fs, k - 1));
        L1(o, os, nx) = L1(f, fs, nx - 1);
        return;
    }
    const double c13 = C13, c14 = C14;
    const double Q[4][5] = {
        {39.0 / 28 + 39.0 / 14 * c13 + 39.0 / 7 * c14, -2.0 / 7 - 39.0 / 7 * c13 - 78.0 / 7 * c14,
         39.0 / 14 * c13 + 39.0 / 7 * c14 - 3.0 / 28, 0, 0},
        {13.0 / 27 - 52.0 / 27 * c13 - 26.0 / 9 * c14, 29.0 / 54 + 104.0 / 27 * c13 + 52.0 / 9 * c14,
         -1.0 / 54 - 52.0 / 27 * c13 - 26.0 / 9 * c14, 0, 0},
        {13.0 / 12 * c13, -13.0 / 6 * c13 + 7.0 / 16, 5.0 / 8 + 13.0 / 12 * c13, -1.0 / 16, 0},
        {78.0 / 71 * c14, -156.0 / 71 * c14 - 4.0 / 71, 39.0 / 71 + 78.0 / 71 * c14, 81.0 / 142, -9.0 / 142}};
    for (int r = 0; r < 4; r++) {
        double sl = 0, sr = 0;
        for (int k = 0; k < 5; k++) {
            sl += Q[r][k] * L1(f, fs, 1 + k);
            sr += Q[r][k] * L1(f, fs, nx - 1 - k);
        }
        L1(o, os, 1 + r) = sl;
        L1(o, os, nx - r) = sr;
    }
    for (int k = 5; k <= nx - 4; k++)
        L1(o, os, k) = (9.0 * (L1(f, fs, k) + L1(f, fs, k - 1)) - (L1(f, fs, k + 1) + L1(f, fs, k - 2))) / 16.0;
}

/* ------------------------------------------------------------------ */
/* geometry: metric/metric_2d_ecs_mod.f90 set_curvilinear_mesh L338-411 */
/* ------------------------------------------------------------------ */
typedef struct {
    double *J, *Q11, *Q12, *Q22, *rx, *ry, *rz, *w; /* packed interior */
} geom;

struct swo {
    int N, order;
    double a, g, H0, omega, axis[3];
    int cor_const;
    double hx;
    mtype mt[3];
    geom G[3];
    double *fJ2;
    double *qb, *qt, *ql, *qr;
    hmap m_xy, m_xyv, m_ch, m_tan;
    fld h, u, v, gx, gy, ut, vt, w1, w2, up, vp, Ju, Jv, Dx, Dy, div;
    /* RK4 registers */
    double *k[4][3], *y[3];
};

static const double PI = 3.14159265358979323846;

int swo_npts(const swo *s, int mesh)
{
    const mtype t = s->mt[mesh];
    return 6 * npts_axis(s->N, t.xt) * npts_axis(s->N, t.yt);
}

static void proto2real(int p, const double v[3], double out[3])
{
    for (int k = 0; k < 3; k++) out[k] = T_EX[p][k] * v[0] + T_EY[p][k] * v[1] - T_N[p][k] * v[2];
}

/* covariant basis a1 = dr/dalpha, a2 = dr/dbeta of r = (X,Y,1)/sigma on the
 * prototype face, and the dual (contravariant) basis b = Qi a */
static void ecs_basis(double al, double be, double a1[3], double a2[3], double b1[3], double b2[3])
{
    double X = tan(al), Y = tan(be);
    double s2 = 1 + X * X + Y * Y, s = sqrt(s2), s3 = s2 * s;
    double cx = 1 + X * X, cy = 1 + Y * Y;
    a1[0] = cx * cy / s3;
    a1[1] = -X * Y * cx / s3;
    a1[2] = -X * cx / s3;
    a2[0] = -X * Y * cy / s3;
    a2[1] = cx * cy / s3;
    a2[2] = -Y * cy / s3;
    double q11 = s2 / cx, q12 = s2 * X * Y / (cx * cy), q22 = s2 / cy;
    for (int k = 0; k < 3; k++) {
        b1[k] = q11 * a1[k] + q12 * a2[k];
        b2[k] = q12 * a1[k] + q22 * a2[k];
    }
}

static void build_geometry(swo *s)
{
    int N = s->N;
    /* quadrature_factory_mod.f90 L89-188 and sbp_operators_collection_mod.f90 L169-197 */
    double Ai[6], Ac[6];
    int ni, nc;
    if (s->order == 63) {
        ni = 5; nc = 6;
        for (int k = 0; k < 5; k++) Ai[k] = H63_I[k];
        for (int k = 0; k < 6; k++) Ac[k] = H63_C[k];
    } else if (s->order == 21) {
        ni = 2; nc = 2;
        Ai[0] = 0.5; Ai[1] = 1;
        Ac[0] = 1; Ac[1] = 1;
    } else {
        ni = 5; nc = 4;
        Ai[0] = 7.0 / 18; Ai[1] = 9.0 / 8; Ai[2] = 1; Ai[3] = 71.0 / 72; Ai[4] = 1;
        Ac[0] = 13.0 / 12; Ac[1] = 7.0 / 8; Ac[2] = 25.0 / 24; Ac[3] = 1;
    }
    for (int m = 0; m < 3; m++) {
        mtype t = s->mt[m];
        int nx = npts_axis(N, t.xt), ny = npts_axis(N, t.yt);
        size_t n = (size_t)6 * nx * ny;
        geom *G = &s->G[m];
        G->J = malloc(n * sizeof(double));
        G->Q11 = malloc(n * sizeof(double));
        G->Q12 = malloc(n * sizeof(double));
        G->Q22 = malloc(n * sizeof(double));
        G->rx = malloc(n * sizeof(double));
        G->ry = malloc(n * sizeof(double));
        G->rz = malloc(n * sizeof(double));
        G->w = malloc(n * sizeof(double));
        double shx = t.xt ? 0.5 : 0.0, shy = t.yt ? 0.5 : 0.0;
        for (int p = 0; p < 6; p++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    size_t id = ((size_t)p * ny + (j - 1)) * nx + (i - 1);
                    double al = -0.25 * PI + (i - 1 + shx) * s->hx;
                    double be = -0.25 * PI + (j - 1 + shy) * s->hx;
                    double X = tan(al), Y = tan(be);
                    double sg = sqrt(1 + X * X + Y * Y);
                    double pr[3] = {X / sg, Y / sg, 1 / sg}, r[3];
                    proto2real(p, pr, r);
                    G->rx[id] = r[0];
                    G->ry[id] = r[1];
                    G->rz[id] = r[2];
                    G->J[id] = (1 + X * X) * (1 + Y * Y) / (sg * sg * sg);
                    G->Q11[id] = sg * sg / (1 + X * X);
                    G->Q12[id] = sg * sg * X * Y / ((1 + X * X) * (1 + Y * Y));
                    G->Q22[id] = sg * sg / (1 + Y * Y);
                    /* sbp_quadrature_mod.f90 L22-41: J*hx*hy*Ax(i)*Ay(j) */
                    const double *Ax = t.xt ? Ac : Ai, *Ay = t.yt ? Ac : Ai;
                    int nex = t.xt ? nc : ni, ney = t.yt ? nc : ni;
                    double wx = 1, wy = 1;
                    if (i <= nex) wx = Ax[i - 1];
                    if (nx - i + 1 <= nex) wx = Ax[nx - i];
                    if (j <= ney) wy = Ay[j - 1];
                    if (ny - j + 1 <= ney) wy = Ay[ny - j];
                    G->w[id] = G->J[id] * s->hx * s->hx * wx * wy;
                }
    }
    /* coriolis_factory_mod.f90 calc_coriolis_parameter L239-280 and the
     * f*J*J product of create_coriolis_Ch_sbp L188-194 */
    int nxy = swo_npts(s, SWO_MESH_XY);
    s->fJ2 = malloc(nxy * sizeof(double));
    double an = sqrt(s->axis[0] * s->axis[0] + s->axis[1] * s->axis[1] + s->axis[2] * s->axis[2]);
    (void)an;
    for (int id = 0; id < nxy; id++) {
        const geom *G = &s->G[SWO_MESH_XY];
        double f;
        if (s->cor_const) {
            f = s->omega;
        } else {
            double r2 = G->rx[id] * G->rx[id] + G->ry[id] * G->ry[id] + G->rz[id] * G->rz[id];
            double pr = (s->axis[0] * G->rx[id] + s->axis[1] * G->ry[id] + s->axis[2] * G->rz[id]) / sqrt(r2);
            f = 2 * s->omega * pr;
        }
        s->fJ2[id] = f * G->J[id] * G->J[id];
    }
    /* covariant edge-sync weights: ecs_Ah_vec_sync_factory_mod.f90 L97-152.
     * Each weight is (own covariant vector across the edge) . (neighbour's
     * contravariant along-edge vector) with the neighbour's prototype face
     * folded onto this one (a 90 degree rotation about the shared edge). */
    s->qb = malloc((N + 1) * sizeof(double));
    s->qt = malloc((N + 1) * sizeof(double));
    s->ql = malloc((N + 1) * sizeof(double));
    s->qr = malloc((N + 1) * sizeof(double));
    for (int i = 1; i <= N + 1; i++) {
        double c = -0.25 * PI + (i - 1) * s->hx, e = 0.25 * PI;
        double a1[3], a2[3], b1[3], b2[3], A1[3], A2[3], B1[3], B2[3], v[3];
        /* bottom: own a2 at beta=-pi/4; neighbour b1 at its top edge */
        ecs_basis(c, -e, a1, a2, b1, b2);
        ecs_basis(c, e, A1, A2, B1, B2);
        v[0] = B1[0]; v[1] = -B1[2]; v[2] = B1[1];
        s->qb[i - 1] = a2[0] * v[0] + a2[1] * v[1] + a2[2] * v[2];
        /* top */
        ecs_basis(c, e, a1, a2, b1, b2);
        ecs_basis(c, -e, A1, A2, B1, B2);
        v[0] = B1[0]; v[1] = B1[2]; v[2] = -B1[1];
        s->qt[i - 1] = a2[0] * v[0] + a2[1] * v[1] + a2[2] * v[2];
        /* left: own a1 at alpha=-pi/4; neighbour b2 at its right edge */
        ecs_basis(-e, c, a1, a2, b1, b2);
        ecs_basis(e, c, A1, A2, B1, B2);
        v[0] = -B2[2]; v[1] = B2[1]; v[2] = B2[0];
        s->ql[i - 1] = a1[0] * v[0] + a1[1] * v[1] + a1[2] * v[2];
        /* right */
        ecs_basis(e, c, a1, a2, b1, b2);
        ecs_basis(-e, c, A1, A2, B1, B2);
        v[0] = B2[2]; v[1] = B2[1]; v[2] = -B2[0];
        s->qr[i - 1] = a1[0] * v[0] + a1[1] * v[1] + a1[2] * v[2];
    }
}

swo *swo_create(int N, int order, double radius, double grav, double h_mean,
                double omega, const double axis[3], int coriolis_const)
{
    if (!(order == 21 || order == 42 || order == 63)) return NULL;
    if (order == 21 && N < 4) return NULL;
    if (order == 42 && N < 8) return NULL;
    if (order == 63 && N < 12) return NULL;
    swo *s = calloc(1, sizeof(swo));
    s->N = N;
    s->order = order;
    s->a = radius;
    s->g = grav;
    s->H0 = h_mean;
    s->omega = omega;
    for (int k = 0; k < 3; k++) s->axis[k] = axis[k];
    s->cor_const = coriolis_const;
    /* mesh_factory_mod.f90 L36-37: hx = (pi/2) / (nx - 1 + 2 shift) = pi/(2N) */
    s->hx = 0.5 * PI / N;
    s->mt[SWO_MESH_XY] = MT_XY;
    s->mt[SWO_MESH_U] = MT_U;
    s->mt[SWO_MESH_V] = MT_V;
    build_geometry(s);
    build_xy_map(&s->m_xy, N, 1, 0);
    build_xy_map(&s->m_xyv, N, 1, 1);
    build_ch_map(&s->m_ch, N, order == 21 ? 2 : (order == 42 ? 3 : 4));
    build_tan_map(&s->m_tan, N);
    fld_init(&s->h, N + 1, N + 1);
    fld_init(&s->u, N, N + 1);
    fld_init(&s->v, N + 1, N);
    fld_init(&s->gx, N, N + 1);
    fld_init(&s->gy, N + 1, N);
    fld_init(&s->ut, N, N + 1);
    fld_init(&s->vt, N + 1, N);
    fld_init(&s->w1, N + 1, N + 1);
    fld_init(&s->w2, N + 1, N + 1);
    fld_init(&s->up, N + 1, N + 1);
    fld_init(&s->vp, N + 1, N + 1);
    fld_init(&s->Ju, N, N + 1);
    fld_init(&s->Jv, N + 1, N);
    fld_init(&s->Dx, N + 1, N + 1);
    fld_init(&s->Dy, N + 1, N + 1);
    fld_init(&s->div, N + 1, N + 1);
    int n3[3] = {swo_npts(s, SWO_MESH_U), swo_npts(s, SWO_MESH_V), swo_npts(s, SWO_MESH_XY)};
    for (int c = 0; c < 3; c++) {
        for (int r = 0; r < 4; r++) s->k[r][c] = malloc(n3[c] * sizeof(double));
        s->y[c] = malloc(n3[c] * sizeof(double));
    }
    return s;
}

void swo_destroy(swo *s)
{
    if (!s) return;
    for (int m = 0; m < 3; m++) {
        geom *G = &s->G[m];
        free(G->J); free(G->Q11); free(G->Q12); free(G->Q22);
        free(G->rx); free(G->ry); free(G->rz); free(G->w);
    }
    free(s->fJ2); free(s->qb); free(s->qt); free(s->ql); free(s->qr);
    free(s->m_xy.e); free(s->m_xyv.e); free(s->m_ch.e); free(s->m_tan.e);
    fld *F[] = {&s->h, &s->u, &s->v, &s->gx, &s->gy, &s->ut, &s->vt, &s->w1, &s->w2,
                &s->up, &s->vp, &s->Ju, &s->Jv, &s->Dx, &s->Dy, &s->div};
    for (size_t k = 0; k < sizeof(F) / sizeof(F[0]); k++) fld_free(F[k]);
    for (int c = 0; c < 3; c++) {
        for (int r = 0; r < 4; r++) free(s->k[r][c]);
        free(s->y[c]);
    }
    free(s);
}

/* packed index of (p,i,j) on mesh m */
static inline size_t PK(const swo *s, int m, int p, int i, int j)
{
    int nx = npts_axis(s->N, s->mt[m].xt), ny = npts_axis(s->N, s->mt[m].yt);
    return ((size_t)p * ny + (j - 1)) * nx + (i - 1);
}

/* ------------------------------------------------------------------ */
/* operators (all act on the halo-padded work fields inside swo)       */
/* ------------------------------------------------------------------ */

/* grad_sbp_sat_mod.f90 L24-76: h (xy) -> gx (mesh y), gy (mesh x) */
static void op_grad(swo *s)
{
    int N = s->N;
    if (g_mutation != 3) apply_map(&s->m_xy, &s->h, &s->h);
    for (int p = 0; p < 6; p++) {
        OMP_FOR
        for (int j = 1; j <= N + 1; j++) {
            d_i2c(s->order, AT(&s->h, p, 0, j), 1, N, AT(&s->gx, p, 0, j), 1);
            for (int i = 1; i <= N; i++) *AT(&s->gx, p, i, j) /= (s->hx * s->a);
        }
        OMP_FOR
        for (int i = 1; i <= N + 1; i++) {
            d_i2c(s->order, AT(&s->h, p, i, 0), s->h.sx, N, AT(&s->gy, p, i, 0), s->gy.sx);
            for (int j = 1; j <= N; j++) *AT(&s->gy, p, i, j) /= (s->hx * s->a);
        }
    }
}

/* co2contra_Ch_mod.f90 L27-94: covariant (u,v) -> contravariant (ut,vt) */
static void op_co2contra(swo *s)
{
    int N = s->N;
    const geom *Gp = &s->G[SWO_MESH_XY], *Gu = &s->G[SWO_MESH_U], *Gv = &s->G[SWO_MESH_V];
    for (int p = 0; p < 6; p++) {
        /* uv -> xy (c2i): uq along x, vq along y */
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            i_c2i(s->order, AT(&s->u, p, 0, j), 1, N, AT(&s->w2, p, 0, j), 1);
        OMP_FOR
        for (int i = 1; i <= N + 1; i++)
            i_c2i(s->order, AT(&s->v, p, i, 0), s->v.sx, N, AT(&s->w1, p, i, 0), s->w1.sx);
        /* at xy: w1 <- vq*J*Qi12 (feeds ut), w2 <- uq*J*Qi12 (feeds vt) */
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N + 1; i++) {
                size_t id = PK(s, SWO_MESH_XY, p, i, j);
                double c = Gp->J[id] * Gp->Q12[id];
                *AT(&s->w1, p, i, j) *= c;
                *AT(&s->w2, p, i, j) *= c;
            }
        /* xy -> uv (i2c) and finalize */
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            i_i2c(s->order, AT(&s->w1, p, 0, j), 1, N, AT(&s->ut, p, 0, j), 1);
        OMP_FOR
        for (int i = 1; i <= N + 1; i++)
            i_i2c(s->order, AT(&s->w2, p, i, 0), s->w2.sx, N, AT(&s->vt, p, i, 0), s->vt.sx);
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N; i++) {
                size_t id = PK(s, SWO_MESH_U, p, i, j);
                *AT(&s->ut, p, i, j) = Gu->Q11[id] * *AT(&s->u, p, i, j) + *AT(&s->ut, p, i, j) / Gu->J[id];
            }
        OMP_FOR
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N + 1; i++) {
                size_t id = PK(s, SWO_MESH_V, p, i, j);
                *AT(&s->vt, p, i, j) = Gv->Q22[id] * *AT(&s->v, p, i, j) + *AT(&s->vt, p, i, j) / Gv->J[id];
            }
    }
}

/* covariant vector sync at xy edge points, ecs_Ah_vec_sync_mod.f90 L34-125
 * called with (v,u) swapped for covariant components (L51-53).  Reads the
 * neighbour's components from the depth-1 halo (the duplicate points). */
static void sync_covariant(swo *s, fld *up, fld *vp)
{
    int N = s->N, e = N + 1;
    const double *qb = s->qb, *qt = s->qt, *ql = s->ql, *qr = s->qr;
    double *zero = NULL;
    if (g_mutation == 2) {
        zero = calloc(N + 1, sizeof(double));
        qb = qt = ql = qr = zero;
    }
    for (int p = 0; p < 6; p++) {
#define U(i, j) (*AT(vp, p, i, j)) /* routine's first component  */
#define V(i, j) (*AT(up, p, i, j)) /* routine's second component */
        double u11 = (U(1, 1) + U(0, 1) + U(1, 0) + qb[0] * V(1, 0)) / 3.0;
        double v11 = (V(1, 1) + V(0, 1) + V(1, 0) + ql[0] * U(0, 1)) / 3.0;
        double ue1 = (U(e, 1) + U(e + 1, 1) + U(e, 0) + qb[e - 1] * V(e, 0)) / 3.0;
        double ve1 = (V(e, 1) + V(e + 1, 1) + V(e, 0) + qr[0] * U(e + 1, 1)) / 3.0;
        double u1e = (U(1, e) + U(0, e) + U(1, e + 1) + qt[0] * V(1, e + 1)) / 3.0;
        double v1e = (V(1, e) + V(0, e) + V(1, e + 1) + ql[e - 1] * U(0, e)) / 3.0;
        double uee = (U(e, e) + U(e + 1, e) + U(e, e + 1) + qt[e - 1] * V(e, e + 1)) / 3.0;
        double vee = (V(e, e) + V(e + 1, e) + V(e, e + 1) + qr[e - 1] * U(e + 1, e)) / 3.0;
        U(1, 1) = u11; V(1, 1) = v11;
        U(e, 1) = ue1; V(e, 1) = ve1;
        U(1, e) = u1e; V(1, e) = v1e;
        U(e, e) = uee; V(e, e) = vee;
        for (int j = 2; j <= N; j++) {
            U(1, j) = 0.5 * (U(0, j) + U(1, j));
            V(1, j) = 0.5 * (V(1, j) + V(0, j) + ql[j - 1] * U(0, j));
            U(e, j) = 0.5 * (U(e, j) + U(e + 1, j));
            V(e, j) = 0.5 * (V(e, j) + V(e + 1, j) + qr[j - 1] * U(e + 1, j));
        }
        for (int i = 2; i <= N; i++) {
            V(i, 1) = 0.5 * (V(i, 0) + V(i, 1));
            U(i, 1) = 0.5 * (U(i, 0) + U(i, 1) + qb[i - 1] * V(i, 0));
            V(i, e) = 0.5 * (V(i, e) + V(i, e + 1));
            U(i, e) = 0.5 * (U(i, e) + U(i, e + 1) + qt[i - 1] * V(i, e + 1));
        }
#undef U
#undef V
    }
    free(zero);
}

/* coriolis_Ch_mod.f90 calc_coriolis L42-69: (ut,vt) -> (cu on y, cv on x) */
static void op_coriolis(swo *s, fld *cu, fld *cv)
{
    int N = s->N;
    const geom *Gu = &s->G[SWO_MESH_U], *Gv = &s->G[SWO_MESH_V];
    for (int p = 0; p < 6; p++) {
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            i_c2i(s->order, AT(&s->ut, p, 0, j), 1, N, AT(&s->w1, p, 0, j), 1);
        OMP_FOR
        for (int i = 1; i <= N + 1; i++)
            i_c2i(s->order, AT(&s->vt, p, i, 0), s->vt.sx, N, AT(&s->w2, p, i, 0), s->w2.sx);
        /* calc_cori_at_p_tile L97-120: (up,vp) <- fJ2*(vp,-up) */
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N + 1; i++) {
                double f = s->fJ2[PK(s, SWO_MESH_XY, p, i, j)];
                *AT(&s->up, p, i, j) = f * *AT(&s->w2, p, i, j);
                *AT(&s->vp, p, i, j) = -f * *AT(&s->w1, p, i, j);
            }
    }
    apply_map(&s->m_xyv, &s->up, &s->vp);
    sync_covariant(s, &s->up, &s->vp);
    for (int p = 0; p < 6; p++) {
        OMP_FOR
        for (int j = 1; j <= N + 1; j++) {
            i_i2c(s->order, AT(&s->up, p, 0, j), 1, N, AT(cu, p, 0, j), 1);
            for (int i = 1; i <= N; i++) *AT(cu, p, i, j) /= Gu->J[PK(s, SWO_MESH_U, p, i, j)];
        }
        OMP_FOR
        for (int i = 1; i <= N + 1; i++) {
            i_i2c(s->order, AT(&s->vp, p, i, 0), s->vp.sx, N, AT(cv, p, i, 0), cv->sx);
            for (int j = 1; j <= N; j++) *AT(cv, p, i, j) /= Gv->J[PK(s, SWO_MESH_V, p, i, j)];
        }
    }
}

/* halo_Ah_scalar_sync_mod.f90 L43-98: average the duplicated edge values */
static void sync_scalar(swo *s, fld *f)
{
    int N = s->N, e = N + 1;
    apply_map(&s->m_xy, f, f);
    for (int p = 0; p < 6; p++) {
#define F(i, j) (*AT(f, p, i, j))
        double c11 = (F(1, 1) + F(0, 1) + F(1, 0)) / 3.0;
        double ce1 = (F(e, 1) + F(e + 1, 1) + F(e, 0)) / 3.0;
        double c1e = (F(1, e) + F(0, e) + F(1, e + 1)) / 3.0;
        double cee = (F(e, e) + F(e + 1, e) + F(e, e + 1)) / 3.0;
        F(1, 1) = c11; F(e, 1) = ce1; F(1, e) = c1e; F(e, e) = cee;
        for (int j = 2; j <= N; j++) {
            F(1, j) = 0.5 * (F(0, j) + F(1, j));
            F(e, j) = 0.5 * (F(e, j) + F(e + 1, j));
        }
        for (int i = 2; i <= N; i++) {
            F(i, 1) = 0.5 * (F(i, 1) + F(i, 0));
            F(i, e) = 0.5 * (F(i, e) + F(i, e + 1));
        }
#undef F
    }
}

/* div_sbp_sat_mod.f90 L29-80 + projection: (ut,vt) -> div (xy) */
static void op_div(swo *s)
{
    int N = s->N;
    const geom *Gp = &s->G[SWO_MESH_XY], *Gu = &s->G[SWO_MESH_U], *Gv = &s->G[SWO_MESH_V];
    for (int p = 0; p < 6; p++) {
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N; i++)
                *AT(&s->Ju, p, i, j) = *AT(&s->ut, p, i, j) * Gu->J[PK(s, SWO_MESH_U, p, i, j)];
        OMP_FOR
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N + 1; i++)
                *AT(&s->Jv, p, i, j) = *AT(&s->vt, p, i, j) * Gv->J[PK(s, SWO_MESH_V, p, i, j)];
    }
    apply_map(&s->m_ch, &s->Ju, &s->Jv);
    for (int p = 0; p < 6; p++) {
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            d_c2i(s->order, AT(&s->Ju, p, 0, j), 1, N, AT(&s->Dx, p, 0, j), 1);
        OMP_FOR
        for (int i = 1; i <= N + 1; i++)
            d_c2i(s->order, AT(&s->Jv, p, i, 0), s->Jv.sx, N, AT(&s->Dy, p, i, 0), s->Dy.sx);
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N + 1; i++)
                *AT(&s->div, p, i, j) = (*AT(&s->Dx, p, i, j) / s->hx + *AT(&s->Dy, p, i, j) / s->hx) /
                                        (Gp->J[PK(s, SWO_MESH_XY, p, i, j)] * s->a);
    }
    if (g_mutation != 5) sync_scalar(s, &s->div);
}

/* operator_swm_lin_mod.f90 apply L66-108 (fields already loaded) */
static void apply_loaded(swo *s, double *du, double *dv, double *dh)
{
    int N = s->N;
    op_grad(s);
    op_co2contra(s);
    /* coriolis writes into the ut/vt-shaped fields cu=Dx-like work: reuse Ju/Jv
     * after div needs them, so use dedicated outputs here */
    fld cu, cv;
    fld_init(&cu, N, N + 1);
    fld_init(&cv, N + 1, N);
    op_coriolis(s, &cu, &cv);
    op_div(s);
    for (int p = 0; p < 6; p++) {
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N; i++)
                du[PK(s, SWO_MESH_U, p, i, j)] = -s->g * *AT(&s->gx, p, i, j) + *AT(&cu, p, i, j);
        OMP_FOR
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N + 1; i++)
                dv[PK(s, SWO_MESH_V, p, i, j)] = -s->g * *AT(&s->gy, p, i, j) + *AT(&cv, p, i, j);
        OMP_FOR
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N + 1; i++)
                dh[PK(s, SWO_MESH_XY, p, i, j)] = -s->H0 * *AT(&s->div, p, i, j);
    }
    fld_free(&cu);
    fld_free(&cv);
}

void swo_apply(swo *s, const double *u, const double *v, const double *h,
               double *du, double *dv, double *dh)
{
    fld_load(&s->u, u);
    fld_load(&s->v, v);
    fld_load(&s->h, h);
    apply_loaded(s, du, dv, dh);
}

void swo_grad(swo *s, const double *h, double *gx, double *gy)
{
    fld_load(&s->h, h);
    op_grad(s);
    fld_store(&s->gx, gx);
    fld_store(&s->gy, gy);
}

void swo_co2contra(swo *s, const double *u, const double *v, double *ut, double *vt)
{
    fld_load(&s->u, u);
    fld_load(&s->v, v);
    op_co2contra(s);
    fld_store(&s->ut, ut);
    fld_store(&s->vt, vt);
}

void swo_coriolis(swo *s, const double *ut, const double *vt, double *cu, double *cv)
{
    fld_load(&s->ut, ut);
    fld_load(&s->vt, vt);
    fld a, b;
    fld_init(&a, s->N, s->N + 1);
    fld_init(&b, s->N + 1, s->N);
    op_coriolis(s, &a, &b);
    fld_store(&a, cu);
    fld_store(&b, cv);
    fld_free(&a);
    fld_free(&b);
}

void swo_div(swo *s, const double *ut, const double *vt, double *div)
{
    fld_load(&s->ut, ut);
    fld_load(&s->vt, vt);
    op_div(s);
    fld_store(&s->div, div);
}

/* rk4_mod.f90 step_rk4 L20-42, with stvec assign/update of
 * grid_field_mod.f90 L460-534 (interior points only) */
void swo_rk4(swo *s, double *u, double *v, double *h, double dt, int nsteps)
{
    int n3[3] = {swo_npts(s, SWO_MESH_U), swo_npts(s, SWO_MESH_V), swo_npts(s, SWO_MESH_XY)};
    double *v0[3] = {u, v, h};
    for (int it = 0; it < nsteps; it++) {
        swo_apply(s, v0[0], v0[1], v0[2], s->k[0][0], s->k[0][1], s->k[0][2]);
        for (int c = 0; c < 3; c++)
            OMP_FOR
            for (int n = 0; n < n3[c]; n++) s->y[c][n] = 1.0 * v0[c][n] + 0.5 * dt * s->k[0][c][n];
        swo_apply(s, s->y[0], s->y[1], s->y[2], s->k[1][0], s->k[1][1], s->k[1][2]);
        for (int c = 0; c < 3; c++)
            OMP_FOR
            for (int n = 0; n < n3[c]; n++) s->y[c][n] = 1.0 * v0[c][n] + 0.5 * dt * s->k[1][c][n];
        swo_apply(s, s->y[0], s->y[1], s->y[2], s->k[2][0], s->k[2][1], s->k[2][2]);
        for (int c = 0; c < 3; c++)
            OMP_FOR
            for (int n = 0; n < n3[c]; n++) s->y[c][n] = 1.0 * v0[c][n] + dt * s->k[2][c][n];
        swo_apply(s, s->y[0], s->y[1], s->y[2], s->k[3][0], s->k[3][1], s->k[3][2]);
        for (int c = 0; c < 3; c++)
            OMP_FOR
            for (int n = 0; n < n3[c]; n++) {
                v0[c][n] = v0[c][n] + dt / 6.0 * s->k[0][c][n] + dt / 3.0 * s->k[1][c][n];
                v0[c][n] = v0[c][n] + dt / 3.0 * s->k[2][c][n] + dt / 6.0 * s->k[3][c][n];
            }
    }
}

/* ------------------------------------------------------------------ */
/* geometry export, integrals, test fields                             */
/* ------------------------------------------------------------------ */
void swo_metric(const swo *s, int mesh, double *J, double *Qi11, double *Qi12,
                double *Qi22, double *rx, double *ry, double *rz)
{
    int n = swo_npts(s, mesh);
    const geom *G = &s->G[mesh];
    size_t b = n * sizeof(double);
    if (J) memcpy(J, G->J, b);
    if (Qi11) memcpy(Qi11, G->Q11, b);
    if (Qi12) memcpy(Qi12, G->Q12, b);
    if (Qi22) memcpy(Qi22, G->Q22, b);
    if (rx) memcpy(rx, G->rx, b);
    if (ry) memcpy(ry, G->ry, b);
    if (rz) memcpy(rz, G->rz, b);
}

void swo_quad_weights(const swo *s, int mesh, double *w)
{
    memcpy(w, s->G[mesh].w, swo_npts(s, mesh) * sizeof(double));
}

void swo_coriolis_fJ2(const swo *s, double *fJ2)
{
    memcpy(fJ2, s->fJ2, swo_npts(s, SWO_MESH_XY) * sizeof(double));
}

void swo_sync_weights(const swo *s, double *qb, double *qt, double *ql, double *qr)
{
    size_t b = (s->N + 1) * sizeof(double);
    memcpy(qb, s->qb, b);
    memcpy(qt, s->qt, b);
    memcpy(ql, s->ql, b);
    memcpy(qr, s->qr, b);
}

double swo_mass(const swo *s, const double *h)
{
    double m = 0;
    int n = swo_npts(s, SWO_MESH_XY);
    for (int k = 0; k < n; k++) m += h[k] * s->G[SWO_MESH_XY].w[k];
    return m;
}

/* Linear SBP energy E = g/2 <h,h>_xy + H0/2 (<u,ut>_y + <v,vt>_x).  The
 * reference keeps its own linear-energy code commented out
 * (operator_swm_lin_mod.f90 L244-269); this is the quadratic form the SBP
 * discretisation conserves. */
double swo_energy(swo *s, const double *u, const double *v, const double *h)
{
    int nu = swo_npts(s, SWO_MESH_U), nv = swo_npts(s, SWO_MESH_V), nh = swo_npts(s, SWO_MESH_XY);
    double *ut = malloc(nu * sizeof(double)), *vt = malloc(nv * sizeof(double));
    swo_co2contra(s, u, v, ut, vt);
    double eh = 0, eu = 0, ev = 0;
    for (int k = 0; k < nh; k++) eh += h[k] * h[k] * s->G[SWO_MESH_XY].w[k];
    for (int k = 0; k < nu; k++) eu += u[k] * ut[k] * s->G[SWO_MESH_U].w[k];
    for (int k = 0; k < nv; k++) ev += v[k] * vt[k] * s->G[SWO_MESH_V].w[k];
    free(ut);
    free(vt);
    return 0.5 * s->g * eh + 0.5 * s->H0 * (eu + ev);
}

/* cell-centre (mesh o) geometry: J, the SBP_C quadrature weight and the
 * Coriolis parameter (metric_2d_ecs_mod.f90 L338-411 at the o points,
 * sbp_quadrature_mod.f90 L22-41 with the centre weights in both
 * directions, coriolis_factory_mod.f90 calc_coriolis_parameter L239-280) */
static void centre_geom(const swo *s, int p, int i, int j, double *J, double *w, double *f)
{
    const int N = s->N;
    const double al = -0.25 * PI + (i - 0.5) * s->hx, be = -0.25 * PI + (j - 0.5) * s->hx;
    const double X = tan(al), Y = tan(be), sg = sqrt(1 + X * X + Y * Y);
    *J = (1 + X * X) * (1 + Y * Y) / (sg * sg * sg);
    double Ac[6] = {1, 1, 1, 1, 1, 1};
    int nc = 0;
    if (s->order == 42) {
        Ac[0] = 13.0 / 12; Ac[1] = 7.0 / 8; Ac[2] = 25.0 / 24; Ac[3] = 1;
        nc = 4;
    } else if (s->order == 63) {
        for (int k = 0; k < 6; k++) Ac[k] = H63_C[k];
        nc = 6;
    }
    double wx = 1, wy = 1;
    if (i <= nc) wx = Ac[i - 1];
    if (N - i + 1 <= nc) wx = Ac[N - i];
    if (j <= nc) wy = Ac[j - 1];
    if (N - j + 1 <= nc) wy = Ac[N - j];
    *w = *J * s->hx * s->hx * wx * wy;
    if (s->cor_const) {
        *f = s->omega;
    } else {
        double pr[3] = {X / sg, Y / sg, 1 / sg}, r[3];
        proto2real(p, pr, r);
        double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        *f = 2 * s->omega * (s->axis[0] * r[0] + s->axis[1] * r[1] + s->axis[2] * r[2]) / rn;
    }
}

/* curl_sbp_sat_mod.f90 calc_curl L26-74 for Ch (curl_ch_sbp_sat_XX, mesh o):
 * Dy = i2c derivative of u along y, Dx = of v along x, each with its SAT
 * against the neighbour's tangential component at the duplicated edge
 * point; curl = (Dx/hx - Dy/hy) / (J a) at the cell centres.  u, v loaded
 * into s->u, s->v; the result into s->div (centres, (N, N) per panel). */
static void op_curl(swo *s)
{
    const int N = s->N;
    apply_map(&s->m_tan, &s->u, &s->v);
    for (int p = 0; p < 6; p++) {
        OMP_FOR
        for (int j = 1; j <= N; j++)  /* Dx at (centre i, centre j): v row j */
            d_i2c(s->order, AT(&s->v, p, 0, j), 1, N, AT(&s->Dx, p, 0, j), 1);
        OMP_FOR
        for (int i = 1; i <= N; i++)  /* Dy: u column i */
            d_i2c(s->order, AT(&s->u, p, i, 0), s->u.sx, N, AT(&s->Dy, p, i, 0), s->Dy.sx);
        OMP_FOR
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N; i++) {
                double J, w, f;
                centre_geom(s, p, i, j, &J, &w, &f);
                *AT(&s->div, p, i, j) = (*AT(&s->Dx, p, i, j) / s->hx - *AT(&s->Dy, p, i, j) / s->hx) / (J * s->a);
            }
    }
}

void swo_curl(swo *s, const double *u, const double *v, double *curl)
{
    const int N = s->N;
    fld_load(&s->u, u);
    fld_load(&s->v, v);
    op_curl(s);
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N; i++) curl[((size_t)p * N + (j - 1)) * N + (i - 1)] = *AT(&s->div, p, i, j);
}

/* print_integrals_diag + calc_energies (swm_output_diag_mod.f90 L156-234)
 * for the Ch staggering with covariant winds and no orography: out = hmin,
 * hmax, mass, TE, KE, PE, enstrophy.  h2uv = interp2d_p2uv_Ch_sbpXX (i2c
 * interpolation of h along x to u points, along y to v points); uv2q =
 * interp2d_uv2qvec_Ch_sbpXX (i2c along y of hu, along x of hv, to the
 * centres, averaged); the quadratures SBP_CXX of each mesh; grav as given
 * (the reference passes Earth_grav). */
void swo_integrals_diag(swo *s, const double *u, const double *v, const double *h, double grav, double out[7])
{
    const int N = s->N;
    const int nh = swo_npts(s, SWO_MESH_XY);
    double hmin = h[0], hmax = h[0], mass = 0, pe = 0, ke_u = 0, ke_v = 0, ens = 0;
    for (int k = 0; k < nh; k++) {
        hmin = h[k] < hmin ? h[k] : hmin;
        hmax = h[k] > hmax ? h[k] : hmax;
        mass += h[k] * s->G[SWO_MESH_XY].w[k];
        pe += 0.5 * grav * h[k] * h[k] * s->G[SWO_MESH_XY].w[k];
    }
    fld_load(&s->u, u);
    fld_load(&s->v, v);
    fld_load(&s->h, h);
    op_co2contra(s); /* s->ut, s->vt */
    /* hu at u points (s->Ju), hv at v points (s->Jv) */
    for (int p = 0; p < 6; p++) {
        for (int j = 1; j <= N + 1; j++) i_i2c(s->order, AT(&s->h, p, 0, j), 1, N, AT(&s->Ju, p, 0, j), 1);
        for (int i = 1; i <= N + 1; i++)
            i_i2c(s->order, AT(&s->h, p, i, 0), s->h.sx, N, AT(&s->Jv, p, i, 0), s->Jv.sx);
    }
    const geom *Gu = &s->G[SWO_MESH_U], *Gv = &s->G[SWO_MESH_V];
    for (int p = 0; p < 6; p++) {
        for (int j = 1; j <= N + 1; j++)
            for (int i = 1; i <= N; i++)
                ke_u += 0.5 * (*AT(&s->Ju, p, i, j) * *AT(&s->ut, p, i, j)) * *AT(&s->u, p, i, j) *
                        Gu->w[PK(s, SWO_MESH_U, p, i, j)];
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N + 1; i++)
                ke_v += 0.5 * (*AT(&s->Jv, p, i, j) * *AT(&s->vt, p, i, j)) * *AT(&s->v, p, i, j) *
                        Gv->w[PK(s, SWO_MESH_V, p, i, j)];
    }
    /* hq at the centres: i2c along y of hu (s->w1), along x of hv (s->w2) */
    for (int p = 0; p < 6; p++) {
        for (int i = 1; i <= N; i++)
            i_i2c(s->order, AT(&s->Ju, p, i, 0), s->Ju.sx, N, AT(&s->w1, p, i, 0), s->w1.sx);
        for (int j = 1; j <= N; j++) i_i2c(s->order, AT(&s->Jv, p, 0, j), 1, N, AT(&s->w2, p, 0, j), 1);
    }
    op_curl(s); /* s->div */
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N; i++) {
                double J, w, f;
                centre_geom(s, p, i, j, &J, &w, &f);
                const double hq = 0.5 * *AT(&s->w1, p, i, j) + 0.5 * *AT(&s->w2, p, i, j);
                const double c = *AT(&s->div, p, i, j) + f;
                ens += (0.5 * c * c) / hq * w;
            }
    out[0] = hmin;
    out[1] = hmax;
    out[2] = mass;
    out[4] = ke_u + ke_v;
    out[5] = pe;
    out[3] = out[4] + out[5];
    out[6] = ens;
}

/* ---- the implicit solve (operator_swm_lin_mod.f90 solve L110-146) ---- */

static double xy_dot(const swo *s, const double *a, const double *b)
{
    const int n = swo_npts(s, SWO_MESH_XY);
    const double *w = s->G[SWO_MESH_XY].w;
    double d = 0;
    for (int k = 0; k < n; k++) d += w[k] * a[k] * b[k];
    return d;
}

/* swm_helm_lin_operator_t%apply (swm_helm_operator_mod.f90 L56-81):
 * out = z - gamma div(co2contra(grad z)) */
static void helm_apply(swo *s, const double *z, double gamma, double *out, double *w)
{
    const int nu = swo_npts(s, SWO_MESH_U), nv = swo_npts(s, SWO_MESH_V), nh = swo_npts(s, SWO_MESH_XY);
    double *gx = w, *gy = gx + nu, *gxt = gy + nv, *gyt = gxt + nu;
    swo_grad(s, z, gx, gy);
    swo_co2contra(s, gx, gy, gxt, gyt);
    swo_div(s, gxt, gyt, out);
    for (int k = 0; k < nh; k++) out[k] = (-1.0 * (g_mutation == 13 ? -gamma : gamma)) * out[k];
    for (int k = 0; k < nh; k++) out[k] = out[k] + 1.0 * z[k];
}

/* operator_swm_lin_t%solve: vout = (I - dt L0)^-1 rhs for the linear
 * operator without Coriolis (L0), through the Helmholtz problem for h,
 * (I - H g dt^2 div co2contra grad) h = rhs_h - H dt div(co2contra(rhs_u,
 * rhs_v)), solved by cg_solver_t%solve (cg_solver_mod.f90 L19-90: x0 = the
 * right-hand side, relative residual in the h quadrature's l2 norm), then
 * u = rhs_u - g dt grad_x h, v = rhs_v - g dt grad_y h.  *iters: the CG
 * iterations done (0 when x0 already meets the tolerance, max_iter + 1 when
 * it never does); *relres: the final relative residual. */
void swo_solve(swo *s, const double *ru, const double *rv, const double *rh, double dt, double rel_tol, int max_iter,
               double *u, double *v, double *h, int *iters, double *relres)
{
    const int nu = swo_npts(s, SWO_MESH_U), nv = swo_npts(s, SWO_MESH_V), nh = swo_npts(s, SWO_MESH_XY);
    double *w = malloc((size_t)(2 * nu + 2 * nv) * sizeof(double));
    double *b = malloc(nh * sizeof(double)), *r = malloc(nh * sizeof(double)), *p = malloc(nh * sizeof(double));
    double *Ap = malloc(nh * sizeof(double)), *ut = malloc(nu * sizeof(double)), *vt = malloc(nv * sizeof(double));
    swo_co2contra(s, ru, rv, ut, vt);
    swo_div(s, ut, vt, b);
    for (int k = 0; k < nh; k++) b[k] = 1.0 * rh[k] + (-s->H0 * dt) * b[k];
    const double gamma = s->H0 * s->g * (dt * dt);
    double *x = h;
    memcpy(x, b, nh * sizeof(double));
    helm_apply(s, x, gamma, r, w);
    for (int k = 0; k < nh; k++) r[k] = -1.0 * r[k] + 1.0 * b[k];
    double rn = sqrt(xy_dot(s, r, r));
    const double bn = sqrt(xy_dot(s, b, b));
    int it = 0;
    /* a zero right-hand side has the solution zero (the reference's 0 / 0
     * would go on with NaN) */
    double rel = bn == 0.0 ? 0.0 : rn / bn;
    if (!(rel < rel_tol)) {
        memcpy(p, r, nh * sizeof(double));
        for (it = 1; it <= max_iter; it++) {
            helm_apply(s, p, gamma, Ap, w);
            const double alpha = rn * rn / xy_dot(s, p, Ap);
            for (int k = 0; k < nh; k++) x[k] = x[k] + alpha * p[k];
            for (int k = 0; k < nh; k++) r[k] = r[k] + (-alpha) * Ap[k];
            const double rnn = sqrt(xy_dot(s, r, r));
            rel = rnn / bn;
            if (rel < rel_tol) break;
            const double beta = (rnn / rn) * (rnn / rn);
            rn = rnn;
            for (int k = 0; k < nh; k++) p[k] = 1.0 * r[k] + beta * p[k];
        }
    }
    *iters = it;
    *relres = rel;
    double *gx = w, *gy = w + nu;
    swo_grad(s, x, gx, gy);
    for (int k = 0; k < nu; k++) u[k] = 1.0 * ru[k] + (-s->g * dt) * gx[k];
    for (int k = 0; k < nv; k++) v[k] = 1.0 * rv[k] + (-s->g * dt) * gy[k];
    free(w); free(b); free(r); free(p); free(Ap); free(ut); free(vt);
}

/* print_CFL_diag (swm_output_diag_mod.f90 L110-154), covariant winds: the
 * gravity-wave CFL sqrt(max(g / dx^2 h)) dt and the advective ones
 * sqrt(max |(ut / dx^2) u|) dt, sqrt(max |(vt / dy^2) v|) dt, ut and vt the
 * contravariant components (co2contra), dx = dy = a pi / (2N). */
void swo_cfl_diag(swo *s, const double *u, const double *v, const double *h, double dt, double grav, double out[3])
{
    const int nu = swo_npts(s, SWO_MESH_U), nv = swo_npts(s, SWO_MESH_V), nh = swo_npts(s, SWO_MESH_XY);
    double *uc = malloc(nu * sizeof(double)), *vc = malloc(nv * sizeof(double));
    swo_co2contra(s, u, v, uc, vc);
    const double dx = s->a * s->hx, cg = grav / (dx * dx), c = 1.0 / (dx * dx);
    double mg = -INFINITY, mu = 0, mv = 0;
    for (int k = 0; k < nh; k++) mg = fmax(mg, cg * h[k]);
    for (int k = 0; k < nu; k++) mu = fmax(mu, fabs(c * uc[k] * u[k]));
    for (int k = 0; k < nv; k++) mv = fmax(mv, fabs(c * vc[k] * v[k]));
    out[0] = sqrt(mg) * dt;
    out[1] = sqrt(mu) * dt;
    out[2] = sqrt(mv) * dt;
    free(uc);
    free(vc);
}

/* print_integral_tends (swm_output_diag_mod.f90 L236-342) of the linear
 * operator's tendency, Ch with covariant winds, no orography: out[0] =
 * TE_tend = sum_u w (I_x(h_t) ut u / 2 + I_x(h) ut u_t) + (the same on v) +
 * sum_xy w g h h_t; out[1] = Enstrophy_tend = sum_o w (zeta + f) zeta_t / hq
 * + sum_o w (-(zeta + f)^2 / 2) hq_t / hq / hq, with hq = I_y(I_x(h)) (the
 * first output of interp2d_uv2qvec), hq_t likewise of h_t, zeta = curl(u,
 * v), zeta_t = curl(u_t, v_t), (u_t, v_t, h_t) = apply(u, v, h).  out[2],
 * out[3]: the sums of the terms' absolute values (the scale against which
 * rounding in the two results is measured, both sums cancel heavily). */
void swo_integral_tends(swo *s, const double *u, const double *v, const double *h, double grav, double out[4])
{
    const int N = s->N;
    const int nu = swo_npts(s, SWO_MESH_U), nv = swo_npts(s, SWO_MESH_V), nh = swo_npts(s, SWO_MESH_XY);
    const int no = 6 * N * N;
    double *ut = malloc(nu * sizeof(double)), *vt = malloc(nv * sizeof(double)), *ht = malloc(nh * sizeof(double));
    double *uc = malloc(nu * sizeof(double)), *vc = malloc(nv * sizeof(double));
    double *hu = malloc(nu * sizeof(double)), *hv = malloc(nv * sizeof(double));
    double *hut = malloc(nu * sizeof(double)), *hvt = malloc(nv * sizeof(double));
    double *z = malloc(no * sizeof(double)), *zt = malloc(no * sizeof(double));
    double *hq = malloc(no * sizeof(double)), *hqt = malloc(no * sizeof(double));
    swo_apply(s, u, v, h, ut, vt, ht);
    swo_co2contra(s, u, v, uc, vc);
    /* h and h_t at the u and v points (interp2d_p2uv_Ch), then at the centres
     * (u_q of interp2d_uv2qvec_Ch: i2c along y of the u-point values) */
    const double *src[2] = {h, ht};
    double *du[2] = {hu, hut}, *dv[2] = {hv, hvt}, *dq[2] = {hq, hqt};
    for (int k = 0; k < 2; k++) {
        fld_load(&s->h, src[k]);
        for (int p = 0; p < 6; p++) {
            for (int j = 1; j <= N + 1; j++) i_i2c(s->order, AT(&s->h, p, 0, j), 1, N, AT(&s->Ju, p, 0, j), 1);
            for (int i = 1; i <= N + 1; i++)
                i_i2c(s->order, AT(&s->h, p, i, 0), s->h.sx, N, AT(&s->Jv, p, i, 0), s->Jv.sx);
            for (int i = 1; i <= N; i++)
                i_i2c(s->order, AT(&s->Ju, p, i, 0), s->Ju.sx, N, AT(&s->w1, p, i, 0), s->w1.sx);
        }
        fld_store(&s->Ju, du[k]);
        fld_store(&s->Jv, dv[k]);
        for (int p = 0; p < 6; p++)
            for (int j = 1; j <= N; j++)
                for (int i = 1; i <= N; i++) dq[k][((size_t)p * N + (j - 1)) * N + (i - 1)] = *AT(&s->w1, p, i, j);
    }
    swo_curl(s, u, v, z);
    swo_curl(s, ut, vt, zt);
    const double *wu = s->G[SWO_MESH_U].w, *wv = s->G[SWO_MESH_V].w, *wh = s->G[SWO_MESH_XY].w;
    double ke1 = 0, ke2 = 0, pe = 0, e1 = 0, e2 = 0, ste = 0, sen = 0;
    if (g_mutation != 6) {
        for (int k = 0; k < nu; k++) ke1 += 0.5 * (hut[k] * uc[k]) * u[k] * wu[k];
        for (int k = 0; k < nv; k++) ke1 += 0.5 * (hvt[k] * vc[k]) * v[k] * wv[k];
    }
    for (int k = 0; k < nu; k++) ke2 += (hu[k] * uc[k]) * ut[k] * wu[k];
    for (int k = 0; k < nv; k++) ke2 += (hv[k] * vc[k]) * vt[k] * wv[k];
    if (g_mutation != 7)
        for (int k = 0; k < nh; k++) pe += grav * h[k] * ht[k] * wh[k];
    for (int k = 0; k < nu; k++) ste += (0.5 * fabs(hut[k] * uc[k] * u[k]) + fabs(hu[k] * uc[k] * ut[k])) * wu[k];
    for (int k = 0; k < nv; k++) ste += (0.5 * fabs(hvt[k] * vc[k] * v[k]) + fabs(hv[k] * vc[k] * vt[k])) * wv[k];
    for (int k = 0; k < nh; k++) ste += fabs(grav * h[k] * ht[k]) * wh[k];
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= N; j++)
            for (int i = 1; i <= N; i++) {
                const size_t id = ((size_t)p * N + (j - 1)) * N + (i - 1);
                double J, w, f;
                centre_geom(s, p, i, j, &J, &w, &f);
                const double c = z[id] + (g_mutation == 9 ? 0.0 : f);
                e1 += (c * zt[id]) / hq[id] * w;
                if (g_mutation != 8) e2 += (((-0.5 * c * c) * hqt[id]) / hq[id]) / hq[id] * w;
                sen += (fabs(c * zt[id] / hq[id]) + fabs(0.5 * c * c * hqt[id] / hq[id] / hq[id])) * w;
            }
    out[0] = (ke2 + ke1) + pe;
    out[1] = e1 + e2;
    out[2] = ste;
    out[3] = sen;
    free(ut); free(vt); free(ht); free(uc); free(vc); free(hu); free(hv); free(hut); free(hvt);
    free(z); free(zt); free(hq); free(hqt);
}

/* ---- lat-lon output of h (latlon_outputer_mod.f90 latlon_write L33-63) ---- */

/* init_halo_xy_interp (ecs_halo_factory_mod.f90 L230-269): the virtual
 * point at along-edge index i, depth j beyond a panel edge, projected onto
 * the neighbour's row j inside; cubic weights over ind-1 .. ind+2 */
static void ecs_xy_weights(int n, double hx, int i, int j, double w[4], int *ind)
{
    const double zb = 0.25 * PI + j * hx, za0 = -0.25 * PI + (i - 1.0) * hx;
    const double za = atan(tan(za0) / tan(zb));
    int k = (int)floor((za + 0.25 * PI) / hx) + 1;
    k = k < 2 ? 2 : k;
    k = k > n - 2 ? n - 2 : k;
    const double z = (za + 0.25 * PI) / hx - (k - 1.0);
    w[0] = -z * (z - 1.0) * (z - 2.0) / 6.0;
    w[1] = (z + 1.0) * (z - 1.0) * (z - 2.0) / 2.0;
    w[2] = -(z + 1.0) * z * (z - 2.0) / 2.0;
    w[3] = (z + 1.0) * z * (z - 1.0) / 6.0;
    *ind = k;
}

/* ext_ecs_tile_halo (ecs_halo_xy_mod.f90 L88-203) on one whole panel of the
 * exchanged field E (xy halo depth hw + 1: the neighbour's duplicate of the
 * edge at depth 1, its row j inside at depth j + 1): the corners' three
 * copies averaged and their diagonals set, then edges bottom, top, left,
 * right: edge points averaged with the duplicate, the virtual rows 1..hw
 * interpolated along the neighbour's rows. */
static void ecs_xy_halo_panel(fld *E, int p, int n, double hx, int hw)
{
#define F(i, j) (*AT(E, p, (i), (j)))
    /* corners 1 (1,1), 2 (n,1), 3 (n,n), 4 (1,n) */
    F(1, 1) = (F(1, 1) + F(1, 0) + F(0, 1)) / 3.0;
    for (int j = 1; j <= hw; j++) {
        F(1 - j, 1 - j) = 0.5 * (F(-j, 1) + F(1, -j));
        F(-j, 1) = F(1 - j, 1 - j);
        F(1, -j) = F(1 - j, 1 - j);
    }
    F(n, 1) = (F(n, 1) + F(n, 0) + F(n + 1, 1)) / 3.0;
    for (int j = 1; j <= hw; j++) {
        F(n + j, 1 - j) = 0.5 * (F(n + j + 1, 1) + F(n, -j));
        F(n + j + 1, 1) = F(n + j, 1 - j);
        F(n, -j) = F(n + j, 1 - j);
    }
    F(n, n) = (F(n, n) + F(n, n + 1) + F(n + 1, n)) / 3.0;
    for (int j = 1; j <= hw; j++) {
        F(n + j, n + j) = 0.5 * (F(n + j + 1, n) + F(n, n + j + 1));
        F(n + j + 1, n) = F(n + j, n + j);
        F(n, n + j + 1) = F(n + j, n + j);
    }
    F(1, n) = (F(1, n) + F(0, n) + F(1, n + 1)) / 3.0;
    for (int j = 1; j <= hw; j++) {
        F(1 - j, n + j) = 0.5 * (F(1, n + j + 1) + F(-j, n));
        F(1, n + j + 1) = F(1 - j, n + j);
        F(-j, n) = F(1 - j, n + j);
    }
    double *buf = malloc((size_t)(n + 2 * HALO) * sizeof(double)), *zb = buf + HALO - 1; /* zb[k], k >= 1 - HALO */
    double w[4];
    int ind;
    for (int e = 0; e < 4; e++) {
        /* edge 1 bottom, 2 top, 3 left, 4 right; along-edge index a, depth j */
        if (g_mutation != 10)
            for (int a = 2; a <= n - 1; a++) {
                double *x = e == 0 ? &F(a, 1) : e == 1 ? &F(a, n) : e == 2 ? &F(1, a) : &F(n, a);
                const double y = e == 0 ? F(a, 0) : e == 1 ? F(a, n + 1) : e == 2 ? F(0, a) : F(n + 1, a);
                *x = 0.5 * (*x + y);
            }
        for (int j = 1; j <= hw; j++) {
            const int jr = g_mutation == 11 ? j - 1 : j; /* 11: the neighbour's row one short */
            for (int a = 1; a <= n; a++)
                zb[a] = e == 0 ? F(a, -jr) : e == 1 ? F(a, n + jr + 1) : e == 2 ? F(-jr, a) : F(n + jr + 1, a);
            for (int a = 2 - j; a <= n + j - 1; a++) {
                ecs_xy_weights(n, hx, a, j, w, &ind);
                const double v = w[0] * zb[ind - 1] + w[1] * zb[ind] + w[2] * zb[ind + 1] + w[3] * zb[ind + 2];
                if (e == 0) F(a, 1 - j) = v;
                else if (e == 1) F(a, n + j) = v;
                else if (e == 2) F(1 - j, a) = v;
                else F(n + j, a) = v;
            }
        }
    }
    free(buf);
#undef F
}

/* h on the lat-lon grid (latlon_grid_generator_mod.f90 init /
 * get_cartesian_coords: lam_i = (i-1) 2 pi / nlon, phi_j = -pi/2 + (j-1) pi /
 * (nlat-1)), as the shallow-water output writes h.dat
 * (swm_output_diag_mod.f90 write_fields L52-73 with h_surf = 0,
 * create_latlon_outputer: cubic): the exchanged, ECS-halo'd xy field
 * (ECS_xy halo, width 2), each point's panel the first of maximal -n.r and
 * its (alpha, beta) by transform_cartesian_to_native_ecs
 * (metric_2d_ecs_mod.f90 L310-336), cubic Lagrange weights at x_ind =
 * int(zdx), zdx = (alpha + pi/4)/hx + 1 (latlon_regrid_mod.f90
 * init_interpolation L298-326, interpolate L330-363), rounded to float as
 * latlon_write stores it.  out[(j-1) nlon + (i-1)]. */
void swo_latlon_h(swo *s, const double *h, int nlat, int nlon, float *out)
{
    const int N = s->N, n = N + 1, hw = 2;
    const double hx = 0.5 * PI / N;
    hmap m = {0, 0};
    build_xy_map(&m, N, hw + 1, 0);
    fld E;
    fld_init(&E, n, n);
    fld_load(&E, h);
    apply_map(&m, &E, &E);
    for (int p = 0; p < 6; p++) ecs_xy_halo_panel(&E, p, n, hx, hw);
    for (int j = 1; j <= nlat; j++) {
        const double phi = -0.5 * PI + (j - 1) * PI / (nlat - 1.0);
        for (int i = 1; i <= nlon; i++) {
            const double lam = (i - 1) * 2.0 * PI / nlon;
            const double r[3] = {cos(phi) * cos(lam), cos(phi) * sin(lam), sin(phi)};
            int pan = 0;
            double rn = 0.0;
            for (int q = 0; q < 6; q++) {
                const double d = -(r[0] * T_N[q][0] + r[1] * T_N[q][1] + r[2] * T_N[q][2]);
                if (d > rn) {
                    pan = q;
                    rn = d;
                }
            }
            const double al = atan((T_EX[pan][0] * r[0] + T_EX[pan][1] * r[1] + T_EX[pan][2] * r[2]) / rn);
            const double be = atan((T_EY[pan][0] * r[0] + T_EY[pan][1] * r[1] + T_EY[pan][2] * r[2]) / rn);
            double zx = (al + 0.25 * PI) / hx + 1.0, zy = (be + 0.25 * PI) / hx + 1.0;
            if (g_mutation == 12) { /* 12: alpha and beta exchanged */
                const double t = zx;
                zx = zy;
                zy = t;
            }
            const int xi = (int)zx, yi = (int)zy;
            zx -= xi;
            zy -= yi;
            const double wx[4] = {-zx * (zx - 1.0) * (zx - 2.0) / 6.0, (zx + 1.0) * (zx - 1.0) * (zx - 2.0) / 2.0,
                                  -(zx + 1.0) * zx * (zx - 2.0) / 2.0, (zx + 1.0) * zx * (zx - 1.0) / 6.0};
            const double wy[4] = {-zy * (zy - 1.0) * (zy - 2.0) / 6.0, (zy + 1.0) * (zy - 1.0) * (zy - 2.0) / 2.0,
                                  -(zy + 1.0) * zy * (zy - 2.0) / 2.0, (zy + 1.0) * zy * (zy - 1.0) / 6.0};
            double v = 0.0;
            for (int b = -1; b <= 2; b++)
                for (int a = -1; a <= 2; a++) v = v + wx[a + 1] * wy[b + 1] * *AT(&E, pan, xi + a, yi + b);
            out[(size_t)(j - 1) * nlon + (i - 1)] = (float)v;
        }
    }
    fld_free(&E);
    free(m.e);
}

/* init_transform_matrix (ecs_halo_vec_xy_factory_mod.f90 L115-187),
 * covariant components: at the exchanged point of edge e (0 bottom, 1 top,
 * 2 left, 3 right), depth j (1: the neighbour's copy of the edge point),
 * along index i, the 2 x 2 map from the neighbour's (frame-aligned)
 * components to the target panel's extended-coordinate components:
 * tm = [b1s.a1t, b2s.a1t; b1s.a2t, b2s.a2t], b the neighbour's contravariant
 * basis rotated into the target's prototype frame, a the target's covariant
 * basis at the same point. */
static void ecs_vec_tm(int e, int i, int j, double hx, double tm[4])
{
    const int bsign = (e == 0 || e == 2) ? 1 : -1, shift = e >= 2;
    const double al = -0.25 * PI + (i - 1.0) * hx, be = bsign * (0.25 * PI - (j - 1.0) * hx);
    const double as = shift ? be : al, bs = shift ? al : be;
    double a1[3], a2[3], b1[3], b2[3];
    ecs_basis(as, bs, a1, a2, b1, b2);
    const double X = tan(as), Y = tan(bs), sg = sqrt(1 + X * X + Y * Y);
    const double r0[3] = {X / sg, Y / sg, 1.0 / sg};
    /* the edge's rotation of the prototype frame: out_k = column k . v */
    double rb1[3], rb2[3], r[3];
    const double *in[3] = {b1, b2, r0};
    double *out[3] = {rb1, rb2, r};
    for (int m = 0; m < 3; m++) {
        const double *v = in[m];
        double *o = out[m];
        if (e == 0) { o[0] = v[0]; o[1] = -v[2]; o[2] = v[1]; }
        else if (e == 1) { o[0] = v[0]; o[1] = v[2]; o[2] = -v[1]; }
        else if (e == 2) { o[0] = -v[2]; o[1] = v[1]; o[2] = v[0]; }
        else { o[0] = v[2]; o[1] = v[1]; o[2] = -v[0]; }
    }
    double A1[3], A2[3], B1[3], B2[3];
    ecs_basis(atan(r[0] / r[2]), atan(r[1] / r[2]), A1, A2, B1, B2);
    tm[0] = rb1[0] * A1[0] + rb1[1] * A1[1] + rb1[2] * A1[2];
    tm[1] = rb2[0] * A1[0] + rb2[1] * A1[1] + rb2[2] * A1[2];
    tm[2] = rb1[0] * A2[0] + rb1[1] * A2[1] + rb1[2] * A2[2];
    tm[3] = rb2[0] * A2[0] + rb2[1] * A2[1] + rb2[2] * A2[2];
}

/* The winds on the lat-lon grid as the shallow-water output writes u.dat and
 * v.dat (swm_output_diag_mod.f90 write_fields L75-76, latlon_vec_write
 * latlon_outputer_mod.f90 L65-107, do_latlon_vector_regrid
 * latlon_regrid_mod.f90 L166-262): the covariant u, v colocated at the xy
 * points by interp2d_uv2pvec_Ch_sbp42 (i2c -> c2i interpolation along x for
 * u, along y for v, the 42 closure whatever the operator's order), the
 * exchanged vector halo turned into the target's extended components
 * (ecs_halo_vec_xy_mod.f90 L60-130), the scalar ECS procedure on each
 * component, cubic interpolation, and the native components turned into
 * zonal / meridional ones by q = [iv.b1, iv.b2; jv.b1, jv.b2]
 * (init_transform_to_latlon_vec_tile, regrid_factory_mod.f90 L338-382; b
 * the panel's contravariant basis, iv / jv the lat-lon unit vectors).  Float
 * outputs as latlon_vec_write stores them. */
void swo_latlon_uv(swo *s, const double *u, const double *v, int nlat, int nlon, float *uo, float *vo)
{
    const int N = s->N, n = N + 1, hw = 2;
    const double hx = 0.5 * PI / N;
    fld EU, EV;
    fld_init(&EU, n, n);
    fld_init(&EV, n, n);
    double *buf = malloc((size_t)(n + 2 * HALO) * sizeof(double));
    for (int p = 0; p < 6; p++) {
        /* (i_c2i reads and writes 1-based lines given by their index-0 address) */
        for (int j = 1; j <= n; j++) /* u rows: N centres -> N + 1 points */
            i_c2i(42, u + ((size_t)p * n + (j - 1)) * N - 1, 1, N, AT(&EU, p, 0, j), 1);
        for (int i = 1; i <= n; i++) { /* v columns */
            for (int j = 1; j <= N; j++) buf[j] = v[((size_t)p * N + (j - 1)) * n + (i - 1)];
            i_c2i(42, buf, 1, N, AT(&EV, p, i, 0), EV.sx);
        }
    }
    free(buf);
    hmap m = {0, 0};
    build_xy_map(&m, N, hw + 1, 1);
    apply_map(&m, &EU, &EV);
    free(m.e);
    for (int p = 0; p < 6; p++) {
        for (int e = 0; e < 4; e++)
            for (int j = 1; j <= hw + 1; j++)
                for (int a = 1; a <= n; a++) {
                    double tm[4];
                    ecs_vec_tm(e, a, j, hx, tm);
                    if (g_mutation == 14) { /* 14: the components left in the neighbour's frame */
                        tm[0] = tm[3] = 1.0;
                        tm[1] = tm[2] = 0.0;
                    }
                    const int ti = e == 0 ? a : e == 1 ? a : e == 2 ? 1 - j : n + j;
                    const int tj = e == 0 ? 1 - j : e == 1 ? n + j : a;
                    double *pu = AT(&EU, p, ti, tj), *pv = AT(&EV, p, ti, tj);
                    const double zu = *pu, zv = *pv;
                    *pu = tm[0] * zu + tm[1] * zv;
                    *pv = tm[2] * zu + tm[3] * zv;
                }
        ecs_xy_halo_panel(&EU, p, n, hx, hw);
        ecs_xy_halo_panel(&EV, p, n, hx, hw);
    }
    for (int j = 1; j <= nlat; j++) {
        const double phi = -0.5 * PI + (j - 1) * PI / (nlat - 1.0);
        for (int i = 1; i <= nlon; i++) {
            const double lam = (i - 1) * 2.0 * PI / nlon;
            const double r[3] = {cos(phi) * cos(lam), cos(phi) * sin(lam), sin(phi)};
            int pan = 0;
            double rn = 0.0;
            for (int q = 0; q < 6; q++) {
                const double d = -(r[0] * T_N[q][0] + r[1] * T_N[q][1] + r[2] * T_N[q][2]);
                if (d > rn) {
                    pan = q;
                    rn = d;
                }
            }
            const double al = atan((T_EX[pan][0] * r[0] + T_EX[pan][1] * r[1] + T_EX[pan][2] * r[2]) / rn);
            const double be = atan((T_EY[pan][0] * r[0] + T_EY[pan][1] * r[1] + T_EY[pan][2] * r[2]) / rn);
            double zx = (al + 0.25 * PI) / hx + 1.0, zy = (be + 0.25 * PI) / hx + 1.0;
            const int xi = (int)zx, yi = (int)zy;
            zx -= xi;
            zy -= yi;
            const double wx[4] = {-zx * (zx - 1.0) * (zx - 2.0) / 6.0, (zx + 1.0) * (zx - 1.0) * (zx - 2.0) / 2.0,
                                  -(zx + 1.0) * zx * (zx - 2.0) / 2.0, (zx + 1.0) * zx * (zx - 1.0) / 6.0};
            const double wy[4] = {-zy * (zy - 1.0) * (zy - 2.0) / 6.0, (zy + 1.0) * (zy - 1.0) * (zy - 2.0) / 2.0,
                                  -(zy + 1.0) * zy * (zy - 2.0) / 2.0, (zy + 1.0) * zy * (zy - 1.0) / 6.0};
            double vu = 0.0, vv = 0.0;
            for (int b = -1; b <= 2; b++)
                for (int a = -1; a <= 2; a++) {
                    vu = vu + wx[a + 1] * wy[b + 1] * *AT(&EU, pan, xi + a, yi + b);
                    vv = vv + wx[a + 1] * wy[b + 1] * *AT(&EV, pan, xi + a, yi + b);
                }
            double a1[3], a2[3], b1[3], b2[3], B1[3], B2[3];
            ecs_basis(al, be, a1, a2, b1, b2);
            proto2real(pan, b1, B1);
            proto2real(pan, b2, B2);
            const double iv[3] = {-sin(lam), cos(lam), 0.0};
            const double jv[3] = {-sin(phi) * cos(lam), -sin(phi) * sin(lam), cos(phi)};
            const double q11 = iv[0] * B1[0] + iv[1] * B1[1] + iv[2] * B1[2];
            const double q12 = iv[0] * B2[0] + iv[1] * B2[1] + iv[2] * B2[2];
            const double q21 = jv[0] * B1[0] + jv[1] * B1[1] + jv[2] * B1[2];
            const double q22 = jv[0] * B2[0] + jv[1] * B2[1] + jv[2] * B2[2];
            const size_t k = (size_t)(j - 1) * nlon + (i - 1);
            uo[k] = (float)(q11 * vu + q12 * vv);
            vo[k] = (float)(q21 * vu + q22 * vv);
        }
    }
    fld_free(&EU);
    fld_free(&EV);
}

/* test_fields_mod.f90 generate_gaussian_hill_scalar_field L491-514 (times amp) */
void swo_gaussian_hill(const swo *s, double lam0, double phi0, double d0,
                       double amp, double *h)
{
    const geom *G = &s->G[SWO_MESH_XY];
    double x0 = cos(phi0) * cos(lam0), y0 = cos(phi0) * sin(lam0), z0 = sin(phi0);
    int n = swo_npts(s, SWO_MESH_XY);
    for (int k = 0; k < n; k++) {
        double c = G->rx[k] * x0 + G->ry[k] * y0 + G->rz[k] * z0;
        c = c > 1 ? 1 : (c < -1 ? -1 : c);
        double r = acos(c);
        h[k] = amp * exp(-(r / d0) * (r / d0));
    }
}

/* covariant component of a Cartesian vector field: V . a1 (mesh u) or
 * V . a2 (mesh v); test_fields_mod.f90 set_vector_test_field L292-329 */
void swo_covariant_from_cart(const swo *s, const double *vx, const double *vy,
                             const double *vz, int mesh, double *out)
{
    int N = s->N;
    mtype t = s->mt[mesh];
    int nx = npts_axis(N, t.xt), ny = npts_axis(N, t.yt);
    double shx = t.xt ? 0.5 : 0.0, shy = t.yt ? 0.5 : 0.0;
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t id = ((size_t)p * ny + (j - 1)) * nx + (i - 1);
                double al = -0.25 * PI + (i - 1 + shx) * s->hx;
                double be = -0.25 * PI + (j - 1 + shy) * s->hx;
                double a1[3], a2[3], b1[3], b2[3], A[3];
                ecs_basis(al, be, a1, a2, b1, b2);
                proto2real(p, mesh == SWO_MESH_U ? a1 : a2, A);
                out[id] = vx[id] * A[0] + vy[id] * A[1] + vz[id] * A[2];
            }
}

/* Williamson test 2 in linear mode: Williamson_test2_mod.f90 L32-112 with
 * test_fields_mod.f90 generate_ts2_height_field L857-894 (metric_term = 0)
 * and gen_solid_rotation_vec_field L586-612 (V = u0 axis x r). */
void swo_ts2_linear(const swo *s, double u0, double h_mean, double *u,
                    double *v, double *h)
{
    double an = sqrt(s->axis[0] * s->axis[0] + s->axis[1] * s->axis[1] + s->axis[2] * s->axis[2]);
    double ax = s->axis[0] / an, ay = s->axis[1] / an, az = s->axis[2] / an;
    const geom *G = &s->G[SWO_MESH_XY];
    int nh = swo_npts(s, SWO_MESH_XY);
    for (int k = 0; k < nh; k++) {
        double r = sqrt(G->rx[k] * G->rx[k] + G->ry[k] * G->ry[k] + G->rz[k] * G->rz[k]);
        double sp = (G->rx[k] * ax + G->ry[k] * ay + G->rz[k] * az) / r;
        double phi = asin(sp);
        h[k] = h_mean - (s->a * s->omega * u0) / s->g * sin(phi) * sin(phi);
    }
    for (int m = SWO_MESH_U; m <= SWO_MESH_V; m++) {
        const geom *Gm = &s->G[m];
        int n = swo_npts(s, m);
        double *vx = malloc(n * sizeof(double)), *vy = malloc(n * sizeof(double)), *vz = malloc(n * sizeof(double));
        for (int k = 0; k < n; k++) {
            double r = sqrt(Gm->rx[k] * Gm->rx[k] + Gm->ry[k] * Gm->ry[k] + Gm->rz[k] * Gm->rz[k]);
            double xn = Gm->rx[k] / r, yn = Gm->ry[k] / r, zn = Gm->rz[k] / r;
            vx[k] = u0 * (ay * zn - az * yn);
            vy[k] = -u0 * (ax * zn - az * xn);
            vz[k] = u0 * (ax * yn - ay * xn);
        }
        swo_covariant_from_cart(s, vx, vy, vz, m, m == SWO_MESH_U ? u : v);
        free(vx); free(vy); free(vz);
    }
}

int swo_halo_map(const swo *s, int kind, int *tp, int *ti, int *tj, int *tc,
                 int *sp, int *si, int *sj, int *sc, int *sign)
{
    const hmap *m = kind == 0 ? &s->m_xy : kind == 1 ? &s->m_ch : kind == 3 ? &s->m_tan : &s->m_xyv;
    if (tp)
        for (int k = 0; k < m->n; k++) {
            const hentry *e = &m->e[k];
            tp[k] = e->tp; ti[k] = e->ti; tj[k] = e->tj; tc[k] = e->tc;
            sp[k] = e->sp; si[k] = e->si; sj[k] = e->sj; sc[k] = e->sc;
            sign[k] = e->sign;
        }
    return m->n;
}

/* 1-D operator on one line for identity tests: op 0 d_i2c, 1 d_c2i,
 * 2 i_i2c, 3 i_c2i.  in points at index 1-HALO of a line padded by HALO on
 * both sides; n = number of centres.  with_sat = 0 drops the SAT terms. */
void swo_line_op(int order, int op, int n, const double *in, double *out, int with_sat)
{
    double *o = calloc(n + 1 + 2 * HALO, sizeof(double));
    const double *f = in + HALO - 1; /* index 0 */
    double *oo = o + HALO - 1;   /* out index 1 at o[HALO] */
    g_no_sat = !with_sat;
    switch (op) {
    case 0: d_i2c(order, f, 1, n, oo, 1); break;
    case 1: d_c2i(order, f, 1, n, oo, 1); break;
    case 2: i_i2c(order, f, 1, n, oo, 1); break;
    default: i_c2i(order, f, 1, n, oo, 1); break;
    }
    g_no_sat = 0;
    int nout = (op == 0 || op == 2) ? n : n + 1;
    for (int k = 1; k <= nout; k++) out[k - 1] = oo[k];
    free(o);
}

/* covariant basis vectors a1, a2 (3-D, unit sphere) at every point of a
 * mesh, packed as [npts][3] */
void swo_basis(const swo *s, int mesh, double *a1out, double *a2out)
{
    int N = s->N;
    mtype t = s->mt[mesh];
    int nx = npts_axis(N, t.xt), ny = npts_axis(N, t.yt);
    double shx = t.xt ? 0.5 : 0.0, shy = t.yt ? 0.5 : 0.0;
    for (int p = 0; p < 6; p++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t id = ((size_t)p * ny + (j - 1)) * nx + (i - 1);
                double al = -0.25 * PI + (i - 1 + shx) * s->hx;
                double be = -0.25 * PI + (j - 1 + shy) * s->hx;
                double a1[3], a2[3], b1[3], b2[3];
                ecs_basis(al, be, a1, a2, b1, b2);
                proto2real(p, a1, a1out + 3 * id);
                proto2real(p, a2, a2out + 3 * id);
            }
}
