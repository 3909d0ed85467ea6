/*
 * CPU oracle for the linear shallow-water RK4 step on the staggered (Ch)
 * equiangular cubed sphere with SBP operators of order 21 and 42.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, scalar restatement of
 * the reference algorithm, used by tests/ (as the parity checker), by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * (ch_swe_amd/, the HIP library) never links or calls it.
 *
 * Reference files restated (paths relative to the reference's src/):
 *   time_schemes/rk4_mod.f90                                   (step_rk4)
 *   models/shallow_water/operator/operator_swm_lin_mod.f90     (apply)
 *   differential_operators/horizontal/gradient/grad_sbp_sat_mod.f90
 *   differential_operators/horizontal/divergence/div_sbp_sat_mod.f90
 *   differential_operators/horizontal/co2contra/co2contra_Ch_mod.f90
 *   differential_operators/horizontal/coriolis/coriolis_Ch_mod.f90
 *   differential_operators/sbp_operators/sbp_{diff,interp}_{i2c,c2i}_{21,42}_mod.f90
 *   equiang_cs/ecs_Ah_vec_sync_mod.f90 (+ _factory_mod.f90, covariant weights)
 *   halo/halo_Ah_scalar_sync_mod.f90
 *   parallel/exchange_factory_mod.f90, buffer_mod.f90, topology/cubed_sphere_topology_mod.f90
 *   metric/metric_2d_ecs_mod.f90, mesh/mesh_factory_mod.f90, domain/domain_factory_mod.f90
 *   quadrature/quadrature_factory_mod.f90, quadrature/sbp_quadrature_mod.f90
 *   test_fields/test_fields_mod.f90 (Gaussian hill, TS2 height, solid rotation)
 *
 * Packed field layout (shared with the product's C-ABI):
 *   u : mesh 'y'  [6][N+1][N]    (panel, j=1..N+1, i=1..N),   i fastest
 *   v : mesh 'x'  [6][N][N+1]    (panel, j=1..N,   i=1..N+1), i fastest
 *   h : mesh 'xy' [6][N+1][N+1]  (panel, j, i),                i fastest
 * This is the reference's tile array p(i,j) (column-major) without halos,
 * one tile per panel, panels 1..6 in order.
 */
#ifndef SWE_ORACLE_H
#define SWE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swo swo;

/* mesh ids */
enum { SWO_MESH_XY = 0, SWO_MESH_U = 1, SWO_MESH_V = 2 };

swo *swo_create(int N, int order, double radius, double grav, double h_mean,
                double omega, const double axis[3], int coriolis_const);
void swo_destroy(swo *s);
int swo_npts(const swo *s, int mesh);

/* full operator and time scheme */
void swo_apply(swo *s, const double *u, const double *v, const double *h,
               double *du, double *dv, double *dh);
void swo_rk4(swo *s, double *u, double *v, double *h, double dt, int nsteps);

/* sub-operators (same packed layouts) */
void swo_grad(swo *s, const double *h, double *gx, double *gy);
void swo_co2contra(swo *s, const double *u, const double *v, double *ut, double *vt);
void swo_coriolis(swo *s, const double *ut, const double *vt, double *cu, double *cv);
void swo_div(swo *s, const double *ut, const double *vt, double *div);

/* geometry, quadrature, test fields */
void swo_metric(const swo *s, int mesh, double *J, double *Qi11, double *Qi12,
                double *Qi22, double *rx, double *ry, double *rz);
void swo_quad_weights(const swo *s, int mesh, double *w);
void swo_coriolis_fJ2(const swo *s, double *fJ2);
void swo_sync_weights(const swo *s, double *qb, double *qt, double *ql, double *qr);
double swo_energy(swo *s, const double *u, const double *v, const double *h);
double swo_mass(const swo *s, const double *h);
/* curl_ch_sbp_sat_XX at the cell centres ((N, N) per panel) */
void swo_curl(swo *s, const double *u, const double *v, double *curl);
/* print_integrals_diag: hmin, hmax, mass, TE, KE, PE, enstrophy */
void swo_integrals_diag(swo *s, const double *u, const double *v, const double *h, double grav, double out[7]);
/* h (packed xy layout) on the nlat x nlon lat-lon grid as the shallow-water
 * output writes h.dat: float, longitude fastest */
void swo_latlon_h(swo *s, const double *h, int nlat, int nlon, float *out);
/* the implicit solve of the linear operator without Coriolis (CG on the
 * Helmholtz problem for h); h must not alias rh */
void swo_solve(swo *s, const double *ru, const double *rv, const double *rh, double dt, double rel_tol, int max_iter,
               double *u, double *v, double *h, int *iters, double *relres);
/* print_CFL_diag: gravity-wave CFL, advective CFL along x and y */
void swo_cfl_diag(swo *s, const double *u, const double *v, const double *h, double dt, double grav, double out[3]);
/* the covariant winds on the lat-lon grid as the output writes u.dat and
 * v.dat: zonal uo, meridional vo (float, longitude fastest) */
void swo_latlon_uv(swo *s, const double *u, const double *v, int nlat, int nlon, float *uo, float *vo);
/* print_integral_tends of the operator's tendency: out = TE_tend,
 * Enstrophy_tend, then the sums of their terms' absolute values */
void swo_integral_tends(swo *s, const double *u, const double *v, const double *h, double grav, double out[4]);
void swo_gaussian_hill(const swo *s, double lam0, double phi0, double d0,
                       double amp, double *h);
void swo_ts2_linear(const swo *s, double u0, double h_mean, double *u,
                    double *v, double *h);
void swo_covariant_from_cart(const swo *s, const double *vx, const double *vy,
                             const double *vz, int mesh, double *out);

/* halo maps: kind 0 = xy scalar depth 1, kind 1 = Ch vector (u x-sides,
 * v y-sides) depth 1..4, kind 2 = xy vector depth 1, kind 3 = tangential
 * components at the duplicated edge points (the curl's SAT values).  Each entry: target (panel, i, j, comp) and source
 * (panel, i, j, comp), sign.  comp: 0 = u-type, 1 = v-type.  Returns count;
 * pass NULL arrays to query the count. */
int swo_halo_map(const swo *s, int kind, int *tp, int *ti, int *tj, int *tc,
                 int *sp, int *si, int *sj, int *sc, int *sign);

/* 1-D line operator (identity tests), basis vectors, planted mistakes */
void swo_line_op(int order, int op, int n, const double *in, double *out, int with_sat);
void swo_basis(const swo *s, int mesh, double *a1, double *a2);
void swo_set_mutation(int m);

/* OpenMP threads used by the row loops (1 when built without OpenMP). */
int swo_num_threads(void);
void swo_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
