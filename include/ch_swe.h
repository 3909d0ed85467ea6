/*
 * ch_swe_amd -- MI355X-native linear shallow-water RK4 step on the staggered
 * (Ch) equiangular cubed sphere with SBP operators of order 21 and 42.
 *
 * C ABI of libchswe.so.  Plain pointers and sizes only.  Every entry point
 * returns 0 on success or a negative SWE_E* code; swe_last_error() gives the
 * message of the last failure on the calling thread.  A plan owns one HIP
 * stream (default priority); calls on one plan are serialised by the caller
 * (one plan per host thread), as the reference runs one thread per MPI rank.
 *
 * Host field layout ("packed"): FP64, one array per field, panels 1..6, each
 * panel stored as the reference's tile array p(i,j) without halos
 * (column-major: i fastest):
 *   u : mesh 'y'  [6][N+1][N]      covariant u at (i=1..N,   j=1..N+1)
 *   v : mesh 'x'  [6][N][N+1]      covariant v at (i=1..N+1, j=1..N)
 *   h : mesh 'xy' [6][N+1][N+1]    height perturbation at vertices
 * (reference: domain/domain_factory_mod.f90 L297-313, partition_mod.f90 L83-97)
 */
#ifndef CH_SWE_H
#define CH_SWE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SWE_OK = 0,
    SWE_EINVAL = -1,   /* bad argument (size, order, null pointer, mismatch) */
    SWE_EHIP = -2,     /* HIP runtime error (no device, launch failure, ...) */
    SWE_ENOMEM = -3,   /* device allocation failed */
    SWE_ENOTFINITE = -4 /* non-finite value detected in a checked result */
};

enum { SWE_FP64 = 0, SWE_FP32 = 1 };

/* swe_config.flags */
enum {
    SWE_FLAG_UNFUSED = 1,      /* run the unfused multi-kernel path (diagnostics) */
    SWE_FLAG_LINEAR_BLOCKS = 2, /* plain block order instead of the XCD-aware one */
    SWE_FLAG_GENERAL_TILES = 4, /* diagnostic: every tile of the fused stage takes the general path */
    SWE_FLAG_COMPENSATED = 8    /* RK4 with compensated (Kahan) summation of each step's update of the
                                   state: states carry a compensation field (twice the device memory) */
};

/* swe_ensemble_rk4 flags */
enum { SWE_ENS_GRAPH = 1 /* replay the step loop from captured HIP graphs */ };

typedef struct swe_plan swe_plan;
typedef struct swe_state swe_state;

/* Operator configuration.  Mirrors the reference's domain/metric config
 * (N, planet_radius, omega, rotation_axis, coriolis_constant:
 * metric/metric_factory_mod.f90 L98-202) and create_linear_swm_operator's
 * arguments (grav, h_mean, SBP order via the operator names:
 * models/shallow_water/operator/operator_swm_factory_mod.f90 L586-685). */
typedef struct {
    int N;              /* cells per panel edge; >= 4 (sbp21), >= 8 (sbp42), >= 12 (sbp63) */
    int order;          /* 21, 42 or 63; 63 runs on the multi-kernel path (as with
                           SWE_FLAG_UNFUSED): ensembles and the server take sbp21 /
                           sbp42 plans */
    int precision;      /* SWE_FP64 or SWE_FP32 (stencils/state); metric is FP64 */
    int coriolis_const; /* 1: f = omega everywhere; 0: f = 2 omega (axis . r) */
    double radius;      /* planet radius [m] */
    double grav;        /* gravity [m/s^2] */
    double h_mean;      /* mean depth H0 [m] */
    double omega;       /* rotation rate [1/s] */
    double axis[3];     /* rotation axis */
    int flags;          /* SWE_FLAG_* bits */
    int block_threads;  /* workgroup size of the fused kernels: 0 or 512 (both mean 512) */
    int tile_x0;        /* width of the first tile along x / y of the fused stage kernel's tiling: */
    int tile_y0;        /*   0 = the library's split; else 8 .. (what the axis leaves) - 9, within the
                             tile limits (ch_swe_amd/autotune.py sweeps these) */
} swe_config;

const char *swe_last_error(void);
const char *swe_version(void);

/* Plan: geometry (FP64 metric built on the device), edge maps, work fields,
 * stream.  Replaces create_domain + create_linear_swm_operator +
 * create_timescheme("rk4") (models/shallow_water/shallow_water_mod.f90
 * L70-93).  The RK4 registers are allocated by the first swe_rk4_steps call. */
int swe_plan_create(const swe_config *cfg, swe_plan **out);
int swe_plan_destroy(swe_plan *plan);
int swe_plan_npts(const swe_plan *plan, int *n_u, int *n_v, int *n_h);
void *swe_plan_stream(swe_plan *plan); /* the plan's hipStream_t */
int swe_synchronize(swe_plan *plan);

/* State vector (u, v, h) resident on the device: stvec_swm_t
 * (models/shallow_water/stvec/stvec_swm_mod.f90 L10-42) and
 * stvec_t%create_similar (stvec/stvec_mod.f90 L10-30). */
int swe_state_create(swe_plan *plan, swe_state **out);
int swe_state_destroy(swe_state *st);
int swe_state_upload(swe_state *st, const double *u, const double *v, const double *h);
int swe_state_download(swe_state *st, double *u, double *v, double *h);
/* The same exchange with the reference's own tile arrays, halos included:
 * grid_field_t%tile(t)%p(is-hw:ie+hw, js-hw:je+hw, 1:1)
 * (grid_field_factory_mod.f90 create_grid_field L9-31, grid_field_mod.f90
 * tile_field_init L78-95) for the
 * partition of one tile per panel, the six tiles back to back (u: (N+2hw) x
 * (N+1+2hw) per panel, v: (N+1+2hw) x (N+2hw), h: (N+1+2hw)^2, column-major).
 * Upload reads only the interior; download writes only the interior and
 * leaves the halo entries as they were (the reference fills halos by
 * exchange when an operator needs them).  0 <= hw <= 64. */
int swe_state_upload_halo(swe_state *st, int hw, const double *u, const double *v, const double *h);
int swe_state_download_halo(swe_state *st, int hw, double *u, double *v, double *h);

/* operator_t%apply (src/operator_mod.f90 L11-41) for operator_swm_lin_t
 * (operator_swm_lin_mod.f90 apply L66-108): vout = A(vin). */
int swe_apply(swe_plan *plan, swe_state *vout, swe_state *vin);

/* stvec_t%assign / %update (stvec/stvec_mod.f90 L10-30; interior points,
 * grid_field_mod.f90 L460-534):
 *   assign: out = s1*v1 + s2*v2      (v2 may be NULL: out = s1*v1)
 *   update: out = out + s1*v1 + s2*v2 (v2 may be NULL) */
int swe_assign(swe_plan *plan, swe_state *out, double s1, swe_state *v1, double s2, swe_state *v2);
int swe_update(swe_plan *plan, swe_state *out, double s1, swe_state *v1, double s2, swe_state *v2);

/* rk4_t%step (time_schemes/rk4_mod.f90 L20-42) applied nsteps times to v0 in
 * place.  The plan owns the k1..k4, y registers (timescheme_factory_mod.f90
 * L90-107). */
int swe_rk4_steps(swe_plan *plan, swe_state *v0, double dt, int nsteps);

/* FP64 SBP-quadrature integrals on the device (quadrature/sbp_quadrature_mod.f90
 * L22-62, abstract_quadrature_mod.f90 L29-66): mass = sum h J hx hy Ax Ay;
 * energy = g/2 <h,h> + H0/2 (<u,ut> + <v,vt>). */
int swe_integrals(swe_plan *plan, swe_state *st, double *mass, double *energy);

/* swe_integrals without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same kernels and results as swe_integrals, with the partial sums in a
 * static device array of the library (24.6 KB per device, any N) instead of
 * the plan's buffers.  Calls that overlap in time on one device share that
 * array and race: serialise them (one host thread, or one stream). */
int swe_integrals_noalloc(swe_plan *plan, swe_state *st, double *mass, double *energy);

/* The integral diagnostics the reference prints every diagnostics step:
 * swm_output_diag_t%print_integrals_diag + calc_energies
 * (models/shallow_water/output_and_diag/swm_output_diag_mod.f90 L156-234)
 * with the Ch operators of the paper's output_diag section
 * (interp2d_p2uv_Ch_sbpXX, interp2d_uv2qvec_Ch_sbpXX, curl_ch_sbp_sat_XX,
 * co2contra_ch_sbpXX, SBP_CXX quadratures), covariant winds, no orography.
 * out[7] = hmin, hmax, mass, TE, KE, PE, enstrophy (FP64 sums on the device);
 * grav is the PE's gravity (the reference uses Earth_grav there whatever the
 * operator's).  Fused sbp21 / sbp42 plans sum tile by tile in one launch,
 * others (SWE_FLAG_UNFUSED, sbp63) point by point on the multi-kernel path;
 * blocks until the values are on the host. */
int swe_integrals_diag(swe_plan *plan, swe_state *st, double grav, double out[7]);

/* operator_swm_lin_t%solve
 * (models/shallow_water/operator/operator_swm_lin_mod.f90 L110-146): vout
 * = (I - dt L0)^-1 rhs for the plan's linear operator without its Coriolis
 * term (L0), as the reference solves it: the Helmholtz problem for h,
 * (I - H g dt^2 div co2contra grad) h = rhs_h - H dt div co2contra(rhs_u,
 * rhs_v), by conjugate gradients (solver/cg_solver_mod.f90 L19-90: initial
 * guess the right-hand side, relative residual in the h quadrature's l2
 * norm below rel_tol, at most max_iter iterations; the helm_solver_cfg
 * keys rel_tol and max_iter), then u = rhs_u - g dt grad_x h, v = rhs_v -
 * g dt grad_y h.  *iters (may be NULL): iterations done, 0 when the initial
 * guess meets rel_tol, max_iter + 1 when the tolerance was never met (the
 * result is then the last iterate, as in the reference, which only prints
 * a message); *relres (may be NULL): the final relative residual.  vout and
 * rhs must differ.  Device scratch (three states, four h fields) is
 * allocated on the first call; blocks once per iteration for the
 * convergence test. */
int swe_solve(swe_plan *plan, swe_state *vout, swe_state *rhs, double dt, int max_iter, double rel_tol, int *iters,
              double *relres);

/* timescheme_t%step of implicit_Eul (time_schemes/implicit_Eul_mod.f90
 * L20-30: v0 = solve(v0, dt)) and crank_nikolson (crank_nikolson_mod.f90
 * L20-32: v0 = solve(v0 + dt/2 apply(v0), dt/2)), nsteps times, with the
 * solver settings of swe_solve; *iters / *relres (may be NULL): the
 * largest over the steps. */
int swe_implicit_euler_steps(swe_plan *plan, swe_state *v0, double dt, int nsteps, int max_iter, double rel_tol,
                             int *iters, double *relres);
int swe_crank_nicolson_steps(swe_plan *plan, swe_state *v0, double dt, int nsteps, int max_iter, double rel_tol,
                             int *iters, double *relres);

/* swm_output_diag_t%print_CFL_diag
 * (models/shallow_water/output_and_diag/swm_output_diag_mod.f90 L110-154),
 * covariant winds: out[3] = the gravity-wave CFL sqrt(max g h / dx^2) dt
 * and the advective CFLs sqrt(max |ut u| / dx^2) dt, sqrt(max |vt v| /
 * dy^2) dt (ut, vt contravariant; dx = dy = radius * pi / (2N)); grav as the
 * reference passes it there (Earth_grav).  Blocks until the values are on
 * the host. */
int swe_cfl_diag(swe_plan *plan, swe_state *st, double dt, double grav, double out[3]);

/* The lat-lon output of h: swm_output_diag_t%write_fields's h.dat
 * (models/shallow_water/output_and_diag/swm_output_diag_mod.f90 L52-73,
 * h_surf = 0) through latlon_outputer_t%write
 * (outputer/latlon_outputer_mod.f90 L33-63) with the cubic lat-lon regrid
 * of the xy points (regridders/latlon_regrid_mod.f90, ECS_xy halo of width
 * 2, equiang_cs/ecs_halo_xy_mod.f90).  out gets nlat * nlon floats,
 * longitude fastest: latitude -90 + (j-1) 180 / (nlat-1) degrees, longitude
 * (i-1) 360 / nlon degrees -- one record of h.dat (the reference writes
 * nlat = 2N+1, nlon = 4N).  The plan's tables for the grid are built on the
 * first call for it (device memory grows by about 60 bytes per lat-lon
 * point); blocks until out is filled. */
int swe_latlon_h(swe_plan *plan, swe_state *st, int nlat, int nlon, float *out);

/* The lat-lon output of the winds: write_fields's u.dat / v.dat
 * (swm_output_diag_mod.f90 L75-76) through latlon_vec_outputer_t%write
 * (latlon_outputer_mod.f90 L65-107) and do_latlon_vector_regrid
 * (regridders/latlon_regrid_mod.f90 L166-262): the covariant u, v colocated
 * at the xy points (interp2d_uv2pvec_Ch_sbp42), the vector ECS halo
 * (equiang_cs/ecs_halo_vec_xy_mod.f90, covariant), cubic interpolation and
 * the turn into zonal (u_out) and meridional (v_out) components, each nlat *
 * nlon floats laid out as swe_latlon_h's.  Tables are built on the first
 * call for a grid (about 150 bytes per lat-lon point); blocks until both
 * outputs are filled. */
int swe_latlon_uv(swe_plan *plan, swe_state *st, int nlat, int nlon, float *u_out, float *v_out);

/* swm_output_diag_t%print_integral_tends
 * (models/shallow_water/output_and_diag/swm_output_diag_mod.f90 L236-342) of
 * the plan's operator: the tendency apply(st) goes to a plan-owned buffer
 * (allocated on the first call), then out[2] = TE_tend, Enstrophy_tend with
 * the same operators as swe_integrals_diag; hq is the first output of
 * interp2d_uv2qvec (i2c along y of h at the u points), as there.  Blocks
 * until the values are on the host. */
int swe_integral_tends(swe_plan *plan, swe_state *st, double grav, double out[2]);

/* Serving: continuous batching of independent solves of mixed size and
 * precision under a device-memory budget (no counterpart in the reference,
 * which runs one solve per program).  Requests are queued with a copy of
 * their initial state.  Each precision has a lane with at most one batch of
 * RK4 steps in flight (all its requests together, one launch per stage and
 * SBP order; the two lanes on their own streams, so their work overlaps).
 * Every request computes exactly what swe_rk4_steps computes for it alone.
 * Single-threaded use. */
typedef struct swe_server swe_server;
int swe_server_create(size_t budget_bytes, swe_server **out);
int swe_server_destroy(swe_server *srv);
/* cfg is checked as swe_plan_create checks it (fused path only); u, v, h
 * packed FP64 of the sizes of cfg->N (copied). */
int swe_server_submit(swe_server *srv, const swe_config *cfg, const double *u, const double *v,
                      const double *h, double dt, int nsteps, long long *id);
/* One scheduling round; never waits for the device.  Retires the requests
 * whose batch has finished, admits queued requests in arrival order while
 * their device memory fits the budget (releasing idle cached plans first),
 * and starts the next batch on each idle lane: the lane's largest request
 * advances chunk_steps steps, a request with fewer cells proportionally more.
 * *completed = requests retired in this call.  A request that cannot be
 * admitted is retired with the error instead of blocking the queue. */
int swe_server_tick(swe_server *srv, int chunk_steps, int *completed);
/* Block until the batches in flight have finished (the next tick retires). */
int swe_server_wait(swe_server *srv);
/* *ready = 0 while request id runs.  Once it has finished: *ready = 1, the
 * request is forgotten, and the return value is its status: SWE_OK with the
 * final (u, v, h) copied out; SWE_ENOTFINITE with the (non-finite) result
 * copied out; or the error that kept it from being admitted (nothing copied).
 * swe_last_error() names the request. */
int swe_server_take(swe_server *srv, long long id, double *u, double *v, double *h, int *ready);
int swe_server_stats(swe_server *srv, int *queued, int *active, size_t *bytes_in_use);
/* Totals since creation: requests admitted, cached plans released to make
 * room, requests retired with an error status. */
int swe_server_counters(swe_server *srv, long long *admitted, long long *evicted_plans, long long *failed);

/* Measurement aid (no counterpart in the reference): with timing enabled,
 * every launch of the fused stage kernel and of the edge-record kernel on
 * this plan carries start / stop HIP events timestamped by the dispatch
 * itself (the kernels' own run times).
 * swe_plan_kernel_times waits for them, returns the summed durations (ms)
 * and counts since the last call, and starts a new interval.  Enabling or
 * disabling drops what was recorded. */
int swe_plan_timing(swe_plan *plan, int enable);
int swe_plan_kernel_times(swe_plan *plan, double *stage_ms, int *n_stage, double *records_ms, int *n_records);

/* Host-only (no device needed): halo map of the in-device edge gathers, for
 * inspection and tests.  kind 0: xy scalar depth 1; 1: Ch vector (u across
 * left/right, v across bottom/top) depth 2, 3 or 4 (sbp21, 42, 63); 2: xy
 * vector depth 1; 3: the tangential components at the duplicated edge points
 * (v across left/right, u across bottom/top) depth 1, read by the curl.  Same entry convention as the exchange builders of
 * parallel/exchange_factory_mod.f90 L166-672: target (panel 0..5, i, j,
 * component) <- sign * source (panel, i, j, component).  Pass NULL arrays to
 * query the count. */
int swe_halo_map(int N, int order, int kind, int *count, int *tp, int *ti, int *tj, int *tc,
                 int *sp, int *si, int *sj, int *sc, int *sign);

/* Host-only self-check of the fused kernels' plan-time tables for a panel
 * grid of N and a precision (no counterpart in the reference; runs without a
 * GPU): the packed cube side table against the topology, the arithmetic
 * edge-segment decode against the segment table, and the tile launch order
 * (a bijection, side tiles first).  0 or SWE_EINVAL with the reason. */
int swe_check_tables(int N, int precision);

/* Ensembles and uneven batches: many independent solves advanced together,
 * every RK4 stage of all members in one launch of the fused stage kernel
 * (a work list of (member, tile) items).  No counterpart in the reference,
 * which runs one solve per program; each member's arithmetic is exactly that
 * of swe_rk4_steps on its own plan, so results do not depend on the mix.
 *   members : distinct states; their plans may differ in N and SBP order but
 *             share one precision and use the fused path; the plans must
 *             outlive the ensemble.  Members may share a plan.
 * The ensemble owns each member's RK4 registers and edge records and one
 * stream (default priority). */
typedef struct swe_ensemble swe_ensemble;
int swe_ensemble_create(int n, swe_state *const *members, swe_ensemble **out);
int swe_ensemble_destroy(swe_ensemble *ens);
/* Advance member i by nsteps[i] >= 0 RK4 steps of size dt[i].  Asynchronous
 * on the ensemble's stream, ordered after earlier work on the members' plan
 * streams and before later work on them (stream events), so
 * swe_state_download / swe_integrals afterwards see the result.  flags:
 * SWE_ENS_GRAPH. */
int swe_ensemble_rk4(swe_ensemble *ens, const double *dt, const int *nsteps, int flags);
void *swe_ensemble_stream(swe_ensemble *ens);
int swe_ensemble_synchronize(swe_ensemble *ens);

#ifdef __cplusplus
}
#endif
#endif
