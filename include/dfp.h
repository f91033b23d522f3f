/*
 * dfp -- MI355X-native pressure projection for dopamine-fdm (C ABI).
 *
 * Replaces the Poisson/projection path of the reference's Fortran module
 * `projection` (src/projection.f90) and its GPU solver (src/poisson_gpu.f90)
 * for z periodic, one rank:
 *   the channel, x periodic (x_bc_type = 0, y_bc_type = 1, z_bc_type = 0);
 *   the inflow/outflow x (x_bc_type = 1: DCT-IV in x, no mean-pressure pin,
 *   optional outlet backflow term, projection.f90:65-96, :126-129, :333-339);
 *   the periodic box (x_bc_type = 0, y_bc_type = 0 on a uniform y grid: a y
 *   FFT over the nym-1 independent rows in place of the wall solve, the last
 *   row a duplicate of the first, projection.f90:246-273, :350-354; FP64).
 *
 * All arrays are the reference's Fortran arrays, passed by address
 * (column-major, x fastest), with these extents (initialization.f90:204-220):
 *     U(nx,nyg,nzg)   V(nxg,ny,nzg)   W(nxg,nyg,nz)   P(nxg,nyg,nzg)
 *     rhs_p(2:nxg, 2:nyg-1, 2:nzg)   ->  (nxg-1) x (nyg-2) x (nzg-1)
 * where nxp, nym, nzp are the Fourier/cell counts and
 *     nx = nxp+2, ny = nym+1, nz = nzp+2, nxg = nxp+3, nyg = nym+2, nzg = nzp+3
 * (x periodic: the last cell duplicates the first), or nx = nxp+1, nxg = nxp+2
 * (x inflow/outflow: every cell is transformed, initialization.f90:331-338).
 *
 * Calls are synchronous unless DFP_ASYNC is given.  A plan is not
 * thread-safe: use one plan per host thread.  Every function returns 0 on
 * success or a negative DFP_E* code; dfp_strerror() names it, and
 * dfp_last_error() gives the plan's last detailed message.
 */
#ifndef DFP_H
#define DFP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfp_plan dfp_plan;

/* Grid of one case.  Mirrors the `global` module data the reference's
 * projection reads (global.f90:130-178, initialization.f90:243-345). */
typedef struct dfp_grid {
    int32_t nxp, nym, nzp;  /* Fourier/cell grid: namelist nx-2, ny-1, nz-2       */
    double dx, dz;          /* FFT spacings: dxmin, dzmin (initialization.f90:326) */
    const double *x;        /* x faces  x(1:nx)   (max-divergence only; may be NULL: */
                            /*          then x(i)-x(i-1) is taken as dx)            */
    const double *y;        /* y faces  y(1:ny)                                     */
    const double *z;        /* z faces  z(1:nz)                                     */
    const double *yg;       /* y centres with ghosts yg(1:nyg)                      */
    const double *zg;       /* z centres with ghosts zg(1:nzg)                      */
    int32_t x_bc_type, y_bc_type, z_bc_type; /* x 0 or 1; y 1 (or 0 with x 0); z 0 */
} dfp_grid;

/* Precision of the spectral solve.
 * DFP_FP64: the reference's algorithm in FP64 (projection.f90:100-382).
 * DFP_FP32_REFINE: FP32 FFTs on an FP32-stored spectrum (the y elimination
 *   runs in FP64 on the stored values), the kx=kz=0 mode, the mean-pressure
 *   pin, the residual b - L_pin phi and the accumulation of phi in FP64, then
 *   dfp_set_refine_sweeps() refinement sweeps (default 1).  No counterpart in
 *   the reference, whose solve is FP64 only. */
enum { DFP_FP64 = 0, DFP_FP32_REFINE = 1 };

/* flags */
enum {
    DFP_FIELDS_ON_DEVICE = 1, /* array arguments are device pointers (resident mode)  */
    DFP_SAVE_P = 2,           /* also write P = phi/(dt*c) with its ghosts (rk_step 3) */
    DFP_ASYNC = 4,            /* do not synchronise the stream before returning       */
    DFP_SOLVE_ONLY = 8,       /* dfp_batch_project: solve rhs_p -> phi as dfp_solve   */
    DFP_CHECK_FINITE = 16     /* dfp_project: refuse u* holding NaN or |u| > 1e8 in the
                                 monitor's ranges (monitor.f90:32-38), DFP_E_NONFINITE,
                                 before anything is written                            */
};

/* error codes */
enum {
    DFP_OK = 0,
    DFP_E_ARG = -1,         /* bad argument (null pointer, wrong size, bad flag)      */
    DFP_E_UNSUPPORTED = -2, /* boundary conditions or FFT lengths outside this build  */
    DFP_E_HIP = -3,         /* HIP runtime error                                      */
    DFP_E_NOMEM = -4,       /* device allocation failed                               */
    DFP_E_NONFINITE = -5    /* NaN or Inf found where a finite value was required     */
};

/* Build a plan: operators (kxx, kzz, Dyy as in initialization.f90:416-539;
 * with x_bc_type = 1 the DCT-IV kxx of :443-445), FFT twiddles and device
 * scratch for one grid on one device.
 * Replaces the reference's set-up in initialization.f90:322-539 and
 * gpu_poisson_init / gpu_gtsv_init (poisson_gpu.f90:60-135, :388-428).
 * DFP_E_UNSUPPORTED for other boundary conditions, odd nxp, nxp/2 or nzp
 * with a prime factor other than 2 and 3, nxp or nzp above 1024, nym
 * above 65535, and DFP_FP32_REFINE with x_bc_type = 1.  With one cell
 * between the walls (nym = 1) Dyy(2,2) is a+b, the last of the reference's
 * two assignments to it (initialization.f90:523-535). */
int dfp_plan_create(const dfp_grid *g, int precision, int device, dfp_plan **out);
int dfp_plan_destroy(dfp_plan *plan);

/* Launch shapes of a plan (no counterpart in the reference, whose FFT and
 * tridiagonal libraries choose their own).  0 = the library's choice.
 *   z_split    radix R in {1,2,3,4,6,8,9,12,16} of the first z-FFT stage,
 *              applied by the x kernels (R must divide nzp; the z kernels
 *              then transform rows of nzp/R <= 256 points, one mode per lane)
 *   x_columns  z columns per x-kernel workgroup (divides nzp/R); given, it
 *              sets both x kernels, else each picks its own (dfp_plan_tuning
 *              reports the forward kernel's)
 *   z_rows     y rows per group of the split z kernels (z_fused = 1)
 *   z_fused    1: split pair of z kernels (FFT + elimination, back
 *              substitution + inverse FFT, spectrum read and written twice);
 *              2: one fused kernel holding a whole (kx, z-split) sub-plane
 *              of nym x nzp/R modes in LDS (read and written once)
 * Results do not depend on x_columns or z_rows; z_split changes the FFT's
 * factorisation and z_fused = 2 the association of the y solve (chunked
 * across lanes), both at round-off level.  DFP_E_ARG for values out of
 * range, DFP_E_UNSUPPORTED for shapes that do not fit the device's LDS. */
typedef struct dfp_tuning {
    int32_t z_split, x_columns, z_rows, z_fused;
} dfp_tuning;
int dfp_plan_create_tuned(const dfp_grid *g, int precision, int device, const dfp_tuning *tune,
                          dfp_plan **out);
int dfp_plan_tuning(const dfp_plan *plan, dfp_tuning *out);

/* Whole projection: divergence -> Poisson solve -> velocity correction.
 * Replaces compute_projection_step (projection.f90:21-34).  On return U, V,
 * W hold the corrected velocity (projection.f90:551-604), rhs_p holds the
 * pseudo-pressure phi with the periodic duplicates filled
 * (projection.f90:441-500) and, with DFP_SAVE_P, P = phi/(dt*c) with its
 * ghosts (projection.f90:361-380, :388-438). */
int dfp_project(dfp_plan *plan, double *U, double *V, double *W, double *rhs_p, double *P,
                double dt, double c, int flags);

/* Outlet backflow stabilisation of the x inflow/outflow case: the module
 * state set_outflow_backflow_pressure reads and writes (global.f90:188-197).
 * beta = outflow_backflow_beta (>= 0; 0 turns the term off), outflow_Uc =
 * the plane-mean outlet speed, Pb(2:nyg-1,2:nzg-1) the outlet face pressure
 * (read) and phib(2:nyg-1,2:nzg-1) the face value of the solve (written by
 * dfp_project and dfp_outflow_backflow, read by dfp_project and
 * dfp_correct), both (nyg-2) x (nzg-2) arrays.  The plan keeps the pointers
 * and reads them at each call: device pointers for calls with
 * DFP_FIELDS_ON_DEVICE, host pointers otherwise.  With beta > 0
 * dfp_project runs compute_projection_step with the term
 * (projection.f90:21-34, :65-96): c is then rk_coef(rk_step,rk_step), which
 * also scales P at the last stage.  DFP_E_ARG for beta > 0 on a periodic x
 * (input_output.f90:312-314) or without Pb and phib.  Call again whenever
 * outflow_Uc changes (boundary_conditions.f90:208). */
int dfp_set_outflow(dfp_plan *plan, double beta, double outflow_Uc, const double *Pb, double *phib);

/* set_outflow_backflow_pressure (projection.f90:65-96) on its own: phib from
 * U, V, W at the outlet and Pb, and rhs_p(nx,:,:) -= 2 phib/dx^2, with
 * c = rk_coef(rk_step,rk_step).  Needs dfp_set_outflow with beta > 0. */
int dfp_outflow_backflow(dfp_plan *plan, const double *U, const double *V, const double *W, double *rhs_p,
                         double dt, double c, int flags);

/* rhs_p = div(U,V,W) on i=2..nxg-1, j=2..nyg-1, k=2..nzg-1.
 * Replaces compute_pseudo_pressure_rhs (projection.f90:37-63). */
int dfp_divergence(dfp_plan *plan, const double *U, const double *V, const double *W,
                   double *rhs_p, int flags);

/* Solve the pressure Poisson equation in place on rhs_p.
 * Replaces solve_poisson_equation (projection.f90:100-382); P may be NULL,
 * and is written only when DFP_SAVE_P is given (the reference's
 * skip_p_save=.True. corresponds to leaving DFP_SAVE_P out). */
int dfp_solve(dfp_plan *plan, double *rhs_p, double *P, double dt, double c, int flags);

/* U,V,W -= grad(rhs_p).  Replaces project_velocity (projection.f90:551-604);
 * with x_bc_type = 1 also the outlet face U(nx) += 2 (rhs_p(nx) - phib)/dx,
 * phib = 0 without the backflow term (:578-589). */
int dfp_correct(dfp_plan *plan, double *U, double *V, double *W, const double *rhs_p, int flags);

/* max |div u| over the interior cells, with the face spacings of
 * check_divergence (projection.f90:607-639, no immersed boundary). */
int dfp_max_divergence(dfp_plan *plan, const double *U, const double *V, const double *W,
                       double *out, int flags);

/* No counterpart in the reference: dfp_max_divergence for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Device-resident fields only
 * (flags must be DFP_FIELDS_ON_DEVICE); the same value as
 * dfp_max_divergence, bit for bit.  Its reduction partials and result live in
 * a fixed array of library device memory (at most 1024 partials + the
 * result, 8200 bytes per GPU, held until the process ends) instead of the
 * plan's workspace, so the call allocates nothing; the result is copied to
 * the host double *out and the plan's stream is synchronised.  All calls on
 * one GPU share that array: calls that overlap on one GPU (several host
 * threads, or several streams) race. */
int dfp_max_divergence_noalloc(dfp_plan *plan, const double *U, const double *V, const double *W,
                               double *out, int flags);

/* The step either side of the projection, for a plan with a periodic x
 * (x_bc_type = 0: the channel, the periodic box or the duct) and device-resident fields (flags must hold DFP_FIELDS_ON_DEVICE;
 * DFP_ASYNC allowed).  Fu, Fv, Fw are host arrays of `stage` device pointers
 * to the right-hand sides Fu_s(2:nx-1,2:nyg-1,2:nzg-1),
 * Fv_s(2:nxg-1,2:ny-1,2:nzg-1), Fw_s(2:nxg-1,2:nyg-1,2:nz-1)
 * (initialization.f90:662-670); coef = rk_coef(stage, 1:stage).
 *
 * dfp_rk_update: the RK3 stage update of the interiors of U, V, W from
 * Uo, Vo, Wo (time_integration.f90:256-258, :394-396, :526-531).
 * dfp_apply_bc: apply_boundary_conditions of the channel on one rank
 * (boundary_conditions.f90:186-260): periodic x and z ghost planes and the
 * Robin (slip-length) wall rows with alpha_x(nx,2,nzg), alpha_y(nxg,2,nzg),
 * alpha_z(nxg,2,nz) (global.f90:328; 0 = no slip), or, for the periodic box,
 * the periodic y ghost rows (:229-233, :762-781; the alphas may be NULL);
 * the duct's z walls take W = 0 on the wall faces and mirrored U, V ghosts
 * (:215-223, :737-758) in place of the z wraps.
 * dfp_substep: one RK3 substage around the projection as time_integration.f90
 * runs it without immersed boundaries, VOF or scalars (:256-304): update,
 * boundary conditions, dfp_project with c = rk_coef(stage,stage) (P saved
 * with DFP_SAVE_P), boundary conditions, all on the plan's stream.
 * DFP_E_UNSUPPORTED for an inflow/outflow plan, whose x boundary conditions
 * (inflow data, convective outflow) are not built here. */
int dfp_rk_update(dfp_plan *plan, int stage, double *U, double *V, double *W, const double *Uo, const double *Vo,
                  const double *Wo, const double *const *Fu, const double *const *Fv, const double *const *Fw,
                  const double *coef, double dt, int flags);
int dfp_apply_bc(dfp_plan *plan, double *U, double *V, double *W, const double *alpha_x, const double *alpha_y,
                 const double *alpha_z, int flags);
int dfp_substep(dfp_plan *plan, int stage, double *U, double *V, double *W, const double *Uo, const double *Vo,
                const double *Wo, const double *const *Fu, const double *const *Fv, const double *const *Fw,
                const double *coef, const double *alpha_x, const double *alpha_y, const double *alpha_z,
                double *rhs_p, double *P, double dt, int flags);

/* Refinement sweeps of a DFP_FP32_REFINE plan, 0..8 (no effect on an FP64
 * plan).  No counterpart in the reference. */
int dfp_set_refine_sweeps(dfp_plan *plan, int sweeps);

/* Launch stream for the plan's kernels (a hipStream_t; NULL = default stream). */
int dfp_set_stream(dfp_plan *plan, void *hip_stream);

/* Stage timing, the analogue of the reference's profiler tags
 * (profiler.f90:14-16).  While enabled, every kernel launch is bracketed by
 * HIP events on the plan's stream; dfp_profile_read returns, per kernel
 * name, the number of launches and the summed milliseconds. */
int dfp_profile_enable(dfp_plan *plan, int enable);
int dfp_profile_read(dfp_plan *plan, int index, const char **name, int64_t *launches, double *ms);
int dfp_profile_reset(dfp_plan *plan);

/* Sizes of the arrays a plan expects, in doubles: U, V, W, P, rhs_p. */
int dfp_field_sizes(const dfp_plan *plan, int64_t sizes[5]);

/* Synthetic channel fields for tests and benchmarks (no counterpart in the
 * reference, whose initial conditions use unseeded noise): U = 1-(y/h)^2 +
 * sigma*N(0,1), V, W = sigma*N(0,1), V = 0 on the walls, periodic x/z ghosts
 * and no-slip y ghosts as boundary_conditions.f90 leaves them.  Normal
 * deviates: SplitMix64 counter hash -> Box-Muller.  Device pointers only. */
int dfp_synth_fields(dfp_plan *plan, double *U, double *V, double *W, uint64_t seed, double sigma);
/* the same on the plan's stream; flags may hold DFP_ASYNC */
int dfp_synth_fields_ex(dfp_plan *plan, double *U, double *V, double *W, uint64_t seed, double sigma, int flags);

/* Batched projections (no counterpart in the reference, which projects one
 * grid per run): a batch executor runs the projections of many plans - any
 * grid sizes, either precision - with one persistent launch per stage on one
 * stream.  Fields are device pointers with each plan's extents; a plan may
 * appear once per call (it owns its spectral scratch).  flags[i] may hold
 * DFP_SAVE_P (and DFP_FIELDS_ON_DEVICE, implied); batch_flags DFP_ASYNC and
 * DFP_SOLVE_ONLY.  Results equal those of dfp_project on each plan, bit for
 * bit; with DFP_SOLVE_ONLY, those of dfp_solve (rhs_p holds the right-hand
 * side and receives phi, U/V/W are not read and may be null): the repeated
 * solve_poisson_equation(skip_p_save=.True.) of a caller such as the
 * reference's VOF preconditioner (vof_pressure.f90:707-731), for many cases
 * in one call. */
typedef struct dfp_batch dfp_batch;
int dfp_batch_create(int device, int max_cases, dfp_batch **out);
int dfp_batch_destroy(dfp_batch *batch);
int dfp_batch_set_stream(dfp_batch *batch, void *hip_stream);
/* workgroups per CU of the persistent x, z and stencil launches (autotuned) */
int dfp_batch_set_occupancy(dfp_batch *batch, int x_per_cu, int z_per_cu, int stencil_per_cu);
/* replay the launch sequence from hipGraphs keyed by its stage list */
int dfp_batch_use_graphs(dfp_batch *batch, int enable);
/* independent launches of one stage (different z splits or z kernels) run
 * side by side on up to 1 + aux_streams streams (0..3, default 3); callers
 * that already overlap several batches on their own streams use 0 */
int dfp_batch_set_parallel(dfp_batch *batch, int aux_streams);
/* plans with x_bc_type = 1 are refused (DFP_E_UNSUPPORTED) */
int dfp_batch_project(dfp_batch *batch, int n, dfp_plan *const *plans, double *const *U, double *const *V,
                      double *const *W, double *const *rhs_p, double *const *P, const double *dt,
                      const double *c, const int *flags, int batch_flags);
const char *dfp_batch_last_error(const dfp_batch *batch);

const char *dfp_strerror(int code);
const char *dfp_last_error(const dfp_plan *plan);

#ifdef __cplusplus
}
#endif
#endif /* DFP_H */
