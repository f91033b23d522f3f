/*
 * C restatement of dopamine-fdm's channel pressure projection (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the product never links or calls this library.
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it as
 * the checker and as the host-CPU baseline.
 *
 * It restates the reference's CPU (FFTW + LAPACK) branch of the projection
 * for the channel case (x periodic, y walls, z periodic), its x
 * inflow/outflow, periodic-y and 4-wall duct branches on one rank; every
 * function cites the reference lines it follows in dfp_oracle.c.
 *
 * Arrays are the reference's Fortran arrays by address (column-major, x
 * fastest):
 *   U(nx,nyg,nzg)  V(nxg,ny,nzg)  W(nxg,nyg,nz)  P(nxg,nyg,nzg)
 *   rhs_p(2:nxg, 2:nyg-1, 2:nzg)  -> extents (nxg-1, nyg-2, nzg-1)
 * with nx=nxp+2, ny=nym+1, nz=nzp+2, nxg=nxp+3, nyg=nym+2, nzg=nzp+3 (x
 * inflow/outflow: nx=nxp+1, nxg=nxp+2).
 */
#ifndef DFP_ORACLE_H
#define DFP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_grid {
    int nxp, nym, nzp;
    double dx, dz;        /* dxmin, dzmin (initialization.f90:279-281, 326-327) */
    const double *x;      /* faces  x(1:nx)   */
    const double *y;      /* faces  y(1:ny)   */
    const double *z;      /* faces  z(1:nz)   */
    const double *yg;     /* centres yg(1:nyg) */
    const double *zg;     /* centres zg(1:nzg) */
    int x_bc_type;        /* 0 periodic; 1 inflow/outflow (nx = nxp+1, nxg = nxp+2) */
    int y_bc_type;        /* 1 walls; 0 periodic (uniform y, nym-1 independent rows) */
    int z_bc_type;        /* 0 periodic; 1 walls (the 4-wall duct: x periodic, y walls;
                             nzp = nzm cells, nz = nzp+1, nzg = nzp+2) */
} ora_grid;

typedef struct ora_plan ora_plan;

/* Grid of genGridandIC.f90:22-61 + initialization.f90:243-281.  Output
 * arrays: x[nx], y[ny], z[nz], xg[nxg], yg[nyg], zg[nzg]; lengths are the
 * namelist (face-to-face) lengths. */
void ora_channel_grid(int nxp, int nym, int nzp, double Lx_i, double Ly_i, double Lz_i,
                      double alpha, int grid_type, double *x, double *y, double *z,
                      double *xg, double *yg, double *zg, double *dx, double *dz);

/* the same with the x boundary type (1: nx = nxp+1 faces over Lx_i) */
void ora_channel_grid_x(int nxp, int nym, int nzp, double Lx_i, double Ly_i, double Lz_i,
                        double alpha, int grid_type, int x_bc_type, double *x, double *y, double *z,
                        double *xg, double *yg, double *zg, double *dx, double *dz);

ora_plan *ora_plan_create(const ora_grid *g, int nthreads);
void ora_plan_destroy(ora_plan *p);

/* operators: kxx[nxp], kzz[nzp]; Dyy tridiagonal as lower[nym-1], diag[nym], upper[nym-1] */
void ora_operators(const ora_plan *p, double *kxx, double *kzz, double *dl, double *d, double *du);

/* x inflow/outflow: outlet backflow state (outflow_backflow_beta, outflow_Uc,
 * Pb(2:nyg-1, 2:nzg-1) as a (nyg-2) x (nzg-2) Fortran array); beta = 0 turns
 * the term off.  ora_project then applies set_outflow_backflow_pressure
 * (projection.f90:65-96) with c = rk_coef(rk_step,rk_step) and the outflow
 * face correction with phib (:578-589).  ora_phib copies phib out. */
int ora_set_outflow(ora_plan *p, double beta, double outflow_Uc, const double *Pb);
void ora_phib(const ora_plan *p, double *phib);

/* DCT-IV with FFTW's REDFT11 scaling along contiguous lines of n points */
void ora_dct4_many(int n, int howmany, double *data);

/* projection.f90:37-63 */
void ora_divergence(const ora_plan *p, const double *U, const double *V, const double *W, double *rhs_p);
/* projection.f90:100-382 CPU branch; P may be NULL (no pressure save) */
int ora_solve(ora_plan *p, double *rhs_p, double *P, double dt, double c);
/* projection.f90:551-604 */
void ora_correct(const ora_plan *p, double *U, double *V, double *W, const double *rhs_p);
/* projection.f90:21-34 */
int ora_project(ora_plan *p, double *U, double *V, double *W, double *rhs_p, double *P,
                double dt, double c);
/* projection.f90:607-639 (no IBM) */
double ora_max_divergence(const ora_plan *p, const double *U, const double *V, const double *W);

/* LAPACK zgtsv on one system, complex stored as interleaved (re,im) doubles;
 * dl, d, du are real here (they are real-valued in the reference) and are
 * overwritten like zgtsv's; returns info. */
int ora_zgtsv(int n, double *dl, double *d, double *du, double *b);

/* batched unnormalised complex FFT, sign -1 forward / +1 backward, FFTW
 * plan_many semantics (interleaved complex doubles) */
int ora_fft_many(int n, int howmany, double *data, long stride, long dist, int sign);

/* RK3 stage update (time_integration.f90:256-258, :394-396, :526-531): the
 * interior of U, V, W from Uo, Vo, Wo and the right-hand sides Fu[s], Fv[s],
 * Fw[s] (interior-shaped arrays, s < stage) with c = rk_coef(stage,1:stage) */
void ora_rk_update(const ora_plan *p, int stage, double *U, double *V, double *W, const double *Uo, const double *Vo,
                   const double *Wo, const double *const *Fu, const double *const *Fv, const double *const *Fw,
                   const double *c, double dt);
/* apply_boundary_conditions of the channel (boundary_conditions.f90:186-260):
 * alpha_x(nx,2,nzg), alpha_y(nxg,2,nzg), alpha_z(nxg,2,nz) */
void ora_apply_bc(const ora_plan *p, double *U, double *V, double *W, const double *alpha_x, const double *alpha_y,
                  const double *alpha_z);

/* threads the plan's loops use (OpenMP); returns the count in effect */
int ora_threads(const ora_plan *p);

#ifdef __cplusplus
}
#endif
#endif
