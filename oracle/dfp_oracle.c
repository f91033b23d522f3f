/*
 * C restatement of dopamine-fdm's channel pressure projection (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the product never links or calls this code.  It
 * is the checker the GPU parity tests compare against and the host-CPU
 * baseline that bench.py times.
 *
 * Reference lines restated (src/...):
 *   grid .............. genGridandIC.f90:22-61, initialization.f90:243-281, :326-345
 *   kxx, kzz .......... initialization.f90:416-434
 *   Dyy ............... initialization.f90:506-535
 *   divergence ........ projection.f90:37-63
 *   solve (CPU) ....... projection.f90:100-382: copy :121, dcmplx + x C2C FFT
 *                       :124-125, z C2C FFT :134-136, Zgtsv per mode :277-297
 *                       with the (0,0) pin :292, inverse z :308-310, inverse x
 *                       + Real()/(nxp*nzp) :325-327, copy back :342,
 *                       periodic duplicates :441-500, P save :361-380,
 *                       finish_pressure_ghosts :388-438 with
 *                       boundary_conditions.f90:277-289, :591-602
 *   correction ........ projection.f90:551-604
 *   max divergence .... projection.f90:606-639
 *   periodic y ........ y_bc_type = 0 with periodic x and z: kyy
 *                       (initialization.f90:467-491, nyp = nym-1, dy =
 *                       y(2)-y(1)), y C2C FFT per z slab, division by
 *                       kxx+kyy+kzz with the (0,0,0) mode set to zero, inverse
 *                       y FFT /nyp (projection.f90:246-273), duplicate row
 *                       rhs_p(:,nyg-1,:) = rhs_p(:,2,:) (:350-354)
 *   RK3 stage update .. time_integration.f90:256-258, :394-396, :526-531
 *   channel BCs ....... apply_boundary_conditions, boundary_conditions.f90:186-260
 *                       (x periodic :263-289, z periodic :580-602, y Robin
 *                       walls :784-804; one rank: the halo exchanges
 *                       :170-183, :809-883 do nothing)
 *   x inflow/outflow .. nxp = nxm (initialization.f90:331-338), DCT-IV
 *                       wavenumbers :436-449, REDFT11 forward
 *                       projection.f90:126-129 and inverse :333-339 with
 *                       /(2*nxp*nzp), no (0,0) pin (:292), backflow term
 *                       :65-96, no x periodic fill :446-468, Neumann x
 *                       ghosts of P :413-422, outflow face :578-589
 * Third-party algorithms restated from their published definitions:
 *   LAPACK zgtsv (Gaussian elimination with partial pivoting on a
 *   tridiagonal matrix; LAPACK version unpinned by the reference,
 *   CMakeLists.txt:50-51) and FFTW's unnormalised complex DFT
 *   (FFTW_FORWARD = e^{-2 pi i jk/n}); the DFT is computed with this file's
 *   own mixed-radix Stockham FFT.  FFTW's REDFT11 (DCT-IV,
 *   y_k = 2 sum_n x_n cos(pi (2n+1)(2k+1) / 4n)) is computed from a
 *   half-length complex DFT with pre- and post-twiddles for even n, by its
 *   definition for odd n.
 * At one rank the 2decomp&fft transposes are data copies that do not change
 * values; they are left out.
 *
 * Pinned against the independent numpy restatement (oracle/ref_numpy.py) and
 * the golden fixtures in tests/golden by tests/test_oracle_c.py.
 */
#include "dfp_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* -------------------------------------------------------------------------
 * mixed-radix Stockham FFT (decimation in frequency, natural-order output)
 * ------------------------------------------------------------------------- */
#define MAXST 40
typedef struct fplan {
    int n, nst;
    int rad[MAXST];
    cplx *tw[MAXST];   /* tw[st][p*(r-1)+(u-1)] = exp(-2 pi i p u / ncur) */
    cplx *root[MAXST]; /* generic radix: exp(-2 pi i q / r), q = 0..r-1 */
} fplan;

static cplx expi_index(long t, long n, int sign) {
    /* exp(sign * 2 pi i t / n) from the exact index, in long double */
    long tt = t % n;
    if (tt < 0) tt += n;
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)tt / (long double)n;
    cplx r = {(double)cosl(a), (double)(sign * sinl(a))};
    return r;
}

static void fplan_init(fplan *f, int n) {
    memset(f, 0, sizeof(*f));
    f->n = n;
    int m = n, nst = 0;
    while (m % 4 == 0) { f->rad[nst++] = 4; m /= 4; }
    while (m % 2 == 0 && m > 1) { f->rad[nst++] = 2; m /= 2; }
    while (m % 3 == 0) { f->rad[nst++] = 3; m /= 3; }
    while (m % 5 == 0) { f->rad[nst++] = 5; m /= 5; }
    for (int p = 7; m > 1; p += 2) {
        while (m % p == 0) { f->rad[nst++] = p; m /= p; }
    }
    f->nst = nst;
    int ncur = n;
    for (int st = 0; st < nst; st++) {
        int r = f->rad[st], mm = ncur / r;
        f->tw[st] = (cplx *)malloc(sizeof(cplx) * (size_t)(mm * (r - 1) + 1));
        for (int p = 0; p < mm; p++)
            for (int u = 1; u < r; u++) f->tw[st][p * (r - 1) + (u - 1)] = expi_index((long)p * u, ncur, -1);
        if (r != 2 && r != 3 && r != 4) {
            f->root[st] = (cplx *)malloc(sizeof(cplx) * (size_t)r);
            for (int q = 0; q < r; q++) f->root[st][q] = expi_index(q, r, -1);
        }
        ncur = mm;
    }
}

static void fplan_free(fplan *f) {
    for (int st = 0; st < f->nst; st++) { free(f->tw[st]); free(f->root[st]); }
}

/* One Stockham pass over "vectors" of V complex values per element.
 * src element e occupies src[e*V .. e*V+V-1]. */
static void stockham(const fplan *f, cplx *x, cplx *y, int V, int sign) {
    int ncur = f->n, s = 1;
    cplx *src = x, *dst = y;
    cplx a[64], b[64];
    for (int st = 0; st < f->nst; st++) {
        const int r = f->rad[st], m = ncur / r;
        const cplx *tw = f->tw[st];
        for (int p = 0; p < m; p++) {
            for (int q = 0; q < s; q++) {
                for (int v = 0; v < V; v++) {
                    for (int t = 0; t < r; t++) a[t] = src[((size_t)q + (size_t)s * (p + (size_t)t * m)) * V + v];
                    if (r == 2) {
                        b[0] = cadd(a[0], a[1]);
                        b[1] = csub(a[0], a[1]);
                    } else if (r == 4) {
                        cplx s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
                        cplx s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
                        /* forward: omega = -i ; backward: omega = +i */
                        cplx jd13 = {-sign * d13.im, sign * d13.re}; /* sign*i*d13 */
                        b[0] = cadd(s02, s13);
                        b[2] = csub(s02, s13);
                        b[1] = cadd(d02, jd13);
                        b[3] = csub(d02, jd13);
                    } else if (r == 3) {
                        const double s3 = 0.86602540378443864676372317075294;
                        cplx t1 = cadd(a[1], a[2]);
                        cplx t2 = {a[0].re - 0.5 * t1.re, a[0].im - 0.5 * t1.im};
                        cplx d = csub(a[1], a[2]);
                        cplx jd = {-sign * s3 * d.im, sign * s3 * d.re}; /* sign*i*(sqrt3/2)*d */
                        b[0] = cadd(a[0], t1);
                        b[1] = cadd(t2, jd);
                        b[2] = csub(t2, jd);
                    } else {
                        const cplx *rt = f->root[st];
                        for (int u = 0; u < r; u++) {
                            cplx acc = {0.0, 0.0};
                            for (int t = 0; t < r; t++) {
                                cplx w = rt[(t * u) % r];
                                if (sign > 0) w.im = -w.im;
                                acc = cadd(acc, cmul(a[t], w));
                            }
                            b[u] = acc;
                        }
                    }
                    dst[((size_t)q + (size_t)s * ((size_t)r * p)) * V + v] = b[0];
                    for (int u = 1; u < r; u++) {
                        cplx w = tw[p * (r - 1) + (u - 1)];
                        if (sign > 0) w.im = -w.im;
                        dst[((size_t)q + (size_t)s * ((size_t)r * p + u)) * V + v] = cmul(b[u], w);
                    }
                }
            }
        }
        cplx *t = src; src = dst; dst = t;
        s *= r;
        ncur = m;
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)f->n * V);
}

#define VB 16 /* lines transformed together when the batch index is contiguous */

/* FFTW plan_many semantics on interleaved complex data */
static void fft_many(const fplan *f, cplx *data, long howmany, long stride, long dist, int sign, int nthreads) {
    const int n = f->n;
    if (dist == 1 && stride >= VB) {
        long nblk = (howmany + VB - 1) / VB;
#pragma omp parallel num_threads(nthreads)
        {
            cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n * VB);
            cplx *y = (cplx *)malloc(sizeof(cplx) * (size_t)n * VB);
#pragma omp for schedule(static)
            for (long blk = 0; blk < nblk; blk++) {
                long b0 = blk * VB;
                int V = (int)((howmany - b0) < VB ? (howmany - b0) : VB);
                for (int e = 0; e < n; e++)
                    memcpy(&x[(size_t)e * V], &data[b0 + e * stride], sizeof(cplx) * (size_t)V);
                stockham(f, x, y, V, sign);
                for (int e = 0; e < n; e++)
                    memcpy(&data[b0 + e * stride], &x[(size_t)e * V], sizeof(cplx) * (size_t)V);
            }
            free(x);
            free(y);
        }
    } else {
#pragma omp parallel num_threads(nthreads)
        {
            cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n);
            cplx *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
            for (long b = 0; b < howmany; b++) {
                for (int e = 0; e < n; e++) x[e] = data[b * dist + e * stride];
                stockham(f, x, y, 1, sign);
                for (int e = 0; e < n; e++) data[b * dist + e * stride] = x[e];
            }
            free(x);
            free(y);
        }
    }
}

int ora_fft_many(int n, int howmany, double *data, long stride, long dist, int sign) {
    if (n <= 0 || howmany < 0) return -1;
    fplan f;
    fplan_init(&f, n);
    fft_many(&f, (cplx *)data, howmany, stride, dist, sign, 1);
    fplan_free(&f);
    return 0;
}

/* -------------------------------------------------------------------------
 * LAPACK zgtsv with real-valued DL, D, DU (the reference's Dyy-derived
 * coefficients are real; their complex storage has zero imaginary parts, so
 * CABS1 reduces to |x| and every coefficient update stays real).
 * ------------------------------------------------------------------------- */
static int zgtsv_real_coef(int n, double *dl, double *d, double *du, cplx *b) {
    if (n == 0) return 0;
    for (int k = 0; k < n - 1; k++) {
        if (dl[k] == 0.0) {
            if (d[k] == 0.0) return k + 1;
        } else if (fabs(d[k]) >= fabs(dl[k])) {
            /* no row interchange */
            double mult = dl[k] / d[k];
            d[k + 1] = d[k + 1] - mult * du[k];
            b[k + 1].re = b[k + 1].re - mult * b[k].re;
            b[k + 1].im = b[k + 1].im - mult * b[k].im;
            if (k < n - 2) dl[k] = 0.0;
        } else {
            /* interchange rows k and k+1 */
            double mult = d[k] / dl[k];
            d[k] = dl[k];
            double temp = d[k + 1];
            d[k + 1] = du[k] - mult * temp;
            if (k < n - 2) {
                dl[k] = du[k + 1];
                du[k + 1] = -mult * dl[k];
            }
            du[k] = temp;
            cplx tb = b[k];
            b[k] = b[k + 1];
            b[k + 1].re = tb.re - mult * b[k + 1].re;
            b[k + 1].im = tb.im - mult * b[k + 1].im;
        }
    }
    if (d[n - 1] == 0.0) return n;
    b[n - 1].re = b[n - 1].re / d[n - 1];
    b[n - 1].im = b[n - 1].im / d[n - 1];
    if (n > 1) {
        b[n - 2].re = (b[n - 2].re - du[n - 2] * b[n - 1].re) / d[n - 2];
        b[n - 2].im = (b[n - 2].im - du[n - 2] * b[n - 1].im) / d[n - 2];
    }
    for (int k = n - 3; k >= 0; k--) {
        b[k].re = (b[k].re - du[k] * b[k + 1].re - dl[k] * b[k + 2].re) / d[k];
        b[k].im = (b[k].im - du[k] * b[k + 1].im - dl[k] * b[k + 2].im) / d[k];
    }
    return 0;
}

int ora_zgtsv(int n, double *dl, double *d, double *du, double *b) {
    return zgtsv_real_coef(n, dl, d, du, (cplx *)b);
}

/* -------------------------------------------------------------------------
 * grid (genGridandIC.f90:22-61, initialization.f90:243-281, :326-327)
 * ------------------------------------------------------------------------- */
static void centres_with_ghosts(int n, const double *f, double *g) {
    /* initialization.f90:249-272 */
    for (int i = 0; i < n - 1; i++) g[i + 1] = 0.5 * (f[i] + f[i + 1]);
    g[0] = g[1] - 2.0 * (g[1] - f[0]);
    g[n] = g[n - 1] + 2.0 * (f[n - 1] - g[n - 1]);
}

void ora_channel_grid(int nxp, int nym, int nzp, double Lx_i, double Ly_i, double Lz_i,
                      double alpha, int grid_type, double *x, double *y, double *z,
                      double *xg, double *yg, double *zg, double *dx, double *dz) {
    ora_channel_grid_x(nxp, nym, nzp, Lx_i, Ly_i, Lz_i, alpha, grid_type, 0, x, y, z, xg, yg, zg, dx, dz);
}

void ora_channel_grid_x(int nxp, int nym, int nzp, double Lx_i, double Ly_i, double Lz_i,
                        double alpha, int grid_type, int x_bc_type, double *x, double *y, double *z,
                        double *xg, double *yg, double *zg, double *dx, double *dz) {
    int nx = nxp + (x_bc_type ? 1 : 2), ny = nym + 1, nz = nzp + 2;
    for (int i = 1; i <= nx; i++) x[i - 1] = Lx_i * (double)(i - 1) / (double)(nx - 1);
    for (int i = 1; i <= nz; i++) z[i - 1] = Lz_i * (double)(i - 1) / (double)(nz - 1);
    if (grid_type == 1) {
        for (int i = 1; i <= ny; i++) y[i - 1] = Ly_i * (double)(i - 1) / (double)(ny - 1);
    } else {
        double th = tanh(0.5 * alpha);
        for (int i = 1; i <= ny; i++) {
            double eta = (double)(i - 1) / (double)(ny - 1);
            y[i - 1] = Ly_i * 0.5 * (1.0 + tanh(alpha * (eta - 0.5)) / th);
        }
    }
    centres_with_ghosts(nx, x, xg);
    centres_with_ghosts(ny, y, yg);
    centres_with_ghosts(nz, z, zg);
    double mx = xg[1] - xg[0], mz = zg[1] - zg[0];
    for (int i = 1; i < nx; i++) if (xg[i + 1] - xg[i] < mx) mx = xg[i + 1] - xg[i];
    for (int i = 1; i < nz; i++) if (zg[i + 1] - zg[i] < mz) mz = zg[i + 1] - zg[i];
    *dx = mx;
    *dz = mz;
}

/* -------------------------------------------------------------------------
 * plan
 * ------------------------------------------------------------------------- */
struct ora_plan {
    int xbc;  /* x_bc_type */
    int ybc;  /* y_bc_type */
    double *kyy;  /* periodic y: nym-1 modified wavenumbers */
    fplan fy;     /* periodic y: length nym-1 */
    double beta, outflow_Uc;
    double *Pb, *phib; /* (nyg-2) x (nzg-2) */
    int nxp, nym, nzp;
    int nx, ny, nz, nxg, nyg, nzg;
    double dx, dz;
    double *x, *y, *z, *yg, *zg;
    double *kxx, *kzz;
    double *dyy_l, *dyy_d, *dyy_u; /* Dyy(j+1,j), Dyy(j,j), Dyy(j,j+1) */
    fplan fx, fz;
    int zbc;          /* z_bc_type: 1 = the duct's z walls */
    double *lam_z;    /* duct: eigenvalues of the z-wall operator, ascending (nzp) */
    double *Qz;       /* duct: orthonormal eigenvectors of S, column m at Qz + m*nzp */
    double *sqrt_w;   /* duct: square roots of the z cell widths */
    double *dzz_l, *dzz_d, *dzz_u; /* z walls with a periodic y: Dzz(k+1,k), Dzz(k,k), Dzz(k,k+1) */
    cplx *work; /* poisson_y_c at np = 1: nxp*nym*nzp */
    int nthreads;
};

static void wavenumbers_dct(int n, double d, double *k) {
    /* initialization.f90:443-445 */
    const double pi = 4.0 * atan(1.0);
    for (int i = 0; i < n; i++) k[i] = 2.0 * (cos(pi * (double)(2 * i + 1) / (2.0 * (double)n)) - 1.0) / pow(d, 2.0);
}

/* DCT-IV, REDFT11 scaling: y_k = 2 sum_n x_n cos(pi (2n+1)(2k+1) / (4N)).
 * Even N = 2M: v_n = (x_{2n} + i x_{N-1-2n}) exp(-i pi (4n+1) / (4N)),
 * V = DFT_M(v), Z_k = V_k exp(-i pi k / N); then y_{2k} = 2 Re Z_k and
 * y_{N-1-2k} = -2 Im Z_k.  Odd N: the definition (O(N^2)). */
static void dct4_many(int n, long howmany, double *data, int nthreads) {
    const long double pil = 3.14159265358979323846264338327950288L;
    if (n % 2) {
#pragma omp parallel num_threads(nthreads)
        {
            double *t = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
            for (long b = 0; b < howmany; b++) {
                double *x = data + b * (long)n;
                for (int k = 0; k < n; k++) {
                    long double acc = 0.0L;
                    for (int q = 0; q < n; q++) {
                        long m8 = ((long)(2 * q + 1) * (2 * k + 1)) % (8L * n);
                        acc += (long double)x[q] * cosl(pil * (long double)m8 / (4.0L * n));
                    }
                    t[k] = (double)(2.0L * acc);
                }
                memcpy(x, t, sizeof(double) * (size_t)n);
            }
            free(t);
        }
        return;
    }
    const int M = n / 2;
    cplx *pre = (cplx *)malloc(sizeof(cplx) * (size_t)M);
    cplx *post = (cplx *)malloc(sizeof(cplx) * (size_t)M);
    for (int q = 0; q < M; q++) {
        /* exp(-i pi (4q+1)/(4N)) = exp(-2 pi i (4q+1) / (8N)); exp(-i pi q / N) = exp(-2 pi i q / (2N)) */
        pre[q] = expi_index(4L * q + 1, 8L * n, -1);
        post[q] = expi_index(q, 2L * n, -1);
    }
    fplan f;
    fplan_init(&f, M);
    cplx *v = (cplx *)malloc(sizeof(cplx) * (size_t)M * (size_t)howmany);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long b = 0; b < howmany; b++) {
        const double *x = data + b * (long)n;
        for (int q = 0; q < M; q++) {
            cplx a = {x[2 * q], x[n - 1 - 2 * q]};
            v[b * M + q] = cmul(a, pre[q]);
        }
    }
    fft_many(&f, v, howmany, 1, M, -1, nthreads);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long b = 0; b < howmany; b++) {
        double *x = data + b * (long)n;
        for (int q = 0; q < M; q++) {
            cplx z = cmul(v[b * M + q], post[q]);
            x[2 * q] = 2.0 * z.re;
            x[n - 1 - 2 * q] = -2.0 * z.im;
        }
    }
    free(v);
    free(pre);
    free(post);
    fplan_free(&f);
}

void ora_dct4_many(int n, int howmany, double *data) { dct4_many(n, howmany, data, 1); }

static double *dup_array(const double *a, int n) {
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(r, a, sizeof(double) * (size_t)n);
    return r;
}

static void wavenumbers(int n, double d, double *k) {
    /* initialization.f90:420-425 / :428-433 */
    const double pi = 4.0 * atan(1.0);
    int half = (int)ceil((double)(float)n / 2.0);
    for (int i = 0; i < n; i++) k[i] = 0.0;
    for (int i = 0; i <= half && i <= n - 1; i++) k[i] = 2.0 * (cos(2.0 * pi * (double)i / (double)n) - 1.0) / pow(d, 2.0);
    for (int i = half + 1; i <= n - 1; i++) k[i] = 2.0 * (cos(2.0 * pi * (double)(-n + i) / (double)n) - 1.0) / pow(d, 2.0);
}

/* Cyclic Jacobi eigenvalue iteration for a dense symmetric n x n matrix a
   (column-major, overwritten): eigenvalues in d, ascending, and orthonormal
   eigenvectors in the columns of v.  Row-cyclic sweeps of symmetric Schur
   rotations (Golub & Van Loan, Matrix Computations, Algorithms 8.4.1-8.4.3)
   until the off-diagonal part vanishes against the diagonal. */
static void jacobi_eigen(int n, double *a, double *d, double *v) {
#define A_(i, j) a[(size_t)(j) * n + (i)]
#define V_E(i, j) v[(size_t)(j) * n + (i)]
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) V_E(i, j) = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                if (i == j) diag += A_(i, j) * A_(i, j);
                else off += A_(i, j) * A_(i, j);
            }
        if (off <= 1e-34 * diag) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A_(p, q);
                if (apq == 0.0) continue;
                const double tau = (A_(q, q) - A_(p, p)) / (2.0 * apq);
                const double t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; k++) {  /* A <- A J */
                    const double akp = A_(k, p), akq = A_(k, q);
                    A_(k, p) = c * akp - s * akq;
                    A_(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {  /* A <- J^T A */
                    const double apk = A_(p, k), aqk = A_(q, k);
                    A_(p, k) = c * apk - s * aqk;
                    A_(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {  /* V <- V J */
                    const double vkp = V_E(k, p), vkq = V_E(k, q);
                    V_E(k, p) = c * vkp - s * vkq;
                    V_E(k, q) = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) d[i] = A_(i, i);
    /* ascending order, eigenvectors along */
    for (int i = 0; i < n - 1; i++) {
        int m = i;
        for (int j = i + 1; j < n; j++)
            if (d[j] < d[m]) m = j;
        if (m != i) {
            const double t = d[i];
            d[i] = d[m];
            d[m] = t;
            for (int k = 0; k < n; k++) {
                const double w = V_E(k, i);
                V_E(k, i) = V_E(k, m);
                V_E(k, m) = w;
            }
        }
    }
#undef A_
#undef V_E
}

/* initialization.f90:544-577 (Dzz, Dyy's construction on z, zg) and
   :583-611: S = W^(1/2) Dzz W^(-1/2) with W the z cell widths, its
   eigenvalues lambda_z and orthonormal eigenvectors Qz, and sqrt_w */
static void duct_eigenbasis(ora_plan *p) {
    const int n = p->nzp;
    const double *z = p->z, *zg = p->zg;
    double *S = (double *)calloc((size_t)n * n, sizeof(double));
    p->lam_z = (double *)malloc(sizeof(double) * (size_t)n);
    p->Qz = (double *)malloc(sizeof(double) * (size_t)n * n);
    p->sqrt_w = (double *)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; k++) p->sqrt_w[k] = sqrt(z[k + 1] - z[k]);
#define ZF(k) z[(k) - 1]
#define ZG(k) zg[(k) - 1]
    for (int k = 2; k <= p->nzg - 1; k++) {
        const double a = 1.0 / (ZF(k) - ZF(k - 1)) / (ZG(k + 1) - ZG(k));
        const double b = 1.0 / (ZF(k) - ZF(k - 1)) * (-1.0 / (ZG(k + 1) - ZG(k)) - 1.0 / (ZG(k) - ZG(k - 1)));
        const double c = 1.0 / (ZF(k) - ZF(k - 1)) / (ZG(k) - ZG(k - 1));
        const int kk = k - 2;
        double dkk;
        if (k == 2 && k == p->nzg - 1) dkk = a + b;
        else if (k == 2) dkk = b + c;
        else if (k == p->nzg - 1) dkk = a + b;
        else dkk = b;
        S[(size_t)kk * n + kk] = dkk;
        if (kk + 1 < n) {
            /* S(k,k+1) = S(k+1,k) = 1 / ((zg(k+2)-zg(k+1)) sqrt(w(k) w(k+1))), 1-based cell k */
            const double off = 1.0 / ((ZG(k + 1) - ZG(k)) * sqrt((ZF(k) - ZF(k - 1)) * (ZF(k + 1) - ZF(k))));
            S[(size_t)(kk + 1) * n + kk] = off;
            S[(size_t)kk * n + kk + 1] = off;
        }
    }
#undef ZF
#undef ZG
    jacobi_eigen(n, S, p->lam_z, p->Qz);
    free(S);
}

ora_plan *ora_plan_create(const ora_grid *g, int nthreads) {
    if (g->nxp < 1 || g->nym < 1 || g->nzp < 1) return NULL;
    if (g->x_bc_type != 0 && g->x_bc_type != 1) return NULL;
    if (g->y_bc_type != 0 && g->y_bc_type != 1) return NULL;
    /* periodic y needs periodic x (initialization.f90:96-97) and two rows */
    if (g->y_bc_type == 0 && (g->x_bc_type != 0 || g->nym < 2)) return NULL;
    /* z walls with a periodic x (initialization.f90:98-100): the 4-wall duct
       (y walls) or the spanwise walls with a periodic y */
    if (g->z_bc_type != 0 && g->z_bc_type != 1) return NULL;
    if (g->z_bc_type == 1 && g->x_bc_type != 0) return NULL;
    ora_plan *p = (ora_plan *)calloc(1, sizeof(ora_plan));
    p->xbc = g->x_bc_type;
    p->ybc = g->y_bc_type;
    p->zbc = g->z_bc_type;
    p->nxp = g->nxp; p->nym = g->nym; p->nzp = g->nzp;
    p->nx = g->nxp + (p->xbc ? 1 : 2); p->ny = g->nym + 1; p->nz = g->nzp + (p->zbc ? 1 : 2);
    p->nxg = p->nx + 1; p->nyg = g->nym + 2; p->nzg = p->nz + 1;
    p->dx = g->dx; p->dz = g->dz;
    p->x = dup_array(g->x, p->nx);
    p->y = dup_array(g->y, p->ny);
    p->z = dup_array(g->z, p->nz);
    p->yg = dup_array(g->yg, p->nyg);
    p->zg = dup_array(g->zg, p->nzg);
    p->kxx = (double *)malloc(sizeof(double) * (size_t)p->nxp);
    p->kzz = (double *)malloc(sizeof(double) * (size_t)p->nzp);
    if (p->xbc) wavenumbers_dct(p->nxp, p->dx, p->kxx);
    else wavenumbers(p->nxp, p->dx, p->kxx);
    wavenumbers(p->nzp, p->dz, p->kzz);
    p->phib = (double *)calloc((size_t)(p->nyg - 2) * (size_t)(p->nzg - 2), sizeof(double));
    /* Dyy, initialization.f90:506-535; index jj = j-2 for reference row j */
    int nym = p->nym, nyg = p->nyg;
    p->dyy_l = (double *)calloc((size_t)(nym > 1 ? nym - 1 : 1), sizeof(double));
    p->dyy_u = (double *)calloc((size_t)(nym > 1 ? nym - 1 : 1), sizeof(double));
    p->dyy_d = (double *)calloc((size_t)nym, sizeof(double));
    const double *y = p->y, *yg = p->yg;
#define YF(j) y[(j) - 1]
#define YG(j) yg[(j) - 1]
    for (int j = 2; j <= nyg - 1; j++) {
        double a = 1.0 / (YF(j) - YF(j - 1)) / (YG(j + 1) - YG(j));
        double b = 1.0 / (YF(j) - YF(j - 1)) * (-1.0 / (YG(j + 1) - YG(j)) - 1.0 / (YG(j) - YG(j - 1)));
        double c = 1.0 / (YF(j) - YF(j - 1)) / (YG(j) - YG(j - 1));
        int jj = j - 2;
        if (j == 2 && j == nyg - 1) {
            /* one cell between the walls: initialization.f90:523-535 assigns
               Dyy(2,2) = b+c and then Dyy(nyg-1,nyg-1) = a+b, the same entry */
            p->dyy_d[jj] = a + b;
        } else if (j == 2) {
            p->dyy_d[jj] = b + c;
            p->dyy_u[jj] = a;
        } else if (j == nyg - 1) {
            p->dyy_d[jj] = a + b;
            p->dyy_l[jj - 1] = c;
        } else {
            p->dyy_u[jj] = a;
            p->dyy_l[jj - 1] = c;
            p->dyy_d[jj] = b;
        }
    }
#undef YF
#undef YG
    fplan_init(&p->fx, p->nxp);
    fplan_init(&p->fz, p->nzp);
    if (!p->ybc) {
        /* initialization.f90:481-489, dy = y(2)-y(1) on the uniform grid */
        p->kyy = (double *)malloc(sizeof(double) * (size_t)(p->nym - 1));
        wavenumbers(p->nym - 1, p->y[1] - p->y[0], p->kyy);
        fplan_init(&p->fy, p->nym - 1);
    }
    if (p->zbc && p->ybc) duct_eigenbasis(p);
    if (p->zbc && !p->ybc) {
        /* initialization.f90:544-577: Dzz (Dyy's construction on z, zg), row kk = k-2 */
        const int n = p->nzp, nzg = p->nzg;
        const double *z = p->z, *zg = p->zg;
        p->dzz_l = (double *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(double));
        p->dzz_u = (double *)calloc((size_t)(n > 1 ? n - 1 : 1), sizeof(double));
        p->dzz_d = (double *)calloc((size_t)n, sizeof(double));
        for (int k = 2; k <= nzg - 1; k++) {
            const double a = 1.0 / (z[k - 1] - z[k - 2]) / (zg[k] - zg[k - 1]);
            const double b = 1.0 / (z[k - 1] - z[k - 2]) * (-1.0 / (zg[k] - zg[k - 1]) - 1.0 / (zg[k - 1] - zg[k - 2]));
            const double c = 1.0 / (z[k - 1] - z[k - 2]) / (zg[k - 1] - zg[k - 2]);
            const int kk = k - 2;
            if (k == 2 && k == nzg - 1) {
                p->dzz_d[kk] = a + b;
            } else if (k == 2) {
                p->dzz_d[kk] = b + c;
                p->dzz_u[kk] = a;
            } else if (k == nzg - 1) {
                p->dzz_d[kk] = a + b;
                p->dzz_l[kk - 1] = c;
            } else {
                p->dzz_u[kk] = a;
                p->dzz_l[kk - 1] = c;
                p->dzz_d[kk] = b;
            }
        }
    }
    p->work = (cplx *)malloc(sizeof(cplx) * (size_t)p->nxp * p->nym * p->nzp);
#ifdef _OPENMP
    p->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    p->nthreads = 1;
#endif
    return p;
}

void ora_plan_destroy(ora_plan *p) {
    if (!p) return;
    free(p->x); free(p->y); free(p->z); free(p->yg); free(p->zg);
    free(p->kxx); free(p->kzz);
    free(p->dyy_l); free(p->dyy_d); free(p->dyy_u);
    free(p->Pb); free(p->phib);
    fplan_free(&p->fx); fplan_free(&p->fz);
    if (!p->ybc) { free(p->kyy); fplan_free(&p->fy); }
    if (p->zbc && p->ybc) { free(p->lam_z); free(p->Qz); free(p->sqrt_w); }
    if (p->zbc && !p->ybc) { free(p->dzz_l); free(p->dzz_d); free(p->dzz_u); }
    free(p->work);
    free(p);
}

int ora_threads(const ora_plan *p) { return p->nthreads; }

void ora_operators(const ora_plan *p, double *kxx, double *kzz, double *dl, double *d, double *du) {
    if (kxx) memcpy(kxx, p->kxx, sizeof(double) * (size_t)p->nxp);
    if (kzz) memcpy(kzz, p->kzz, sizeof(double) * (size_t)p->nzp);
    if (d) memcpy(d, p->dyy_d, sizeof(double) * (size_t)p->nym);
    if (p->nym > 1) {
        if (dl) memcpy(dl, p->dyy_l, sizeof(double) * (size_t)(p->nym - 1));
        if (du) memcpy(du, p->dyy_u, sizeof(double) * (size_t)(p->nym - 1));
    }
}

/* Fortran-order element addresses (1-based reference indices) */
#define IDX3(a, n1, n2, i, j, k) ((a)[((size_t)((k) - 1) * (size_t)(n2) + (size_t)((j) - 1)) * (size_t)(n1) + (size_t)((i) - 1)])
#define U_(i, j, k) IDX3(U, p->nx, p->nyg, i, j, k)
#define V_(i, j, k) IDX3(V, p->nxg, p->ny, i, j, k)
#define W_(i, j, k) IDX3(W, p->nxg, p->nyg, i, j, k)
#define P_(i, j, k) IDX3(P, p->nxg, p->nyg, i, j, k)
/* rhs_p(2:nxg, 2:nyg-1, 2:nzg) */
#define R_(i, j, k) IDX3(rhs_p, p->nxg - 1, p->nyg - 2, (i) - 1, (j) - 1, (k) - 1)

void ora_divergence(const ora_plan *p, const double *U, const double *V, const double *W, double *rhs_p) {
    /* projection.f90:42-61 */
    const double inv_dx_p = 1.0 / p->dx;
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
    for (int k = 2; k <= p->nzg - 1; k++) {
        for (int j = 2; j <= p->nyg - 1; j++) {
            double inv_dzk = 1.0 / (p->z[k - 1] - p->z[k - 2]);
            double inv_dyj = 1.0 / (p->y[j - 1] - p->y[j - 2]);
            for (int i = 2; i <= p->nxg - 1; i++) {
                R_(i, j, k) = (U_(i, j, k) - U_(i - 1, j, k)) * inv_dx_p +
                              (V_(i, j, k) - V_(i, j - 1, k)) * inv_dyj +
                              (W_(i, j, k) - W_(i, j, k - 1)) * inv_dzk;
            }
        }
    }
}

int ora_set_outflow(ora_plan *p, double beta, double outflow_Uc, const double *Pb) {
    if (beta < 0.0 || (beta > 0.0 && (!p->xbc || !Pb))) return -1;  /* input_output.f90:312-314 */
    p->beta = beta;
    p->outflow_Uc = outflow_Uc;
    size_t n = (size_t)(p->nyg - 2) * (size_t)(p->nzg - 2);
    free(p->Pb);
    p->Pb = NULL;
    if (Pb) {
        p->Pb = (double *)malloc(sizeof(double) * n);
        memcpy(p->Pb, Pb, sizeof(double) * n);
    }
    return 0;
}

void ora_phib(const ora_plan *p, double *phib) {
    memcpy(phib, p->phib, sizeof(double) * (size_t)(p->nyg - 2) * (size_t)(p->nzg - 2));
}

#define PB_(j, k) p->Pb[(size_t)((k) - 2) * (size_t)(p->nyg - 2) + (size_t)((j) - 2)]
#define PHIB_(j, k) p->phib[(size_t)((k) - 2) * (size_t)(p->nyg - 2) + (size_t)((j) - 2)]

static void set_outflow_backflow_pressure(ora_plan *p, const double *U, const double *V, const double *W,
                                          double *rhs_p, double c) {
    /* projection.f90:73-94, c = dt*rk_coef(rk_step,rk_step) */
    const int nx = p->nx;
    const double uref = 0.05 * (p->outflow_Uc > 1e-2 ? p->outflow_Uc : 1e-2);
    const double inv_dx2 = 1.0 / pow(p->dx, 2);
    for (int k = 2; k <= p->nzg - 1; k++)
        for (int j = 2; j <= p->nyg - 1; j++) {
            double un = U_(nx, j, k);
            double vt = 0.5 * (V_(nx, j, k) + V_(nx, j - 1, k));
            double wt = 0.5 * (W_(nx, j, k) + W_(nx, j, k - 1));
            double th = 0.5 * (1.0 - tanh(un / uref));
            PHIB_(j, k) = c * (-0.5 * p->beta * (un * un + vt * vt + wt * wt) * th - PB_(j, k));
            R_(nx, j, k) = R_(nx, j, k) - 2.0 * PHIB_(j, k) * inv_dx2;
        }
}

static void apply_periodic_xz_pressure(const ora_plan *p, double *rhs_p) {
    /* projection.f90:453-456 and :479-482 (single rank); no x fill for the
       inflow/outflow x (:446-448) */
    if (!p->xbc)
        for (int k = 2; k <= p->nzg; k++)
            for (int j = 2; j <= p->nyg - 1; j++) R_(p->nxg - 1, j, k) = R_(2, j, k);
    if (p->zbc) return; /* z walls: no z fill (:472-497) */
    for (int j = 2; j <= p->nyg - 1; j++)
        for (int i = 2; i <= p->nxg - 1; i++) R_(i, j, p->nzp + 2) = R_(i, j, 2);
}

static void finish_pressure_ghosts(const ora_plan *p, double *P) {
    /* projection.f90:396-410, boundary_conditions.f90:283-285 and :597-599 (id = 4) */
    int nxg = p->nxg, nyg = p->nyg, nzg = p->nzg;
    for (int k = 1; k <= nzg; k++)
        for (int i = 1; i <= nxg; i++) {
            P_(i, 1, k) = P_(i, 2, k);
            P_(i, nyg, k) = P_(i, nyg - 1, k);
        }
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) {
            if (p->xbc) {
                /* :413-422 Neumann copies at the two x ends */
                P_(1, j, k) = P_(2, j, k);
                P_(nxg, j, k) = P_(nxg - 1, j, k);
            } else {
                P_(1, j, k) = P_(nxg - 2, j, k);
                P_(nxg - 1, j, k) = P_(2, j, k);
                P_(nxg, j, k) = P_(3, j, k);
            }
        }
    for (int j = 1; j <= nyg; j++)
        for (int i = 1; i <= nxg; i++) {
            if (p->zbc) {
                /* :423-436 Neumann copies at the z walls */
                P_(i, j, 1) = P_(i, j, 2);
                P_(i, j, nzg) = P_(i, j, nzg - 1);
            } else {
                P_(i, j, 1) = P_(i, j, nzg - 2);
                P_(i, j, nzg - 1) = P_(i, j, 2);
                P_(i, j, nzg) = P_(i, j, 3);
            }
        }
}

/* projection.f90:172-218, the duct: per x mode i, z_hat = (p * sqrt_w) Qz,
   one Zgtsv per (i, m) with kxx(i) + lambda_z(m) (the pin where i = 0 and
   |lambda_z(m)| < 1e-8 max|lambda_z|), p = z_hat Qz^T / sqrt_w */
static int duct_solve(ora_plan *p, cplx *wk) {
    const int nxp = p->nxp, nym = p->nym, nzm = p->nzp;
    double lmax = 0.0;
    for (int m = 0; m < nzm; m++) lmax = fmax(lmax, fabs(p->lam_z[m]));
    int info_all = 0;
#pragma omp parallel num_threads(p->nthreads)
    {
        cplx *zh = (cplx *)malloc(sizeof(cplx) * (size_t)nym * nzm);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)nym);
        double *DL = (double *)malloc(sizeof(double) * (size_t)(nym > 1 ? nym - 1 : 1));
        double *DU = (double *)malloc(sizeof(double) * (size_t)(nym > 1 ? nym - 1 : 1));
        double *D = (double *)malloc(sizeof(double) * (size_t)nym);
#pragma omp for schedule(static)
        for (int i = 0; i < nxp; i++) {
            for (int j = 0; j < nym; j++)
                for (int m = 0; m < nzm; m++) {
                    cplx acc = {0.0, 0.0};
                    for (int k = 0; k < nzm; k++) {
                        const cplx v = wk[((size_t)k * nym + j) * nxp + i];
                        const double q = p->Qz[(size_t)m * nzm + k];
                        acc.re += v.re * p->sqrt_w[k] * q;
                        acc.im += v.im * p->sqrt_w[k] * q;
                    }
                    zh[(size_t)m * nym + j] = acc;
                }
            for (int m = 0; m < nzm; m++) {
                const double ws = p->kxx[i] + p->lam_z[m];
                for (int j = 0; j < nym - 1; j++) {
                    DL[j] = p->dyy_l[j];
                    DU[j] = p->dyy_u[j];
                }
                for (int j = 0; j < nym; j++) D[j] = p->dyy_d[j] + ws;
                if (i == 0 && fabs(p->lam_z[m]) < 1e-8 * lmax) D[0] = 3.0 / 2.0 * D[0];
                for (int j = 0; j < nym; j++) b[j] = zh[(size_t)m * nym + j];
                const int info = zgtsv_real_coef(nym, DL, D, DU, b);
                if (info != 0) {
#pragma omp atomic write
                    info_all = info;
                }
                for (int j = 0; j < nym; j++) zh[(size_t)m * nym + j] = b[j];
            }
            for (int j = 0; j < nym; j++)
                for (int k = 0; k < nzm; k++) {
                    cplx acc = {0.0, 0.0};
                    for (int m = 0; m < nzm; m++) {
                        const cplx v = zh[(size_t)m * nym + j];
                        const double q = p->Qz[(size_t)m * nzm + k];
                        acc.re += v.re * q;
                        acc.im += v.im * q;
                    }
                    cplx *o = wk + ((size_t)k * nym + j) * nxp + i;
                    o->re = acc.re / p->sqrt_w[k];
                    o->im = acc.im / p->sqrt_w[k];
                }
        }
        free(zh); free(b); free(DL); free(DU); free(D);
    }
    return info_all;
}

int ora_solve(ora_plan *p, double *rhs_p, double *P, double dt, double c) {
    const int nxp = p->nxp, nym = p->nym, nzp = p->nzp;
    cplx *wk = p->work;
    if (p->xbc) {
        /* :121 copy, :127 REDFT11 along x, :128 dcmplx */
        double *xr = (double *)malloc(sizeof(double) * (size_t)nxp * nym * nzp);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 0; k < nzp; k++)
            for (int j = 0; j < nym; j++)
                for (int i = 0; i < nxp; i++) xr[((size_t)k * nym + j) * nxp + i] = R_(i + 2, j + 2, k + 2);
        dct4_many(nxp, (long)nym * nzp, xr, p->nthreads);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (size_t e = 0; e < (size_t)nxp * nym * nzp; e++) {
            cplx v = {xr[e], 0.0};
            wk[e] = v;
        }
        free(xr);
    } else {
        /* :121 poisson_y_r = rhs_p(2:nxp+1, 2:nym+1, 2:nzp+1); :124 dcmplx */
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 0; k < nzp; k++)
            for (int j = 0; j < nym; j++)
                for (int i = 0; i < nxp; i++) {
                    cplx v = {R_(i + 2, j + 2, k + 2), 0.0};
                    wk[((size_t)k * nym + j) * nxp + i] = v;
                }
        /* :125 C2C forward along x, batch nym*nzp, contiguous lines */
        fft_many(&p->fx, wk, (long)nym * nzp, 1, nxp, -1, p->nthreads);
    }
    int info_all = 0;
    if (p->zbc && p->ybc) {
        info_all = duct_solve(p, wk);
        goto inverse;
    }
    if (p->zbc) {
        /* spanwise walls with a periodic y: :143-150 y FFT over the nyp rows of
           each z slab; :219-245 one Zgtsv along z per (i, j) with Dzz + kxx +
           kyy, D(2) *= 3/2 where i = j = 0; :311-315 inverse y FFT, /nyp */
        const int nyp = nym - 1;
        for (int k = 0; k < nzp; k++) fft_many(&p->fy, wk + (size_t)k * nym * nxp, nxp, nxp, 1, -1, p->nthreads);
#pragma omp parallel num_threads(p->nthreads)
        {
            double *DL = (double *)malloc(sizeof(double) * (size_t)(nzp > 1 ? nzp - 1 : 1));
            double *DU = (double *)malloc(sizeof(double) * (size_t)(nzp > 1 ? nzp - 1 : 1));
            double *D = (double *)malloc(sizeof(double) * (size_t)nzp);
            cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)nzp);
#pragma omp for schedule(static)
            for (int j = 0; j < nyp; j++)
                for (int i = 0; i < nxp; i++) {
                    const double ws = p->kxx[i] + p->kyy[j];
                    for (int k = 0; k < nzp - 1; k++) {
                        DL[k] = p->dzz_l[k];
                        DU[k] = p->dzz_u[k];
                    }
                    for (int k = 0; k < nzp; k++) D[k] = p->dzz_d[k] + ws;
                    if (i == 0 && j == 0) D[0] = 3.0 / 2.0 * D[0];
                    cplx *line = wk + (size_t)j * nxp + i;
                    for (int k = 0; k < nzp; k++) b[k] = line[(size_t)k * nym * nxp];
                    const int info = zgtsv_real_coef(nzp, DL, D, DU, b);
                    if (info != 0) {
#pragma omp atomic write
                        info_all = info;
                    }
                    for (int k = 0; k < nzp; k++) line[(size_t)k * nym * nxp] = b[k];
                }
            free(DL); free(DU); free(D); free(b);
        }
        for (int k = 0; k < nzp; k++) fft_many(&p->fy, wk + (size_t)k * nym * nxp, nxp, nxp, 1, +1, p->nthreads);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 0; k < nzp; k++)
            for (int j = 0; j < nyp; j++)
                for (int i = 0; i < nxp; i++) {
                    cplx *v = wk + ((size_t)k * nym + j) * nxp + i;
                    v->re = v->re / (double)nyp;
                    v->im = v->im / (double)nyp;
                }
        goto inverse;
    }
    /* :134-136 C2C forward along z, stride nxp*nym, batch nxp*nym */
    fft_many(&p->fz, wk, (long)nxp * nym, (long)nxp * nym, 1, -1, p->nthreads);
    if (!p->ybc) {
        /* periodic y (:246-273): per z slab, C2C forward along y over the nyp =
           nym-1 independent rows (stride nxp, batch nxp), divide by
           kxx+kyy+kzz with the (0,0,0) mode set to zero, backward, /nyp */
        const int nyp = nym - 1;
        for (int k = 0; k < nzp; k++) fft_many(&p->fy, wk + (size_t)k * nym * nxp, nxp, nxp, 1, -1, p->nthreads);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 0; k < nzp; k++)
            for (int j = 0; j < nyp; j++)
                for (int i = 0; i < nxp; i++) {
                    cplx *v = wk + ((size_t)k * nym + j) * nxp + i;
                    const double ws = p->kxx[i] + p->kyy[j] + p->kzz[k];
                    if (i == 0 && j == 0 && k == 0) {
                        v->re = 0.0;
                        v->im = 0.0;
                    } else {
                        v->re = v->re / ws;
                        v->im = v->im / ws;
                    }
                }
        for (int k = 0; k < nzp; k++) fft_many(&p->fy, wk + (size_t)k * nym * nxp, nxp, nxp, 1, +1, p->nthreads);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 0; k < nzp; k++)
            for (int j = 0; j < nyp; j++)
                for (int i = 0; i < nxp; i++) {
                    cplx *v = wk + ((size_t)k * nym + j) * nxp + i;
                    v->re = v->re / (double)nyp;
                    v->im = v->im / (double)nyp;
                }
        goto inverse;
    }
    /* :277-297 one Zgtsv per (i,k) mode */
#pragma omp parallel num_threads(p->nthreads)
    {
        double *DL = (double *)malloc(sizeof(double) * (size_t)(nym > 1 ? nym - 1 : 1));
        double *DU = (double *)malloc(sizeof(double) * (size_t)(nym > 1 ? nym - 1 : 1));
        double *D = (double *)malloc(sizeof(double) * (size_t)nym);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)nym);
#pragma omp for schedule(static)
        for (int k = 0; k < nzp; k++) {
            for (int i = 0; i < nxp; i++) {
                double wavenum_sum = p->kxx[i] + p->kzz[k];
                for (int j = 0; j < nym - 1; j++) {
                    DL[j] = p->dyy_l[j];
                    DU[j] = p->dyy_u[j];
                }
                for (int j = 0; j < nym; j++) D[j] = p->dyy_d[j] + wavenum_sum;
                if (!p->xbc && i == 0 && k == 0) D[0] = 3.0 / 2.0 * D[0];
                cplx *line = wk + (size_t)k * nym * nxp + i;
                for (int j = 0; j < nym; j++) b[j] = line[(size_t)j * nxp];
                int info = zgtsv_real_coef(nym, DL, D, DU, b);
                if (info != 0) {
#pragma omp atomic write
                    info_all = info;
                }
                for (int j = 0; j < nym; j++) line[(size_t)j * nxp] = b[j];
            }
        }
        free(DL); free(DU); free(D); free(b);
    }
inverse:
    /* :308-310 inverse z (the duct solved z in physical space) */
    if (!p->zbc) fft_many(&p->fz, wk, (long)nxp * nym, (long)nxp * nym, 1, +1, p->nthreads);
    if (p->xbc) {
        /* :333 Real(), :334 REDFT11, :336 /(2*nxp*nzp); :342 back into rhs_p */
        double *xr = (double *)malloc(sizeof(double) * (size_t)nxp * nym * nzp);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (size_t e = 0; e < (size_t)nxp * nym * nzp; e++) xr[e] = wk[e].re;
        dct4_many(nxp, (long)nym * nzp, xr, p->nthreads);
        const double norm = (double)(2L * nxp * nzp);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 0; k < nzp; k++)
            for (int j = 0; j < nym; j++)
                for (int i = 0; i < nxp; i++) R_(i + 2, j + 2, k + 2) = xr[((size_t)k * nym + j) * nxp + i] / norm;
        free(xr);
    } else {
        /* :325 inverse x; :327 Real()/(nxp*nzp), the duct Real()/nxp (:330); :342 back into rhs_p */
        fft_many(&p->fx, wk, (long)nym * nzp, 1, nxp, +1, p->nthreads);
        const double norm = p->zbc ? (double)nxp : (double)((long)nxp * nzp);
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 0; k < nzp; k++)
            for (int j = 0; j < nym; j++)
                for (int i = 0; i < nxp; i++) R_(i + 2, j + 2, k + 2) = wk[((size_t)k * nym + j) * nxp + i].re / norm;
    }
    apply_periodic_xz_pressure(p, rhs_p);
    if (!p->ybc) {
        /* :350-354 periodic y: the duplicate row, over the whole x-z extent */
        for (int k = 2; k <= p->nzg; k++)
            for (int i = 2; i <= p->nxg; i++) R_(i, p->nyg - 1, k) = R_(i, 2, k);
    }
    /* :357-358 single rank: no interior halo neighbours */
    if (P) {
        /* :377 */
        const double den = dt * c;
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
        for (int k = 2; k <= p->nzg - 1; k++)
            for (int j = 2; j <= p->nyg - 1; j++)
                for (int i = 2; i <= p->nxg - 1; i++) P_(i, j, k) = R_(i, j, k) / den;
        finish_pressure_ghosts(p, P);
    }
    return info_all ? -100 - info_all : 0;
}

void ora_correct(const ora_plan *p, double *U, double *V, double *W, const double *rhs_p) {
    /* projection.f90:558-601, one rank (the seam face nx is the domain edge) */
    const double inv_dx = 1.0 / p->dx;
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
    for (int k = 2; k <= p->nzg - 1; k++)
        for (int j = 2; j <= p->nyg - 1; j++)
            for (int i = 2; i <= p->nx - 1; i++) U_(i, j, k) = U_(i, j, k) - (R_(i + 1, j, k) - R_(i, j, k)) * inv_dx;
    if (p->xbc) {
        /* :579-589 outflow face */
        const int nx = p->nx;
        for (int k = 2; k <= p->nzg - 1; k++)
            for (int j = 2; j <= p->nyg - 1; j++) {
                if (p->beta > 0.0) U_(nx, j, k) = U_(nx, j, k) + 2.0 * (R_(nx, j, k) - PHIB_(j, k)) * inv_dx;
                else U_(nx, j, k) = U_(nx, j, k) + 2.0 * R_(nx, j, k) * inv_dx;
            }
    }
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
    for (int k = 2; k <= p->nzg - 1; k++)
        for (int j = 2; j <= p->ny - 1; j++) {
            double inv_dygj = 1.0 / (p->yg[j] - p->yg[j - 1]);
            for (int i = 2; i <= p->nxg - 1; i++) V_(i, j, k) = V_(i, j, k) - (R_(i, j + 1, k) - R_(i, j, k)) * inv_dygj;
        }
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
    for (int k = 2; k <= p->nz - 1; k++) {
        double inv_dzgk = 1.0 / (p->zg[k] - p->zg[k - 1]);
        for (int j = 2; j <= p->nyg - 1; j++)
            for (int i = 2; i <= p->nxg - 1; i++) W_(i, j, k) = W_(i, j, k) - (R_(i, j, k + 1) - R_(i, j, k)) * inv_dzgk;
    }
}

int ora_project(ora_plan *p, double *U, double *V, double *W, double *rhs_p, double *P, double dt, double c) {
    ora_divergence(p, U, V, W, rhs_p);
    if (p->beta > 0.0) set_outflow_backflow_pressure(p, U, V, W, rhs_p, dt * c);
    int rc = ora_solve(p, rhs_p, P, dt, c);
    ora_correct(p, U, V, W, rhs_p);
    return rc;
}

double ora_max_divergence(const ora_plan *p, const double *U, const double *V, const double *W) {
    double div = 0.0;
#pragma omp parallel for schedule(static) reduction(max : div) num_threads(p->nthreads)
    for (int k = 2; k <= p->nzg - 1; k++)
        for (int j = 2; j <= p->nyg - 1; j++)
            for (int i = 2; i <= p->nxg - 1; i++) {
                double d = fabs((U_(i, j, k) - U_(i - 1, j, k)) / (p->x[i - 1] - p->x[i - 2]) +
                                (V_(i, j, k) - V_(i, j - 1, k)) / (p->y[j - 1] - p->y[j - 2]) +
                                (W_(i, j, k) - W_(i, j, k - 1)) / (p->z[k - 1] - p->z[k - 2]));
                if (d > div) div = d;
            }
    return div;
}


/* -------------------------------------------------------------------------
 * the step either side of the projection (channel case, one rank)
 * ------------------------------------------------------------------------- */
/* Interior arrays Fu(2:nx-1,2:nyg-1,2:nzg-1), Fv(2:nxg-1,2:ny-1,2:nzg-1),
 * Fw(2:nxg-1,2:nyg-1,2:nz-1) (initialization.f90:662-670). */
#define FI_(a, n1, n2, i, j, k) IDX3(a, (n1) - 2, (n2) - 2, (i) - 1, (j) - 1, (k) - 1)

void ora_rk_update(const ora_plan *p, int stage, double *U, double *V, double *W, const double *Uo, const double *Vo,
                   const double *Wo, const double *const *Fu, const double *const *Fv, const double *const *Fw,
                   const double *c, double dt) {
    /* time_integration.f90:256-258 (stage 1: Uo + dt*c11*Fu1), :394-396
       (stage 2: Uo + dt*(c21*Fu1 + c22*Fu2)), :526-531 (stage 3:
       Uo + dt*(c31*Fu1 + c32*Fu2 + c33*Fu3)), operation for operation */
    for (int f = 0; f < 3; f++) {
        double *F = f == 0 ? U : (f == 1 ? V : W);
        const double *Fo = f == 0 ? Uo : (f == 1 ? Vo : Wo);
        const double *const *R = f == 0 ? Fu : (f == 1 ? Fv : Fw);
        const int n1 = f == 0 ? p->nx : p->nxg, n2 = f == 1 ? p->ny : p->nyg, n3 = f == 2 ? p->nz : p->nzg;
        for (int k = 2; k <= n3 - 1; k++)
            for (int j = 2; j <= n2 - 1; j++)
                for (int i = 2; i <= n1 - 1; i++) {
                    double v;
                    if (stage == 1) {
                        v = IDX3(Fo, n1, n2, i, j, k) + dt * c[0] * FI_(R[0], n1, n2, i, j, k);
                    } else if (stage == 2) {
                        v = IDX3(Fo, n1, n2, i, j, k) +
                            dt * (c[0] * FI_(R[0], n1, n2, i, j, k) + c[1] * FI_(R[1], n1, n2, i, j, k));
                    } else {
                        v = IDX3(Fo, n1, n2, i, j, k) +
                            dt * (c[0] * FI_(R[0], n1, n2, i, j, k) + c[1] * FI_(R[1], n1, n2, i, j, k) +
                                  c[2] * FI_(R[2], n1, n2, i, j, k));
                    }
                    IDX3(F, n1, n2, i, j, k) = v;
                }
    }
}

/* boundary_conditions.f90:263-289 (id 1: x faces, else x centres), one rank */
static void periodic_x(double *F, int n1, int n2, int n3, int faces) {
    for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++) {
            if (faces) {
                IDX3(F, n1, n2, 1, j, k) = IDX3(F, n1, n2, n1 - 1, j, k);
                IDX3(F, n1, n2, n1, j, k) = IDX3(F, n1, n2, 2, j, k);
            } else {
                IDX3(F, n1, n2, 1, j, k) = IDX3(F, n1, n2, n1 - 2, j, k);
                IDX3(F, n1, n2, n1 - 1, j, k) = IDX3(F, n1, n2, 2, j, k);
                IDX3(F, n1, n2, n1, j, k) = IDX3(F, n1, n2, 3, j, k);
            }
        }
}

/* boundary_conditions.f90:580-602 (id 3: z faces, else z centres), one rank */
static void periodic_z(double *F, int n1, int n2, int n3, int faces) {
    for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
            if (faces) {
                IDX3(F, n1, n2, i, j, 1) = IDX3(F, n1, n2, i, j, n3 - 1);
                IDX3(F, n1, n2, i, j, n3) = IDX3(F, n1, n2, i, j, 2);
            } else {
                IDX3(F, n1, n2, i, j, 1) = IDX3(F, n1, n2, i, j, n3 - 2);
                IDX3(F, n1, n2, i, j, n3 - 1) = IDX3(F, n1, n2, i, j, 2);
                IDX3(F, n1, n2, i, j, n3) = IDX3(F, n1, n2, i, j, 3);
            }
        }
}

/* boundary_conditions.f90:737-758: z walls (one rank): W faces no-penetration,
   U and V centres no-slip by ghost mirroring */
static void dirichlet_z(double *F, int n1, int n2, int n3, int faces) {
    for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
            if (faces) {
                IDX3(F, n1, n2, i, j, 1) = 0.0;
                IDX3(F, n1, n2, i, j, n3) = 0.0;
            } else {
                IDX3(F, n1, n2, i, j, 1) = -IDX3(F, n1, n2, i, j, 2);
                IDX3(F, n1, n2, i, j, n3) = -IDX3(F, n1, n2, i, j, n3 - 1);
            }
        }
}

/* boundary_conditions.f90:762-781: periodic y (one rank); faces: V (id 1,
   1 <- ny-1, ny <- 2), else U and W centres (1 <- nyg-2, nyg-1 <- 2, nyg <- 3) */
static void periodic_y(double *F, int n1, int n2, int n3, int faces) {
    for (int k = 1; k <= n3; k++)
        for (int i = 1; i <= n1; i++) {
            if (faces) {
                IDX3(F, n1, n2, i, 1, k) = IDX3(F, n1, n2, i, n2 - 1, k);
                IDX3(F, n1, n2, i, n2, k) = IDX3(F, n1, n2, i, 2, k);
            } else {
                IDX3(F, n1, n2, i, 1, k) = IDX3(F, n1, n2, i, n2 - 2, k);
                IDX3(F, n1, n2, i, n2 - 1, k) = IDX3(F, n1, n2, i, 2, k);
                IDX3(F, n1, n2, i, n2, k) = IDX3(F, n1, n2, i, 3, k);
            }
        }
}

/* boundary_conditions.f90:784-804: Robin (slip-length) walls; alpha(n1,2,n3)
   holds the lower (1) and upper (2) wall coefficients */
static void robin_y(const ora_plan *p, double *F, int n1, int n2, int n3, const double *alpha, int faces) {
    for (int k = 1; k <= n3; k++)
        for (int i = 1; i <= n1; i++) {
            const double a1 = IDX3(alpha, n1, 2, i, 1, k), a2 = IDX3(alpha, n1, 2, i, 2, k);
            if (faces) {
                const double d1 = p->y[1] - p->y[0], d2 = p->y[p->ny - 1] - p->y[p->ny - 2];
                IDX3(F, n1, n2, i, 1, k) = a1 * IDX3(F, n1, n2, i, 2, k) / d1 / (a1 / d1 + 1.0);
                IDX3(F, n1, n2, i, n2, k) = a2 * IDX3(F, n1, n2, i, n2 - 1, k) / d2 / (a2 / d2 + 1.0);
            } else {
                const double g1 = p->yg[1] - p->yg[0], g2 = p->yg[p->nyg - 1] - p->yg[p->nyg - 2];
                IDX3(F, n1, n2, i, 1, k) = (2.0 * a1 / g1 - 1.0) * IDX3(F, n1, n2, i, 2, k) / (2.0 * a1 / g1 + 1.0);
                IDX3(F, n1, n2, i, n2, k) =
                    (2.0 * a2 / g2 - 1.0) * IDX3(F, n1, n2, i, n2 - 1, k) / (2.0 * a2 / g2 + 1.0);
            }
        }
}

void ora_apply_bc(const ora_plan *p, double *U, double *V, double *W, const double *alpha_x, const double *alpha_y,
                  const double *alpha_z) {
    /* boundary_conditions.f90:186-260 for x_bc_type = 0: periodic x, periodic
       z (or the z walls, :215-223), then Robin y (y_bc_type = 1) or periodic y
       (y_bc_type = 0, :229-233; alphas unused), then the x wrap and, for a
       periodic z, the z wrap again (:247-258) */
    for (int pass = 0; pass < 2; pass++) {
        periodic_x(U, p->nx, p->nyg, p->nzg, 1);
        periodic_x(V, p->nxg, p->ny, p->nzg, 0);
        periodic_x(W, p->nxg, p->nyg, p->nz, 0);
        if (!p->zbc) {
            periodic_z(U, p->nx, p->nyg, p->nzg, 0);
            periodic_z(V, p->nxg, p->ny, p->nzg, 0);
            periodic_z(W, p->nxg, p->nyg, p->nz, 1);
        } else if (pass == 0) {
            /* :218-222 z walls (the second pass has no z part, :253-257) */
            dirichlet_z(U, p->nx, p->nyg, p->nzg, 0);
            dirichlet_z(V, p->nxg, p->ny, p->nzg, 0);
            dirichlet_z(W, p->nxg, p->nyg, p->nz, 1);
        }
        if (pass == 0 && !p->ybc) {
            periodic_y(U, p->nx, p->nyg, p->nzg, 0);
            periodic_y(V, p->nxg, p->ny, p->nzg, 1);
            periodic_y(W, p->nxg, p->nyg, p->nz, 0);
        } else if (pass == 0) {
            robin_y(p, U, p->nx, p->nyg, p->nzg, alpha_x, 0);
            robin_y(p, V, p->nxg, p->ny, p->nzg, alpha_y, 1);
            robin_y(p, W, p->nxg, p->nyg, p->nz, alpha_z, 0);
        }
    }
}
