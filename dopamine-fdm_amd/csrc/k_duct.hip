// The 4-wall duct (x periodic, y walls, z walls; projection.f90:172-218,
// set-up initialization.f90:544-618).  z has no FFT: the z-wall operator Dzz
// (Dyy's construction on z, zg) is diagonalised once on the host through the
// symmetric S = W^(1/2) Dzz W^(-1/2) = Qz diag(lambda_z) Qz^T (W = the z cell
// widths), and after the x transform every x mode's (y, z) plane is taken to
// the z eigenbasis, solved along y per eigenmode and taken back:
//   k_zeig_mfma    C(j, n) = sum_k A(j, k) B(k, n) for every x mode: A the
//                  complex plane, B real: Qf = W^(1/2) Qz forward,
//                  Qi = Qz^T W^(-1/2) back (a batched complex-by-real GEMM
//                  on the FP64 matrix cores)
//   k_duct_lu      inverse pivots of each (kx, m) system Dyy + kxx + lambda_z(m)
//                  (the pin where kx = 0 on the zero eigenvalue)
//   k_duct_ysolve  elimination and back substitution along y, one lane per
//                  (kx, m), in place
// Spectrum layout: spec[(kx*nym + j)*nzp + k] (the x kernels' layout with no
// z split), and the eigenmode planes the same with m in place of k.
//
// The spanwise walls with a periodic y (projection.f90:143-150, :219-245,
// :311-315) use the same z eigenbasis: between the two GEMMs each x mode's
// columns are transformed along y (the nyp = nym-1 independent rows) and
// divided by kxx + kyy + lambda_z(m) (k_yfft_solve).  The reference solves
// the kx = ky = 0 line along z with its pin D(2) *= 3/2, which fixes that
// line's additive constant; k_zwall00 solves exactly that pinned system
// (Thomas along z) and hands its eigenbasis image to k_yfft_solve, so phi
// carries the reference's constant.
#include "fft_lds.h"

namespace dfp {

constexpr int kGemmThreads = 256;
constexpr int kGemmTJ = 64;  // rows (j) per tile
constexpr int kGemmTN = 64;  // columns (n) per tile
constexpr int kGemmKC = 32;  // k per LDS stage

// C = A B for every x mode on the matrix cores: v_mfma_f64_16x16x4f64 (D += A B on
// 16x16x4 FP64 tiles; lane l holds A(l%16, l/16), B(l/16, l%16) and
// D(l/16 + 4v, l%16) in its v-th accumulator, as measured on gfx950).  The
// complex rows of A are taken as pairs of real rows (re, im), so the real
// GEMM has M = 2 nym.  Within a 16-row tile, real row i holds part
// (i/4) % 2 of complex row (i%4) + 4 (i/8): lane l's accumulators v = 0, 1
// are then the re and im parts of complex row l/16 and v = 2, 3 those of
// row l/16 + 4, each pair stored as one 16-byte value.  A block's tile is 64
// complex rows (128 real) x 64 columns; wave w computes real rows
// 32w..32w+31 (two tile rows) x all 64 columns (four tile columns).  Rows
// j < nrow of each x mode's plane (nym rows apart) are transformed: a
// periodic y leaves out its duplicate row nym-1.
typedef double v4d __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kGemmThreads) k_zeig_mfma(const double2* __restrict__ A,
                                                           const double* __restrict__ B, double2* __restrict__ C,
                                                           int nym, int nrow, int K, int N) {
    __shared__ double As[kGemmKC][2 * kGemmTJ + 2];  // real rows, padded
    __shared__ double Bs[kGemmKC][kGemmTN];
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int n0 = blockIdx.x * kGemmTN, j0 = blockIdx.y * kGemmTJ;
    const long plane = (long)blockIdx.z * nym;
    const auto a = gptr(A) + plane * K;
    const auto b = gptr(B);
    constexpr int QA = kGemmTJ * kGemmKC / kGemmThreads, QB = kGemmKC * kGemmTN / kGemmThreads;
    // the next stage's tile values, in flight while the current stage multiplies
    double2 pa[QA];
    double pb[QB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = t + q * kGemmThreads;
            const int jj = e / kGemmKC, kk = e - jj * kGemmKC;
            const int j = j0 + jj, k = k0 + kk;
            pa[q] = (j < nrow && k < K) ? gld<double2>(a + (long)j * K + k) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = t + q * kGemmThreads;
            const int kk = e / kGemmTN, nn = e - kk * kGemmTN;
            const int k = k0 + kk, n = n0 + nn;
            pb[q] = (k < K && n < N) ? b[(long)k * N + n] : 0.0;
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = t + q * kGemmThreads;
            const int jj = e / kGemmKC, kk = e - jj * kGemmKC;
            const int cr = jj & 7;                                     // complex row within its tile
            const int i0 = 16 * (jj >> 3) + (cr & 3) + 8 * (cr >> 2);  // its re row; im is 4 rows on
            As[kk][i0] = pa[q].x;
            As[kk][i0 + 4] = pa[q].y;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = t + q * kGemmThreads;
            const int kk = e / kGemmTN, nn = e - kk * kGemmTN;
            Bs[kk][nn] = pb[q];
        }
    };
    v4d acc[2][4];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = v4d{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int k0 = 0; k0 < K; k0 += kGemmKC) {
        stage();
        __syncthreads();
        if (k0 + kGemmKC < K) fetch(k0 + kGemmKC);
#pragma unroll
        for (int ks = 0; ks < kGemmKC / 4; ++ks) {
            const int kk = 4 * ks + lk;
            double av[2], bv[4];
#pragma unroll
            for (int r = 0; r < 2; ++r) av[r] = As[kk][32 * w + 16 * r + li];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][16 * c + li];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r], bv[c], acc[r][c], 0, 0, 0);
        }
        __syncthreads();
    }
    const auto cp = gptr(C) + plane * N;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        // complex rows j0 + 16w + 8r + lk (v = 0, 1) and + 4 (v = 2, 3)
        const int j = j0 + 16 * w + 8 * r + lk;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int n = n0 + 16 * c + li;
            if (n >= N) continue;
            if (j < nrow) gst<double2>(cp + (long)j * N + n, make_double2(acc[r][c][0], acc[r][c][1]));
            if (j + 4 < nrow) gst<double2>(cp + (long)(j + 4) * N + n, make_double2(acc[r][c][2], acc[r][c][3]));
        }
    }
}

// inverse pivots lu[(kx*nym + j)*nzp + m] = 1/D'(j) of the elimination
// (zgtsv's no-interchange branch, as the channel's table), one lane per mode
__global__ void __launch_bounds__(256) k_duct_lu(double* __restrict__ lu, Ops o, const double* __restrict__ lam_z,
                                                 int nxh, int nym, int nzp, int pin_m) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nxh * nzp) return;
    const int kx = (int)(t / nzp);
    const int m = (int)(t - (long)kx * nzp);
    const double lam = o.kxx[kx] + lam_z[m];
    double dp = o.dd[0] + lam;
    if (kx == 0 && m == pin_m) dp = 3.0 / 2.0 * dp;
    double* col = lu + (long)kx * nym * nzp + m;
    col[0] = 1.0 / dp;
    for (int j = 1; j < nym; ++j) {
        const double mult = o.dl[j - 1] / dp;
        dp = (o.dd[j] + lam) - mult * o.du[j - 1];
        col[(long)j * nzp] = 1.0 / dp;
    }
}

constexpr int kDuctAhead = 4;  // rows read ahead by each lane of the y sweeps

// y(j) = b(j) - DL(j-1)/D'(j-1) y(j-1), then x(j) = (y(j) - DU(j) x(j+1)) / D'(j),
// in place on the mode's column (stride nzp); rows and pivots are read
// kDuctAhead rows ahead into a ring of registers (clamped, unconditional)
__global__ void __launch_bounds__(256) k_duct_ysolve(double2* __restrict__ S, const double* __restrict__ lu, Ops o,
                                                     int nxh, int nym, int nzp) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nxh * nzp) return;
    const int kx = (int)(t / nzp);
    const int m = (int)(t - (long)kx * nzp);
    const long base = (long)kx * nym * nzp + m;
    const auto col = gptr(S) + base;
    const auto piv = gptr(lu) + base;
    const auto dl = gptr(o.dl);
    const auto du = gptr(o.du);
    double2 rb[kDuctAhead];
    double rp[kDuctAhead];
    // forward elimination
#pragma unroll
    for (int q = 0; q < kDuctAhead; ++q) {
        const int jq = min(q, nym - 1);
        rb[q] = gld<double2>(col + (long)jq * nzp);
        rp[q] = piv[(long)jq * nzp];
    }
    double yx = 0.0, yy = 0.0, ipp = 0.0;
    for (int j0 = 0; j0 < nym; j0 += kDuctAhead) {
#pragma unroll
        for (int q = 0; q < kDuctAhead; ++q) {
            const int j = j0 + q;
            if (j < nym) {
                const double2 bj = rb[q];
                const double pj = rp[q];
                const int jn = min(j + kDuctAhead, nym - 1);
                rb[q] = gld<double2>(col + (long)jn * nzp);
                rp[q] = piv[(long)jn * nzp];
                const double mult = j > 0 ? dl[j - 1] * ipp : 0.0;
                yx = bj.x - mult * yx;
                yy = bj.y - mult * yy;
                gst<double2>(col + (long)j * nzp, make_double2(yx, yy));
                ipp = pj;
            }
        }
    }
    // back substitution, rows from the top
#pragma unroll
    for (int q = 0; q < kDuctAhead; ++q) {
        const int jq = max(nym - 1 - q, 0);
        rb[q] = gld<double2>(col + (long)jq * nzp);
        rp[q] = piv[(long)jq * nzp];
    }
    double xx = 0.0, xy = 0.0;
    for (int t0 = 0; t0 < nym; t0 += kDuctAhead) {
#pragma unroll
        for (int q = 0; q < kDuctAhead; ++q) {
            const int tt = t0 + q;
            if (tt < nym) {
                const int j = nym - 1 - tt;
                const double2 yj = rb[q];
                const double ip = rp[q];
                const int jn = max(j - kDuctAhead, 0);
                rb[q] = gld<double2>(col + (long)jn * nzp);
                rp[q] = piv[(long)jn * nzp];
                const double dr = j < nym - 1 ? du[j] : 0.0;
                xx = (yj.x - dr * xx) * ip;
                xy = (yj.y - dr * xy) * ip;
                gst<double2>(col + (long)j * nzp, make_double2(xx, xy));
            }
        }
    }
}

// The pinned kx = ky = 0 line, one workgroup of 1024 lanes: r(k) = sum over
// the nyp rows of the kx = 0 plane (its ky = 0 y-mode; four row groups per
// column, eight independent partial sums each, combined in a fixed order);
// p = (Dzz with D(2) *= 3/2)^-1 r by zgtsv's no-interchange elimination and
// back substitution (one lane, with the plan's multipliers DL/D' and inverse
// pivots of that fixed system staged in LDS first); z00(m) = sum_k p(k)
// Qf(k, m) (eight partial sums per lane).  LDS: r, the multipliers, the
// inverse pivots and the row-group partials.
constexpr int kZw00Threads = 1024;
constexpr int kZw00Groups = 4;
__global__ void __launch_bounds__(kZw00Threads) k_zwall00(const double2* __restrict__ S, const double* __restrict__ qf,
                                                          const double* __restrict__ zmul, const double* __restrict__ zip,
                                                          const double* __restrict__ dzu, double2* __restrict__ z00,
                                                          int nym, int nzm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* r = reinterpret_cast<double2*>(smem_raw);     // nzm
    double* ip = reinterpret_cast<double*>(r + nzm);       // nzm: inverse pivots
    double* mu = ip + nzm;                                 // nzm: multipliers (mu[0] unused)
    double* uu = mu + nzm;                                 // nzm: DU(k), uu[nzm-1] = 0
    double2* part = reinterpret_cast<double2*>(uu + nzm);  // kZw00Groups x (kZw00Threads / kZw00Groups)
    for (int k = threadIdx.x; k < nzm; k += kZw00Threads) {
        ip[k] = zip[k];
        mu[k] = k > 0 ? zmul[k] : 0.0;
        uu[k] = k < nzm - 1 ? dzu[k] : 0.0;
    }
    constexpr int KB = kZw00Threads / kZw00Groups;       // columns per pass
    const int nyp = nym - 1;
    const int tk = threadIdx.x % KB, grp = threadIdx.x / KB;
    for (int k0 = 0; k0 < nzm; k0 += KB) {
        const int k = k0 + tk;
        double2 acc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = make_double2(0.0, 0.0);
        if (k < nzm) {
            int j = grp;
            for (; j + 7 * kZw00Groups < nyp; j += 8 * kZw00Groups) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const double2 v = S[(long)(j + u * kZw00Groups) * nzm + k];
                    acc[u].x += v.x;
                    acc[u].y += v.y;
                }
            }
            for (; j < nyp; j += kZw00Groups) {
                const double2 v = S[(long)j * nzm + k];
                acc[0].x += v.x;
                acc[0].y += v.y;
            }
        }
        double2 sgrp = acc[0];
#pragma unroll
        for (int u = 1; u < 8; ++u) {
            sgrp.x += acc[u].x;
            sgrp.y += acc[u].y;
        }
        part[grp * KB + tk] = sgrp;
        __syncthreads();
        if (grp == 0 && k < nzm) {
            double2 s = part[tk];
            for (int q = 1; q < kZw00Groups; ++q) {
                s.x += part[q * KB + tk].x;
                s.y += part[q * KB + tk].y;
            }
            r[k] = s;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double yx = r[0].x, yy = r[0].y;
        for (int k = 1; k < nzm; ++k) {
            yx = r[k].x - mu[k] * yx;
            yy = r[k].y - mu[k] * yy;
            r[k] = make_double2(yx, yy);
        }
        double xx = 0.0, xy = 0.0;
        for (int k = nzm - 1; k >= 0; --k) {
            xx = (r[k].x - uu[k] * xx) * ip[k];
            xy = (r[k].y - uu[k] * xy) * ip[k];
            r[k] = make_double2(xx, xy);
        }
    }
    __syncthreads();
    for (int m = threadIdx.x; m < nzm; m += kZw00Threads) {
        double2 acc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = make_double2(0.0, 0.0);
        int k = 0;
        for (; k + 7 < nzm; k += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const double q = qf[(long)(k + u) * nzm + m];
                acc[u].x = fma(r[k + u].x, q, acc[u].x);
                acc[u].y = fma(r[k + u].y, q, acc[u].y);
            }
        }
        for (; k < nzm; ++k) {
            const double q = qf[(long)k * nzm + m];
            acc[0].x = fma(r[k].x, q, acc[0].x);
            acc[0].y = fma(r[k].y, q, acc[0].y);
        }
        double2 s = acc[0];
#pragma unroll
        for (int u = 1; u < 8; ++u) {
            s.x += acc[u].x;
            s.y += acc[u].y;
        }
        z00[m] = s;
    }
}

// One workgroup per (kx, group of G eigenmode columns): the nyp x G tile of
// the independent rows in LDS, the y FFT down the columns, each mode divided
// by kxx + kyy + lambda_z(m) (the kx = ky = 0 row taken from z00), the
// inverse FFT scaled by 1/nyp, stored back.  The duplicate row nym-1
// (projection.f90:350-354) is not carried: the inverse x kernel reads row 0
// for it.  Columns past nzm are zero-filled and not stored.
constexpr int kYsThreads = 256;
__global__ void __launch_bounds__(kYsThreads) k_yfft_solve(double2* __restrict__ S, const double2* __restrict__ twy_g,
                                                           FftDesc fy, int ntwy, const double* __restrict__ kxx,
                                                           const double* __restrict__ kyy,
                                                           const double* __restrict__ lam_z,
                                                           const double2* __restrict__ z00, int nym, int nzm, int G,
                                                           FastDiv div_g) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double2* A = reinterpret_cast<double2*>(smem_raw);  // nym x G
    double2* twy = A + (long)nym * G;                   // ntwy
    const int ngrp = (nzm + G - 1) / G;
    const int kx = blockIdx.x / ngrp;
    const int m0 = (blockIdx.x - kx * ngrp) * G;
    const int nyp = nym - 1;
    const auto plane = gptr(S) + (long)kx * nym * nzm;
    for (int e = threadIdx.x; e < nyp * G; e += kYsThreads) {
        const int j = (int)div_g.div((uint32_t)e), c = e - j * G;
        const int m = m0 + c;
        A[e] = m < nzm ? gld<double2>(plane + (long)j * nzm + m) : make_double2(0.0, 0.0);
    }
    for (int e = threadIdx.x; e < ntwy; e += kYsThreads) twy[e] = twy_g[e];
    __syncthreads();
    lds_fft_cols<-1, 4>(A, 0, G, G, fy, twy, div_g);
    const double lx = kxx[kx];
    const double inv_ny = 1.0 / (double)nyp;
    for (int e = threadIdx.x; e < nyp * G; e += kYsThreads) {
        const int ky = (int)div_g.div((uint32_t)e), c = e - ky * G;
        const int m = m0 + c;
        if (m >= nzm) continue;
        if (kx == 0 && ky == 0) {
            // the pinned line's y-mode (the inverse FFT's 1/nyp applies to it too)
            const double2 v = z00[m];
            A[e] = make_double2(v.x * inv_ny, v.y * inv_ny);
        } else {
            const double2 v = A[e];
            const double s = inv_ny / (lx + kyy[ky] + lam_z[m]);
            A[e] = make_double2(v.x * s, v.y * s);
        }
    }
    __syncthreads();
    lds_fft_cols<1, 4>(A, 0, G, G, fy, twy, div_g);
    for (int e = threadIdx.x; e < nyp * G; e += kYsThreads) {
        const int j = (int)div_g.div((uint32_t)e), c = e - j * G;
        const int m = m0 + c;
        if (m < nzm) gst<double2>(plane + (long)j * nzm + m, A[e]);
    }
}

size_t yfft_solve_lds_bytes(int nym, int G, int ntwy) { return sizeof(double2) * ((size_t)nym * G + ntwy); }

// columns per k_yfft_solve workgroup: the largest power of two up to 16
// whose columns one LDS transform call holds
int yfft_solve_columns(const FftDesc& fy, int nym, size_t lds_limit) {
    for (int G = 16; G >= 1; G /= 2)
        if ((long)G * fy.n <= inplace_capacity(fy, kYsThreads, 4) &&
            yfft_solve_lds_bytes(nym, G, twiddle_count(fy)) <= lds_limit)
            return G;
    return 0;
}

hipError_t launch_duct_lu(dfp_plan* p) {
    const Geo& g = p->geo;
    const long n = (long)g.nxh * g.nzp;
    hipLaunchKernelGGL(k_duct_lu, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, p->d_lu, p->ops, p->duct_lam,
                       g.nxh, g.nym, g.nzp, p->pin_m);
    return hipGetLastError();
}

static hipError_t launch_zeig(dfp_plan* p, const double2* A, const double* B, double2* C, const char* name) {
    const Geo& g = p->geo;
    const int nrow = g.ybc == 0 ? g.nym - 1 : g.nym;  // the duplicate row of a periodic y is not transformed
    const dim3 grid(ceil_div(g.nzp, kGemmTN), ceil_div(nrow, kGemmTJ), g.nxh);
    const int slot = prof_begin(p, name);
    hipLaunchKernelGGL(k_zeig_mfma, grid, dim3(kGemmThreads), 0, p->stream, A, B, C, g.nym, nrow, g.nzp, g.nzp);
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_duct_solve(dfp_plan* p) {
    const Geo& g = p->geo;
    hipError_t e;
    if (g.ybc == 0) {
        // spanwise walls with a periodic y: the pinned kx = ky = 0 line, the
        // forward z eigenbasis GEMM, the y transforms and divisions, the inverse GEMM
        // the pinned line (one workgroup) runs on the plan's second stream
        // beside the forward GEMM: both only read the spectrum, and the y
        // solve joins them.  With per-kernel profiling on, both stay in order
        // on the plan's stream.
        const size_t l00 = (sizeof(double2) + 3 * sizeof(double)) * (size_t)g.nzp + sizeof(double2) * kZw00Threads;
        const bool fork = !p->prof && p->aux;
        hipStream_t s00 = p->stream;
        if (fork) {
            if ((e = hipEventRecord(p->fork_ev, p->stream)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(p->aux, p->fork_ev, 0)) != hipSuccess) return e;
            s00 = p->aux;
        }
        int slot = prof_begin(p, "zwall00");
        hipLaunchKernelGGL(k_zwall00, dim3(1), dim3(kZw00Threads), l00, s00, p->d_spec, p->duct_qf, p->zw_mult,
                           p->zw_ip, p->dzz_u, p->d_z00, g.nym, g.nzp);
        prof_end(p, slot);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (fork && (e = hipEventRecord(p->join_ev, p->aux)) != hipSuccess) return e;
        if ((e = launch_zeig(p, p->d_spec, p->duct_qf, p->d_spec2, "zeig_fwd")) != hipSuccess) return e;
        if (fork && (e = hipStreamWaitEvent(p->stream, p->join_ev, 0)) != hipSuccess) return e;
        const int G = p->gy, ngrp = ceil_div(g.nzp, G), ntwy = twiddle_count(p->fy);
        slot = prof_begin(p, "yfft_solve");
        hipLaunchKernelGGL(k_yfft_solve, dim3(g.nxh * ngrp), dim3(kYsThreads), yfft_solve_lds_bytes(g.nym, G, ntwy),
                           p->stream, p->d_spec2, p->tw_y, p->fy, ntwy, p->ops.kxx, p->ops.kyy, p->duct_lam, p->d_z00,
                           g.nym, g.nzp, G, FastDiv((uint32_t)G));
        prof_end(p, slot);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        return launch_zeig(p, p->d_spec2, p->duct_qi, p->d_spec, "zeig_inv");
    }
    e = launch_zeig(p, p->d_spec, p->duct_qf, p->d_spec2, "zeig_fwd");
    if (e != hipSuccess) return e;
    const long n = (long)g.nxh * g.nzp;
    const int slot = prof_begin(p, "duct_ysolve");
    hipLaunchKernelGGL(k_duct_ysolve, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, p->d_spec2, p->d_lu, p->ops,
                       g.nxh, g.nym, g.nzp);
    prof_end(p, slot);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_zeig(p, p->d_spec2, p->duct_qi, p->d_spec, "zeig_inv");
}

}  // namespace dfp
