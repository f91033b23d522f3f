// The 4-wall duct (x periodic, y walls, z walls; projection.f90:172-218,
// set-up initialization.f90:544-618).  z has no FFT: the z-wall operator Dzz
// (Dyy's construction on z, zg) is diagonalised once on the host through the
// symmetric S = W^(1/2) Dzz W^(-1/2) = Qz diag(lambda_z) Qz^T (W = the z cell
// widths), and after the x transform every x mode's (y, z) plane is taken to
// the z eigenbasis, solved along y per eigenmode and taken back:
//   k_zeig         C(j, n) = sum_k A(j, k) B(k, n) for every x mode: A the
//                  complex plane, B real: Qf = W^(1/2) Qz forward,
//                  Qi = Qz^T W^(-1/2) back (a batched complex-by-real GEMM)
//   k_duct_lu      inverse pivots of each (kx, m) system Dyy + kxx + lambda_z(m)
//                  (the pin where kx = 0 on the zero eigenvalue)
//   k_duct_ysolve  elimination and back substitution along y, one lane per
//                  (kx, m), in place
// Spectrum layout: spec[(kx*nym + j)*nzp + k] (the x kernels' layout with no
// z split), and the eigenmode planes the same with m in place of k.
#include "fft_lds.h"

namespace dfp {

constexpr int kGemmThreads = 256;
constexpr int kGemmTJ = 64;  // rows (j) per tile
constexpr int kGemmTN = 64;  // columns (n) per tile
constexpr int kGemmKC = 16;  // k per LDS stage

// One (n-tile, j-tile, kx) tile of C = A B.  Thread (tx, ty) = (t % 16,
// t / 16) accumulates rows ty + 16 r and columns tx + 16 c (r, c < 4): in
// the inner loop a wave reads 4 distinct A values (broadcast) and 16
// consecutive B values, and the stores of a row run over 16 consecutive
// columns.  A is staged transposed (As[k][j], rows padded by one value so
// the transposing writes fall in distinct banks).
__global__ void __launch_bounds__(kGemmThreads) k_zeig(const double2* __restrict__ A, const double* __restrict__ B,
                                                      double2* __restrict__ C, int nym, int K, int N) {
    __shared__ double2 As[kGemmKC][kGemmTJ + 1];
    __shared__ double Bs[kGemmKC][kGemmTN];
    const int t = threadIdx.x;
    const int tx = t % 16, ty = t / 16;
    const int n0 = blockIdx.x * kGemmTN, j0 = blockIdx.y * kGemmTJ;
    const long plane = (long)blockIdx.z * nym;
    const auto a = gptr(A) + plane * K;
    const auto b = gptr(B);
    double2 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < K; k0 += kGemmKC) {
        // A tile: 64 rows x 16 k, consecutive lanes along k (256-byte runs)
#pragma unroll
        for (int q = 0; q < kGemmTJ * kGemmKC / kGemmThreads; ++q) {
            const int e = t + q * kGemmThreads;
            const int jj = e / kGemmKC, kk = e - jj * kGemmKC;
            const int j = j0 + jj, k = k0 + kk;
            As[kk][jj] = (j < nym && k < K) ? gld<double2>(a + (long)j * K + k) : make_double2(0.0, 0.0);
        }
        // B tile: 16 k x 64 columns, consecutive lanes along n
#pragma unroll
        for (int q = 0; q < kGemmKC * kGemmTN / kGemmThreads; ++q) {
            const int e = t + q * kGemmThreads;
            const int kk = e / kGemmTN, nn = e - kk * kGemmTN;
            const int k = k0 + kk, n = n0 + nn;
            Bs[kk][nn] = (k < K && n < N) ? b[(long)k * N + n] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGemmKC; ++kk) {
            double2 av[4];
            double bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) av[r] = As[kk][ty + 16 * r];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][tx + 16 * c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc[r][c].x = fma(av[r].x, bv[c], acc[r][c].x);
                    acc[r][c].y = fma(av[r].y, bv[c], acc[r][c].y);
                }
        }
        __syncthreads();
    }
    const auto cp = gptr(C) + plane * N;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = j0 + ty + 16 * r;
        if (j >= nym) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int n = n0 + tx + 16 * c;
            if (n < N) gst<double2>(cp + (long)j * N + n, acc[r][c]);
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

hipError_t launch_duct_lu(dfp_plan* p) {
    const Geo& g = p->geo;
    const long n = (long)g.nxh * g.nzp;
    hipLaunchKernelGGL(k_duct_lu, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, p->d_lu, p->ops, p->duct_lam,
                       g.nxh, g.nym, g.nzp, p->pin_m);
    return hipGetLastError();
}

static hipError_t launch_zeig(dfp_plan* p, const double2* A, const double* B, double2* C, const char* name) {
    const Geo& g = p->geo;
    const dim3 grid(ceil_div(g.nzp, kGemmTN), ceil_div(g.nym, kGemmTJ), g.nxh);
    const int slot = prof_begin(p, name);
    hipLaunchKernelGGL(k_zeig, grid, dim3(kGemmThreads), 0, p->stream, A, B, C, g.nym, g.nzp, g.nzp);
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_duct_solve(dfp_plan* p) {
    const Geo& g = p->geo;
    hipError_t e = launch_zeig(p, p->d_spec, p->duct_qf, p->d_spec2, "zeig_fwd");
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
