// FP64 pieces of the FP32+refine form of the projection.
//
// The pinned operator the spectral solve inverts splits exactly into the x-z
// mean of each y row (its (0,0) Fourier mode) and the rest:
//   phi = phibar(j) + phi',   mean_xz(L_pin phi) = Dyy_pin phibar,
// because the periodic x and z stencils have zero mean.  The FP32 form
// solves the mean part in FP64 here (k_mode00) and the rest with the FP32
// spectral chain (FP32 FFTs, FP64 elimination on FP32-stored modes), then
// refines:  r = b - L_pin phi (FP64, k_div_rfft_x mode 2),
//           phi += FP32 solve of r' + FP64 solve of rbar  (k_irfft_x ACCUM).
// The velocity correction is the FP64 kernel of the plain path.
#include "dfp_internal.h"

namespace dfp {

// Row means of the field from the x kernel's per-block partial sums, then the
// (0,0)-mode solve Dyy_pin x = mean with the inverse pivots the plan holds for
// kx = kz = 0 (lu[j*nzs], pin of projection.f90:292 included) and the
// elimination order of the spectral kernels.  One block; the solve itself is
// one thread (nym sequential steps).
__global__ void __launch_bounds__(256) k_mode00(const double* __restrict__ rowpart, int nparts, double count,
                                                const double* __restrict__ lu, Ops o, int nym, int nzs,
                                                double* __restrict__ x) {
    for (int j = threadIdx.x; j < nym; j += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < nparts; ++q) s += rowpart[(long)j * nparts + q];
        x[j] = s / count;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double yp = x[0];
    for (int j = 1; j < nym; ++j) {
        const double mult = o.dl[j - 1] * lu[(long)(j - 1) * nzs];
        yp = x[j] - mult * yp;
        x[j] = yp;
    }
    double xn = x[nym - 1] * lu[(long)(nym - 1) * nzs];
    x[nym - 1] = xn;
    for (int j = nym - 2; j >= 0; --j) {
        xn = (x[j] - o.du[j] * xn) * lu[(long)j * nzs];
        x[j] = xn;
    }
}

hipError_t launch_mode00(dfp_plan* p, const double* rowpart, int nparts, double* x) {
    const Geo& g = p->geo;
    const int slot = prof_begin(p, "mode00");
    hipLaunchKernelGGL(k_mode00, dim3(1), dim3(256), 0, p->stream, rowpart, nparts, (double)g.nxp * (double)g.nzp,
                       p->d_lu, p->ops, g.nym, g.nzp / 2 + 1, x);
    prof_end(p, slot);
    return hipGetLastError();
}

// P = phi/(dt*c) with every ghost (projection.f90:361-380, :388-438 on one
// rank): each P element copies the periodic / wall-clamped interior cell.
__global__ void __launch_bounds__(256) k_save_p(const double* __restrict__ rhs, double* __restrict__ P, Geo g,
                                                double dtc, FastDiv div_x, FastDiv div_y) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.nxg * g.nyg * g.nzg;
    if (t >= n) return;
    const uint32_t tj = div_x.div((uint32_t)t);
    const int i = (int)(t - (long)tj * g.nxg) + 1;  // reference 1-based indices
    const uint32_t tk = div_y.div(tj);
    const int j = (int)(tj - tk * (uint32_t)g.nyg) + 1;
    const int k = (int)tk + 1;
    const int ii = (i - 2 + g.nxp) % g.nxp;
    const int jj = j < 2 ? 0 : (j > g.nym + 1 ? g.nym - 1 : j - 2);
    const int kk = (k - 2 + g.nzp) % g.nzp;
    P[t] = rhs[ii + jj * (long)(g.nxg - 1) + kk * (long)(g.nxg - 1) * g.nym] / dtc;
}

hipError_t launch_save_p(dfp_plan* p, const double* rhs, double* P, double dtc) {
    const Geo& g = p->geo;
    const long n = (long)g.nxg * g.nyg * g.nzg;
    const int slot = prof_begin(p, "save_p");
    hipLaunchKernelGGL(k_save_p, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, rhs, P, g, dtc, FastDiv(g.nxg),
                       FastDiv(g.nyg));
    prof_end(p, slot);
    return hipGetLastError();
}

}  // namespace dfp
