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
// elimination order of the spectral kernels.  Rows are staged through LDS in
// chunks by the whole block; one thread runs each recurrence over a chunk
// from LDS (distinct arrays, so its loads run ahead of its stores).
constexpr int kM00Chunk = 1024;

__device__ __forceinline__ void mode00_body(const MPar& m) {
    __shared__ double s_in[kM00Chunk], s_mul[kM00Chunk], s_cu[kM00Chunk], s_out[kM00Chunk];
    const auto rowpart = gptr(m.rowpart);
    const auto lu = gptr(m.lu);
    const auto dl = gptr(m.o.dl);
    const auto du = gptr(m.o.du);
    const auto x = gptr(m.x);
    const int nym = m.nym, nparts = m.nparts;
    const long nzs = m.nzs;
    // forward: y(0) = mean(0); y(j) = mean(j) - DL(j-1)/D'(j-1) * y(j-1)
    double yp = 0.0;
    for (int c0 = 0; c0 < nym; c0 += kM00Chunk) {
        const int len = min(kM00Chunk, nym - c0);
        for (int t = threadIdx.x; t < len; t += blockDim.x) {
            const int j = c0 + t;
            // the parts in order, eight loads in flight at a time
            const auto rp = rowpart + (long)j * nparts;
            double s = 0.0;
            int q = 0;
            for (; q + 8 <= nparts; q += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = rp[q + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += v[u];
            }
            for (; q < nparts; ++q) s += rp[q];
            s_in[t] = s / m.count;
            s_mul[t] = j > 0 ? dl[j - 1] * lu[(j - 1) * nzs] : 0.0;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // eight rows' operands are read together, so one LDS round trip
            // serves eight steps of the recurrence
            for (int t0 = 0; t0 < len; t0 += 8) {
                double a[8], mu[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int t = min(t0 + q, len - 1);
                    a[q] = s_in[t];
                    mu[q] = s_mul[t];
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (t0 + q < len) {
                        yp = a[q] - mu[q] * yp;
                        s_out[t0 + q] = yp;
                    }
                }
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < len; t += blockDim.x) x[c0 + t] = s_out[t];
        __syncthreads();
    }
    // backward: x(nym-1) = y/D'; x(j) = (y(j) - DU(j) x(j+1)) / D'(j)
    double xn = 0.0;
    for (int c1 = nym; c1 > 0; c1 -= kM00Chunk) {
        const int c0 = max(0, c1 - kM00Chunk), len = c1 - c0;
        for (int t = threadIdx.x; t < len; t += blockDim.x) {
            const int j = c0 + t;
            s_in[t] = x[j];
            s_mul[t] = lu[j * nzs];
            s_cu[t] = j < nym - 1 ? du[j] : 0.0;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int t1 = len - 1; t1 >= 0; t1 -= 8) {
                double a[8], ip[8], cu[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int t = max(t1 - q, 0);
                    a[q] = s_in[t];
                    ip[q] = s_mul[t];
                    cu[q] = s_cu[t];
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int t = t1 - q;
                    if (t >= 0) {
                        xn = c0 + t == nym - 1 ? a[q] * ip[q] : (a[q] - cu[q] * xn) * ip[q];
                        s_out[t] = xn;
                    }
                }
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < len; t += blockDim.x) x[c0 + t] = s_out[t];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_mode00(MPar m) { mode00_body(m); }

// cases c = blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ void __launch_bounds__(256) kb_mode00(const MPar* __restrict__ cs, const int* __restrict__ idx,
                                                 const int* __restrict__ pnc) {
    const int nc = *pnc;
    for (int c = blockIdx.x; c < nc; c += gridDim.x) {
        mode00_body(cs[idx[c]]);
        __syncthreads();
    }
}

MPar make_mpar(const dfp_plan* p) {
    MPar m{};
    const Geo& g = p->geo;
    m.rowpart = p->d_rowpart;
    m.lu = p->d_lu;
    m.o = p->ops;
    m.x = p->d_pbar;
    m.count = (double)g.nxp * (double)g.nzp;
    m.nparts = (g.nzp / p->zr) / p->tk_x;
    m.nym = g.nym;
    m.nzs = g.nzp / 2 + 1;
    return m;
}

hipError_t launch_mode00(dfp_plan* p, hipStream_t s) {
    const int slot = prof_begin(p, "mode00");
    hipLaunchKernelGGL(k_mode00, dim3(1), dim3(256), 0, s, make_mpar(p));
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_mode00_batch(hipStream_t s, const MPar* cs, const int* idx, const int* nc, int grid) {
    hipLaunchKernelGGL(kb_mode00, dim3(grid), dim3(256), 0, s, cs, idx, nc);
    return hipGetLastError();
}

// P = phi/(dt*c) with every ghost (projection.f90:361-380, :388-438 on one
// rank): each P element copies the periodic / wall-clamped interior cell.
// Used by batches (a single plan writes P in its last inverse transform).
__device__ __forceinline__ void save_p_body(const SPar& sp, int bx) {
    const Geo& g = sp.g;
    const long t = (long)bx * 256 + threadIdx.x;
    const long n = (long)g.nxg * g.nyg * g.nzg;
    if (t >= n) return;
    const uint32_t tj = sp.div_x.div((uint32_t)t);
    const int i = (int)(t - (long)tj * g.nxg) + 1;  // reference 1-based indices
    const uint32_t tk = sp.div_y.div(tj);
    const int j = (int)(tj - tk * (uint32_t)g.nyg) + 1;
    const int k = (int)tk + 1;
    const int ii = (i - 2 + g.nxp) % g.nxp;
    const int jj = j < 2 ? 0 : (j > g.nym + 1 ? g.nym - 1 : j - 2);
    const int kk = (k - 2 + g.nzp) % g.nzp;
    gptr(sp.P)[t] = gptr(sp.rhs)[ii + jj * (long)(g.nxg - 1) + kk * (long)(g.nxg - 1) * g.nym] / sp.dtc;
}

__global__ void __launch_bounds__(256) kb_save_p(const SPar* __restrict__ cs, const int* __restrict__ off,
                                                 const int* __restrict__ idx, const int* __restrict__ pnc) {
    const int nc = *pnc;
    const int total = off[nc];
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        int lo = 0, hi = nc;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= w) lo = mid;
            else hi = mid;
        }
        save_p_body(cs[idx[lo]], w - off[lo]);
    }
}

SPar make_spar(const dfp_plan* p, const double* rhs, double* P, double dtc) {
    SPar sp{};
    const Geo& g = p->geo;
    sp.rhs = rhs;
    sp.P = P;
    sp.g = g;
    sp.dtc = dtc;
    sp.div_x = FastDiv(g.nxg);
    sp.div_y = FastDiv(g.nyg);
    sp.nblocks = ceil_div((long)g.nxg * g.nyg * g.nzg, 256);
    return sp;
}

hipError_t launch_save_p_batch(hipStream_t s, const SPar* cs, const int* off, const int* idx, const int* nc,
                               int grid) {
    hipLaunchKernelGGL(kb_save_p, dim3(resident_grid(kb_save_p, grid, 256, 0)), dim3(256), 0, s, cs, off, idx, nc);
    return hipGetLastError();
}

}  // namespace dfp
