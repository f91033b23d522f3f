// Physical-space stencils of the projection (FP64, HBM-bound):
//   k_divergence   rhs_p = div(U,V,W)            projection.f90:37-63
//   k_correct      U,V,W -= grad(rhs_p)          projection.f90:551-604
//   k_max_div      max |div(U,V,W)|              projection.f90:607-639
//   k_outflow_phib outlet backflow face values   projection.f90:65-96
//   k_synth        seeded synthetic channel fields (benchmark/test inputs)
// Threads run over the (x,y) plane of one z index, x fastest (the contiguous
// index of every field), blockIdx.y is the z index; consecutive lanes read
// consecutive addresses of one row.  The
// stencil arithmetic is kept operation-for-operation as the reference writes
// it (no FMA contraction), so these kernels reproduce the CPU path bit for bit.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "dfp_internal.h"

namespace dfp {

// rhs_p(2:nxg,2:nyg-1,2:nzg) element of reference cell (i,j,k)
__device__ __forceinline__ long ridx(const Geo& g, int i, int j, int k) {
    return (long)(i - 2) + (long)(j - 2) * (g.nxg - 1) + (long)(k - 2) * (long)(g.nxg - 1) * g.nym;
}

__global__ void __launch_bounds__(256) k_divergence(const double* __restrict__ U, const double* __restrict__ V,
                                                    const double* __restrict__ W, double* __restrict__ rhs, Geo g,
                                                    FastDiv div_ni) {
#pragma clang fp contract(off)
    const int ni = g.nxg - 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * g.nym) return;
    const int j = 2 + (int)div_ni.div((uint32_t)t);                 // 2..nyg-1
    const int i = 2 + t - (j - 2) * ni;       // 2..nxg-1
    const int k = 2 + blockIdx.y;             // 2..nzg-1
    const long iu = (i - 1) + (long)(j - 1) * g.nx + (long)(k - 1) * g.nx * g.nyg;
    const long iv = (i - 1) + (long)(j - 1) * g.nxg + (long)(k - 1) * (long)g.nxg * g.ny;
    const long iw = (i - 1) + (long)(j - 1) * g.nxg + (long)(k - 1) * (long)g.nxg * g.nyg;
    const double du = (U[iu] - U[iu - 1]) * g.inv_dx;
    const double dv = (V[iv] - V[iv - g.nxg]) * g.inv_dyf[j - 2];
    const double dw = (W[iw] - W[iw - (long)g.nxg * g.nyg]) * g.inv_dzf[k - 2];
    rhs[ridx(g, i, j, k)] = du + dv + dw;
}

hipError_t launch_divergence(dfp_plan* p, const double* U, const double* V, const double* W, double* rhs) {
    const Geo& g = p->geo;
    dim3 grid(ceil_div((long)(g.nxg - 2) * g.nym, 256), g.nzg - 2);
    const int slot = prof_begin(p, "divergence");
    hipLaunchKernelGGL(k_divergence, grid, dim3(256), 0, p->stream, U, V, W, rhs, g, FastDiv(g.nxg - 2));
    prof_end(p, slot);
    return hipGetLastError();
}

// U(i,j,k) -= (phi(i+1)-phi(i))/dx          i=2..nx-1,  j=2..nyg-1, k=2..nzg-1
//   inflow/outflow x, outlet face i = nx: U += 2 (phi(nx) - phib)/dx (phib = 0
//   without the backflow term)
// V(i,j,k) -= (phi(j+1)-phi(j))/dyg(j)      i=2..nxg-1, j=2..ny-1,  k=2..nzg-1
// W(i,j,k) -= (phi(k+1)-phi(k))/dzg(k)      i=2..nxg-1, j=2..nyg-1, k=2..nz-1
// Each thread takes two consecutive x points (i, i+1) of one row j, with
// 16-byte loads and stores (8-byte aligned where the row length is odd), and
// walks kCorrectK consecutive z planes, so phi(k+1) read for W at one step is
// phi(k) of the next: phi is read from memory once.
constexpr int kCorrectK = 4;

__device__ __forceinline__ void correct_body(const CPar& c, int bx, int bz) {
#pragma clang fp contract(off)
    const Geo& g = c.g;
    // U, V, W and phi are distinct arrays: the loads of later planes may be
    // issued ahead of the stores of earlier ones
    auto* __restrict__ U = gptr(c.U);
    auto* __restrict__ V = gptr(c.V);
    auto* __restrict__ W = gptr(c.W);
    const auto* __restrict__ rhs = gptr(c.rhs);
    const auto inv_dyg = gptr(g.inv_dyg);
    const auto inv_dzg = gptr(g.inv_dzg);
    const int np = (g.nxg - 1) / 2;  // pairs per row: ceil((nxg-2)/2)
    const int t = bx * 256 + threadIdx.x;
    if (t >= np * g.nym) return;
    const int j = 2 + (int)c.div_ni.div((uint32_t)t);
    const int i = 2 + 2 * (t - (j - 2) * np);  // this lane's points: i, i+1
    const int k0 = 2 + bz * kCorrectK;
    const int k1 = min(k0 + kCorrectK - 1, g.nzg - 1);
    const long pr_k = (long)(g.nxg - 1) * g.nym;
    const long pu_k = (long)g.nx * g.nyg, pv_k = (long)g.nxg * g.ny, pw_k = (long)g.nxg * g.nyg;
    // per point e = 0, 1 (i+e): U on i+e <= nx-1, V and W on i+e <= nxg-1
    const bool u0 = i <= g.nx - 1, u1 = i + 1 <= g.nx - 1;
    const bool in1 = i + 1 <= g.nxg - 1;
    const bool do_v = j <= g.ny - 1;
    // inflow/outflow x: the outlet face U(nx) takes the face value of phi,
    // zero or phib (projection.f90:578-589)
    const bool f0 = g.xbc == 1 && i == g.nx, f1 = g.xbc == 1 && i + 1 == g.nx;
    const double idyg = do_v ? inv_dyg[j - 2] : 0.0;
    const long r = ridx(g, i, j, k0);
    const long iu = (i - 1) + (long)(j - 1) * g.nx + (long)(k0 - 1) * pu_k;
    const long iv = (i - 1) + (long)(j - 1) * g.nxg + (long)(k0 - 1) * pv_k;
    const long iw = (i - 1) + (long)(j - 1) * g.nxg + (long)(k0 - 1) * pw_k;
    const int nk = k1 - k0 + 1;  // planes of this block, 1..kCorrectK
    // Every load of the block's planes is issued before any store (the
    // arrays are distinct): phi(i..i+1) on planes k0..k1+1 (rhs_p holds the
    // planes up to nzg and the points up to nxg, so i+1 <= nxg is in its
    // row), phi(i+2) (clamped to nxg), phi(j+1) (row j on the last row), u,
    // v, w.  A pair (i, i+1) of U, V or W reads one point past the row when
    // i+1 is beyond it: the next row's first element, which exists because
    // j and k stay below the arrays' last row and plane (U: j <= nyg-1 of
    // 1..nyg, k <= nzg-1 of 1..nzg; V: k <= nzg-1, its last row j = ny only
    // reached with do_v false and i+1 <= nxg; W: j <= nyg-1).  Plane
    // indices are clamped to the block's last plane, so every address is in
    // range; values of unused planes and points are never stored.
    dbl2u pc[kCorrectK + 1], py[kCorrectK], uo[kCorrectK], vo[kCorrectK], wo[kCorrectK];
    double px2[kCorrectK];
    const int i2 = min(i + 2, g.nxg) - i;  // offset of phi(i+2) in the row
#pragma unroll
    for (int q = 0; q <= kCorrectK; ++q) pc[q] = ld2(rhs + r + min(q, nk) * pr_k);
    const long oy = do_v ? (long)(g.nxg - 1) : 0;
#pragma unroll
    for (int q = 0; q < kCorrectK; ++q) {
        const int qq = min(q, nk - 1);
        px2[q] = rhs[r + qq * pr_k + i2];
        py[q] = ld2(rhs + r + qq * pr_k + oy);
        uo[q] = ld2(U + iu + qq * pu_k);
        vo[q] = ld2(V + iv + qq * pv_k);
        wo[q] = ld2(W + iw + qq * pw_k);
    }
    const auto phib = gptr(c.phib);  // null without the backflow term
#pragma unroll
    for (int q = 0; q < kCorrectK; ++q) {
        if (q >= nk) break;
        const int k = k0 + q;
        // U: interior points, or the outlet face
        double ua = 0.0, ub = 0.0;
        if (u0) {
            ua = uo[q].x - (pc[q].y - pc[q].x) * g.inv_dx;
        } else if (f0) {
            ua = phib ? uo[q].x + 2.0 * (pc[q].x - phib[(j - 2) + (long)(k - 2) * g.nym]) * g.inv_dx
                      : uo[q].x + 2.0 * pc[q].x * g.inv_dx;
        }
        if (u1) {
            ub = uo[q].y - (px2[q] - pc[q].y) * g.inv_dx;
        } else if (f1) {
            ub = phib ? uo[q].y + 2.0 * (pc[q].y - phib[(j - 2) + (long)(k - 2) * g.nym]) * g.inv_dx
                      : uo[q].y + 2.0 * pc[q].y * g.inv_dx;
        }
        st2(U + iu + q * pu_k, ua, ub, u0 || f0, u1 || f1);
        if (do_v)
            st2(V + iv + q * pv_k, vo[q].x - (py[q].x - pc[q].x) * idyg, vo[q].y - (py[q].y - pc[q].y) * idyg, true,
                in1);
        // phi(k+1) is used for W only, for k <= nz-1
        if (k <= g.nz - 1) {
            const double dz = inv_dzg[k - 2];
            st2(W + iw + q * pw_k, wo[q].x - (pc[q + 1].x - pc[q].x) * dz, wo[q].y - (pc[q + 1].y - pc[q].y) * dz,
                true, in1);
        }
    }
}

__global__ void __launch_bounds__(256) k_correct(CPar c) {
    const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int bz = L / gridDim.x;
    correct_body(c, L - bz * gridDim.x, bz);
}

__global__ void __launch_bounds__(256) kb_correct(const CPar* __restrict__ cs, const int* __restrict__ off,
                                                  const int* __restrict__ idx, const int* __restrict__ pnc) {
    const int nc = *pnc;
    const int total = off[nc];
    const int me = xcd_remap(blockIdx.x, gridDim.x);
    for (int base = 0; base < total; base += gridDim.x) {
        const int w = base + me;
        if (w >= total) break;
        int lo = 0, hi = nc;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= w) lo = mid;
            else hi = mid;
        }
        const CPar& c = cs[idx[lo]];
        const int li = w - off[lo];
        const int bz = li / c.nbx;
        correct_body(c, li - bz * c.nbx, bz);
    }
}

CPar make_cpar(const dfp_plan* p, double* U, double* V, double* W, const double* rhs) {
    CPar c{};
    c.U = U; c.V = V; c.W = W; c.rhs = rhs;
    c.phib = nullptr;
    c.g = p->geo;
    c.div_ni = FastDiv((p->geo.nxg - 1) / 2);  // pairs of x points per row
    c.nbx = ceil_div((long)((p->geo.nxg - 1) / 2) * p->geo.nym, 256);
    return c;
}

int correct_blocks_z(const Geo& g) { return ceil_div(g.nzg - 2, kCorrectK); }

hipError_t launch_correct(dfp_plan* p, double* U, double* V, double* W, const double* rhs, const double* phib) {
    CPar c = make_cpar(p, U, V, W, rhs);
    c.phib = phib;
    dim3 grid(c.nbx, correct_blocks_z(p->geo));
    const int slot = prof_begin(p, "correct");
    hipLaunchKernelGGL(k_correct, grid, dim3(256), 0, p->stream, c);
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_correct_batch(hipStream_t s, const CPar* cs, const int* off, const int* idx, const int* nc,
                                int grid) {
    hipLaunchKernelGGL(kb_correct, dim3(resident_grid(kb_correct, grid, 256, 0)), dim3(256), 0, s, cs, off, idx, nc);
    return hipGetLastError();
}

// Outlet backflow term of the inflow/outflow x, set_outflow_backflow_pressure
// (projection.f90:73-94) on one rank (which owns the outlet): one thread per
// outlet face (j,k), j = 2..nyg-1, k = 2..nzg-1,
//   phib(j,k) = dtc*(-beta/2*(un^2+vt^2+wt^2)*th - Pb(j,k)),
//   th = (1 - tanh(un/uref))/2, un = U(nx,j,k), vt, wt the face averages of
//   V and W, and, with rhs given, rhs_p(nx,j,k) -= 2*phib(j,k)/dx^2.
// Pb and phib are (nyg-2) x (nzg-2) arrays, (j,k) at (j-2) + (k-2)*(nyg-2).
__global__ void __launch_bounds__(256) k_outflow_phib(const double* __restrict__ U, const double* __restrict__ V,
                                                      const double* __restrict__ W, const double* __restrict__ Pb,
                                                      double* __restrict__ phib, double* __restrict__ rhs, Geo g,
                                                      double beta, double uref, double dtc, double inv_dx2) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= g.nym * (g.nzp + 1)) return;
    const int k = 2 + t / g.nym;
    const int j = 2 + t - (k - 2) * g.nym;
    const int nx = g.nx;
    const double un = U[(nx - 1) + (long)(j - 1) * g.nx + (long)(k - 1) * g.nx * g.nyg];
    const long iv = (nx - 1) + (long)(j - 1) * g.nxg + (long)(k - 1) * (long)g.nxg * g.ny;
    const long iw = (nx - 1) + (long)(j - 1) * g.nxg + (long)(k - 1) * (long)g.nxg * g.nyg;
    const double vt = 0.5 * (V[iv] + V[iv - g.nxg]);
    const double wt = 0.5 * (W[iw] + W[iw - (long)g.nxg * g.nyg]);
    const double th = 0.5 * (1.0 - tanh(un / uref));
    const double ph = dtc * (-0.5 * beta * (un * un + vt * vt + wt * wt) * th - Pb[t]);
    phib[t] = ph;
    if (rhs) {
        const long r = ridx(g, nx, j, k);
        rhs[r] = rhs[r] - 2.0 * ph * inv_dx2;
    }
}

hipError_t launch_outflow_phib(dfp_plan* p, const double* U, const double* V, const double* W, const double* Pb,
                               double* phib, double* rhs, double dtc) {
    const Geo& g = p->geo;
    const double uref = 0.05 * std::max(p->outflow_uc, 1e-2);
    const double inv_dx2 = 1.0 / std::pow(p->dx, 2.0);
    const int n = g.nym * (g.nzp + 1);
    const int slot = prof_begin(p, "outflow_phib");
    hipLaunchKernelGGL(k_outflow_phib, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, U, V, W, Pb, phib, rhs, g,
                       p->beta, uref, dtc, inv_dx2);
    prof_end(p, slot);
    return hipGetLastError();
}

// |div u| of cell (i,j,k) (reference indices) with the face spacings of
// check_divergence (projection.f90:631-633)
__device__ __forceinline__ double abs_div(const double* __restrict__ U, const double* __restrict__ V,
                                          const double* __restrict__ W, const Geo& g, int i, int j, int k) {
#pragma clang fp contract(off)
    const long iu = (i - 1) + (long)(j - 1) * g.nx + (long)(k - 1) * g.nx * g.nyg;
    const long iv = (i - 1) + (long)(j - 1) * g.nxg + (long)(k - 1) * (long)g.nxg * g.ny;
    const long iw = (i - 1) + (long)(j - 1) * g.nxg + (long)(k - 1) * (long)g.nxg * g.nyg;
    const double dxi = g.x ? (g.x[i - 1] - g.x[i - 2]) : 1.0 / g.inv_dx;
    return fabs((U[iu] - U[iu - 1]) / dxi + (V[iv] - V[iv - g.nxg]) / (g.y[j - 1] - g.y[j - 2]) +
                (W[iw] - W[iw - (long)g.nxg * g.nyg]) / (g.z[k - 1] - g.z[k - 2]));
}

// max over the block of one value per thread; thread 0 receives it
__device__ __forceinline__ double block_max(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// Block-level max of |div| over one (j,k) row segment; partials[block] holds it.
__global__ void __launch_bounds__(256) k_max_div(const double* __restrict__ U, const double* __restrict__ V,
                                                 const double* __restrict__ W, double* __restrict__ partials, Geo g) {
    __shared__ double red[256];
    const int ni = g.nxg - 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + t / ni;
    const int i = 2 + t - (j - 2) * ni;
    const int k = 2 + blockIdx.y;
    const double v = t < ni * g.nym ? abs_div(U, V, W, g, i, j, k) : 0.0;
    const double m = block_max(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x + (long)gridDim.x * blockIdx.y] = m;
}

// The same maximum with a bounded number of blocks: block b covers the
// cells c = b*256 + tid, + gridDim.x*256, ... of the interior in (i, j, k)
// order; partials[b] receives its maximum.
__global__ void __launch_bounds__(256) k_max_div_strided(const double* __restrict__ U, const double* __restrict__ V,
                                                         const double* __restrict__ W, double* __restrict__ partials,
                                                         Geo g) {
    __shared__ double red[256];
    const long ni = g.nxg - 2, nj = g.nym;
    const long total = ni * nj * (g.nzg - 2);
    double v = 0.0;
    for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < total; c += (long)gridDim.x * 256) {
        const long q = c / ni;  // interior row (j, k), j fastest
        const long kq = q / nj;
        const int i = 2 + (int)(c - q * ni);
        const int j = 2 + (int)(q - kq * nj);
        const int k = 2 + (int)kq;
        v = fmax(v, abs_div(U, V, W, g, i, j, k));
    }
    const double m = block_max(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

__global__ void __launch_bounds__(256) k_reduce_max(const double* __restrict__ partials, long n,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    double v = 0.0;
    for (long e = threadIdx.x; e < n; e += blockDim.x) v = fmax(v, partials[e]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_max_div(dfp_plan* p, const double* U, const double* V, const double* W, double* out_dev) {
    const Geo& g = p->geo;
    dim3 grid(ceil_div((long)(g.nxg - 2) * g.nym, 256), g.nzg - 2);
    const long nb = (long)grid.x * grid.y;
    if (nb > p->red_len) return hipErrorInvalidValue;
    const int slot = prof_begin(p, "max_div");
    hipLaunchKernelGGL(k_max_div, grid, dim3(256), 0, p->stream, U, V, W, p->d_red, g);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, p->stream, p->d_red, nb, out_dev);
    prof_end(p, slot);
    return hipGetLastError();
}

// Field guard of the reference's monitor (monitor.f90:32-38): a velocity value
// in U(2:nx-1,2:nyg-1,2:nzg-1), V(2:nxg-1,2:ny-1,2:nzg-1) or
// W(2:nxg-1,2:nyg-1,2:nz-1) that is NaN or exceeds 1e8 in magnitude.  Block b
// strides over the three ranges and writes 1 (found) or 0 to partials[b].
__global__ void __launch_bounds__(256) k_guard_finite(const double* __restrict__ U, const double* __restrict__ V,
                                                      const double* __restrict__ W, double* __restrict__ partials,
                                                      Geo g) {
    __shared__ double red[256];
    const long nk = g.nzg - 2;
    const long su = (long)(g.nx - 2) * g.nym * nk;             // U interior
    const long sv = (long)(g.nxg - 2) * (g.ny - 2) * nk;       // V interior faces
    const long sw = (long)(g.nxg - 2) * g.nym * (g.nz - 2);    // W interior faces
    double bad = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < su + sv + sw; e += (long)gridDim.x * 256) {
        const double* F;
        long r, n1, n2, s1, s2;
        if (e < su) {
            F = U; r = e; n1 = g.nx - 2; n2 = g.nym; s1 = g.nx; s2 = (long)g.nx * g.nyg;
        } else if (e < su + sv) {
            F = V; r = e - su; n1 = g.nxg - 2; n2 = g.ny - 2; s1 = g.nxg; s2 = (long)g.nxg * g.ny;
        } else {
            F = W; r = e - su - sv; n1 = g.nxg - 2; n2 = g.nym; s1 = g.nxg; s2 = (long)g.nxg * g.nyg;
        }
        const long q = r / n1, k = q / n2;
        const long i = r - q * n1, j = q - k * n2;
        const double x = F[(i + 1) + (j + 1) * s1 + (k + 1) * s2];
        if (x != x || fabs(x) > 1e8) bad = 1.0;
    }
    const double m = block_max(bad, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

hipError_t launch_guard_finite(dfp_plan* p, const double* U, const double* V, const double* W, double* out_dev) {
    const int blocks = std::max(1, std::min(1024, p->red_len));
    const int slot = prof_begin(p, "guard_finite");
    hipLaunchKernelGGL(k_guard_finite, dim3(blocks), dim3(256), 0, p->stream, U, V, W, p->d_red, p->geo);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, p->stream, p->d_red, (long)blocks, out_dev);
    prof_end(p, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free max-divergence (dfp_max_divergence_noalloc).  Its partial
// maxima and result live in a fixed array of device memory instead of the
// plan's workspace: kNoAllocParts partials + the result, 8200 bytes, held
// by every GPU the library runs on until the process ends.  Host code never
// uses the array's name as an address: at a device's first call, with that
// device current, hipGetSymbolAddress gives the device's own copy, kept in
// scratch_dev[device].  Every call on a device shares its copy; calls that
// overlap on one device race (each call overwrites all partials it reads,
// so nothing has to start at zero).
// ---------------------------------------------------------------------------
constexpr int kNoAllocParts = 1024;
constexpr int kNoAllocDoubles = kNoAllocParts + 1;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

static std::mutex scratch_mu;
static double* scratch_dev[kMaxDevices] = {};

// address of this device's copy of scratch; the device must be current
static hipError_t scratch_ptr(int device, double** out) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_dev[device] = static_cast<double*>(ptr);
    }
    *out = scratch_dev[device];
    return hipSuccess;
}

}  // namespace dfp

extern "C" int dfp_max_divergence_noalloc(dfp_plan* p, const double* U, const double* V, const double* W,
                                          double* out, int flags) {
    using namespace dfp;
    if (!p) return DFP_E_ARG;
    auto fail = [&](int code, const std::string& msg) {
        p->last_error = msg;
        return code;
    };
    if (!U || !V || !W || !out) return fail(DFP_E_ARG, "U, V, W and out are required");
    if (flags != DFP_FIELDS_ON_DEVICE)
        return fail(DFP_E_ARG, "dfp_max_divergence_noalloc takes device-resident fields (DFP_FIELDS_ON_DEVICE) only");
    if (p->device < 0 || p->device >= kMaxDevices) return fail(DFP_E_UNSUPPORTED, "device index beyond 63");
    const Geo& g = p->geo;
    const long cells = (long)(g.nxg - 2) * g.nym * (g.nzg - 2);
    // at most kNoAllocParts blocks, each striding over its share of the
    // cells: any grid a plan takes fits the fixed array
    const long blocks = std::max<long>(1, std::min<long>(kNoAllocParts, (cells + 255) / 256));
    if (blocks + 1 > kNoAllocDoubles)
        return fail(DFP_E_UNSUPPORTED, "max-divergence partials exceed the fixed workspace of 1025 doubles");
    hipError_t e = hipSetDevice(p->device);
    double* ws = nullptr;
    if (e == hipSuccess) e = scratch_ptr(p->device, &ws);
    if (e != hipSuccess) return fail(DFP_E_HIP, std::string("workspace address: ") + hipGetErrorString(e));
    const int slot = prof_begin(p, "max_div_noalloc");
    hipLaunchKernelGGL(k_max_div_strided, dim3((unsigned)blocks), dim3(256), 0, p->stream, U, V, W, ws, g);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, p->stream, ws, blocks, ws + kNoAllocParts);
    prof_end(p, slot);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, ws + kNoAllocParts, sizeof(double), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return fail(DFP_E_HIP, std::string("max divergence: ") + hipGetErrorString(e));
    return DFP_OK;
}

namespace dfp {

// ---------------------------------------------------------------------------
// synthetic fields: SplitMix64 counter hash -> Box-Muller normal deviates
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double gauss(unsigned long long key, unsigned long long e) {
    const unsigned long long G = 0x9E3779B97F4A7C15ull;
    const unsigned long long a = mix64(key + (2ull * e + 1ull) * G);
    const unsigned long long b = mix64(key + (2ull * e + 2ull) * G);
    const double u1 = (double)((a >> 11) + 1ull) * 0x1.0p-53;  // (0,1]
    const double u2 = (double)(b >> 11) * 0x1.0p-53;           // [0,1)
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

// periodic source index of a cell-centred ghost (1-based): 1 -> n-2, n-1 -> 2, n -> 3
__device__ __forceinline__ int per_cell(int i, int n) { return i == 1 ? n - 2 : (i == n - 1 ? 2 : (i == n ? 3 : i)); }
// periodic source index of a face ghost: 1 -> n-1, n -> 2
__device__ __forceinline__ int per_face(int i, int n) { return i == 1 ? n - 1 : (i == n ? 2 : i); }

// field f: 0 = U, 1 = V, 2 = W.  One thread per array element (ghosts included).
__global__ void __launch_bounds__(256) k_synth(double* __restrict__ F, int f, Geo g, unsigned long long key,
                                               double sigma, double yc0, double h) {
    const int n1 = (f == 0) ? g.nx : g.nxg;
    const int n2 = (f == 1) ? g.ny : g.nyg;
    const int n3 = (f == 2) ? g.nz : g.nzg;
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > n1) return;
    // the inflow/outflow x has no periodic ghosts: every x index is its own source
    const int is = g.xbc == 1 ? i : ((f == 0) ? per_face(i, n1) : per_cell(i, n1));
    const int ks = (f == 2) ? per_face(k, n3) : per_cell(k, n3);
    double sgn = 1.0;
    int js = j;
    if (f == 1) {
        if (j == 1 || j == n2) {
            F[(i - 1) + (long)(j - 1) * n1 + (long)(k - 1) * n1 * n2] = 0.0;
            return;
        }
    } else if (j == 1) {
        js = 2; sgn = -1.0;
    } else if (j == n2) {
        js = n2 - 1; sgn = -1.0;
    }
    const unsigned long long e = (unsigned long long)(is - 1) + (unsigned long long)(js - 1) * n1 +
                                 (unsigned long long)(ks - 1) * n1 * n2;
    double v = sigma * gauss(key, e);
    if (f == 0) {
        const double yc = (g.yg[js - 1] - yc0) / h;
        v = (1.0 - yc * yc) + v;
    }
    F[(i - 1) + (long)(j - 1) * n1 + (long)(k - 1) * n1 * n2] = sgn * v;
}

static unsigned long long mix64_host(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

hipError_t launch_synth(dfp_plan* p, double* U, double* V, double* W, uint64_t seed, double sigma) {
    const Geo& g = p->geo;
    const double yc0 = 0.5 * (p->hy.front() + p->hy.back());
    const double h = 0.5 * (p->hy.back() - p->hy.front());
    double* F[3] = {U, V, W};
    const int slot = prof_begin(p, "synth");
    for (int f = 0; f < 3; ++f) {
        const int n1 = (f == 0) ? g.nx : g.nxg;
        const int n2 = (f == 1) ? g.ny : g.nyg;
        const int n3 = (f == 2) ? g.nz : g.nzg;
        const unsigned long long key = mix64_host(seed + (unsigned long long)(f + 1) * 0x9E3779B97F4A7C15ull);
        hipLaunchKernelGGL(k_synth, dim3(ceil_div(n1, 256), n2, n3), dim3(256), 0, p->stream, F[f], f, g, key, sigma,
                           yc0, h);
    }
    prof_end(p, slot);
    return hipGetLastError();
}

}  // namespace dfp
