// Spectral stages of the projection, x direction, for the x inflow/outflow
// case (x_bc_type = 1): a DCT-IV along x (FFTW's REDFT11,
// y_k = 2 sum_n x_n cos(pi (2n+1)(2k+1) / 4N), projection.f90:126-129, and
// the same transform for the inverse with 1/(2 nxp nzp), :333-339) in place
// of the periodic real FFT.  The z split, the spectrum layout and the z
// kernels are those of the periodic case (k_xfft_impl.h), with all N = nxp
// DCT modes kept as x planes: spec[((kx*R + u)*nym + j)*nz2 + m], kx < nxp.
// The DCT-IV values are real; the first z stage makes them complex.
//
// DCT-IV of length N = 2M through a length-M complex FFT (both directions):
//   v_n = x_{2n} + i x_{N-1-2n}               (n < M; packed in place: real
//                                              slot i of x_i even, N-i odd)
//   c_k = 2 e^{-i pi k/N} FFT_M[v_n e^{-i pi (4n+1)/(4N)}]_k
//   y_{2k} = Re c_k,  y_{N-1-2k} = -Im c_k.
// The pre-twiddles e^{-i pi (4n+1)/(4N)} and post-twiddles 2 e^{-i pi k/N}
// follow the x stage twiddles in the plan's twiddle array, then the first z
// stage's: one contiguous LDS copy per workgroup.
#pragma once

#include "k_xfft_impl.h"

namespace dfp {

// real slot of x_i in the packed line of v_n = x_{2n} + i x_{N-1-2n}
__device__ __forceinline__ int dct_slot(int i, int N) { return (i & 1) ? N - i : i; }

// Divergence (MODE 0, with the outlet backflow term when a.phib is given:
// projection.f90:55-57, :91) or rhs_p as given (MODE 1), DCT-IV in x, first
// radix-R z stage; FP64.
template <int MODE, int R>
__device__ __forceinline__ void xfwd_dct_body(const XPar<double2>& q, int bx, int j, unsigned char* smem_raw) {
    using C = double2;
    const Geo& g = q.g;
    const XArgs& a = q.a;
    const auto aU = gptr(a.U);
    const auto aV = gptr(a.V);
    const auto aW = gptr(a.W);
    const auto arhs = gptr(a.rhs);
    const auto aphib = gptr(a.phib);
    const auto inv_dyf = gptr(g.inv_dyf);
    const auto inv_dzf = gptr(g.inv_dzf);
    const auto spec = gptr(q.spec);
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int M = q.fd.n;
    const int nxp = g.nxp;
    const int nz2 = g.nzp / R;
    const int k0 = bx * q.T;
    const int nt = q.T;  // T divides nz2
    const int nl = R * nt;
    const int ntw = q.ntw;
    C* A = smem;                 // nl lines of M packed points
    C* twl = smem + nl * M;      // FFT stage twiddles
    C* prl = twl + ntw;          // pre-twiddles, M
    C* pol = prl + M;            // post-twiddles, M
    C* tzl = pol + M;            // first z stage twiddles, nz2*(R-1)
    double* Ar = reinterpret_cast<double*>(A);
    for (int e = threadIdx.x; e < ntw + 2 * M + nz2 * (R - 1); e += kXThreads) twl[e] = gld<C>(gptr(q.tw) + e);

    const long pu_j = (long)g.nx, pu_k = (long)g.nx * g.nyg;
    const long pv_j = (long)g.nxg, pv_k = (long)g.nxg * g.ny;
    const long pw_j = (long)g.nxg, pw_k = (long)g.nxg * g.nyg;
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const int total = nl * nxp;
    const double inv_dy = inv_dyf[j];
    // two pairs of neighbouring points (i, i+1), i even, per thread per batch
    // (16-byte field loads); every load is issued before any is used
    const int total2 = total >> 1;  // nxp is even
    for (int base = 0; base < total2; base += 2 * kXThreads) {
        double v[2][2];
        int dst[2], di[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e0 = base + (int)threadIdx.x + u * kXThreads;
            const int e = 2 * min(e0, total2 - 1);  // first point of the pair
            const int l = (int)q.div_nxp.div((uint32_t)e);
            const int i = e - l * nxp;
            const int t = (int)q.div_t.div((uint32_t)l);
            const int k = k0 + (l - t * nt) + t * nz2;
            dst[u] = e0 < total2 ? l * nxp : -1;  // line l; the points go to slots dct_slot(i), dct_slot(i+1)
            di[u] = i;
            if (MODE == 1) {
                const dbl2u r2 = ld2(arhs + i + j * pr_j + k * pr_k);
                v[u][0] = r2.x;
                v[u][1] = r2.y;
            } else {
                const long iu = (i + 1) + (j + 1) * pu_j + (k + 1) * pu_k;  // U(i+2,j+2,k+2)
                const long iv = (i + 1) + (j + 1) * pv_j + (k + 1) * pv_k;  // V(i+2,j+2,k+2)
                const long iw = (i + 1) + (j + 1) * pw_j + (k + 1) * pw_k;  // W(i+2,j+2,k+2)
                const dbl2u ua = ld2(aU + iu - 1);  // U(i+1), U(i+2)
                const double ub = aU[iu + 1];       // U(i+3)
                const dbl2u v1 = ld2(aV + iv), v0 = ld2(aV + iv - pv_j);
                const dbl2u w1 = ld2(aW + iw), w0 = ld2(aW + iw - pw_k);
                const double idz = inv_dzf[k];
                // the outlet face value of the backflow term (projection.f90:91), at i+1 = nxp-1
                const bool last = a.phib && i + 1 == nxp - 1;
                const double fb = last ? aphib[j + (long)k * g.nym] : 0.0;
                {
#pragma clang fp contract(off)
                    // projection.f90:55-57
                    v[u][0] = (ua.y - ua.x) * g.inv_dx + (v1.x - v0.x) * inv_dy + (w1.x - w0.x) * idz;
                    double d = (ub - ua.y) * g.inv_dx + (v1.y - v0.y) * inv_dy + (w1.y - w0.y) * idz;
                    if (last) d = d - 2.0 * fb * a.idx2;
                    v[u][1] = d;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (dst[u] >= 0) {
                Ar[dst[u] + dct_slot(di[u], nxp)] = v[u][0];
                Ar[dst[u] + dct_slot(di[u] + 1, nxp)] = v[u][1];
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nl * M; e += kXThreads) {
        const int n = e - (int)q.div_half.div((uint32_t)e) * M;
        A[e] = cmul(A[e], prl[n]);
    }
    __syncthreads();
    lds_fft_inplace<-1, 2, C, true>(A, nl, M, q.fd, twl);

    // post-twiddle: modes 2k and N-1-2k of each line, first z stage, store
    for (int pidx = threadIdx.x; pidx < M * nt; pidx += kXThreads) {
        const int kq = (int)q.div_t.div((uint32_t)pidx);
        const int c = pidx - kq * nt;
        C ae[R], ao[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const C cv = cmul(A[(t * nt + c) * M + kq], pol[kq]);
            ae[t] = make_double2(cv.x, 0.0);
            ao[t] = make_double2(-cv.y, 0.0);
        }
        dft_reg<-1, R>(ae);
        dft_reg<-1, R>(ao);
        const int kp = k0 + c;
        const auto oe = spec + ((long)(2 * kq) * R * g.nym + j) * nz2 + kp;
        const auto oo = spec + ((long)(nxp - 1 - 2 * kq) * R * g.nym + j) * nz2 + kp;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const C tz = u == 0 ? make_double2(1.0, 0.0) : tzl[kp * (R - 1) + (u - 1)];
            gst<C>(oe + (long)u * g.nym * nz2, u == 0 ? ae[0] : cmul(ae[u], tz));
            gst<C>(oo + (long)u * g.nym * nz2, u == 0 ? ao[0] : cmul(ao[u], tz));
        }
    }
}

template <int MODE, int R>
__global__ void __launch_bounds__(kXThreads) k_div_dct_x(XPar<double2> q) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    int bx, j;
    xtile(L, gridDim.x, gridDim.y, q.strip, bx, j);
    xfwd_dct_body<MODE, R>(q, bx, j, smem_raw);
}

// Largest tile of the inverse kernel: nl*M points, two butterfly slots per
// thread in the transform (radix 4 at most): the post-twiddle holds its
// share in registers between a read and a write phase.
constexpr int kDctSlots = 8;

// Last inverse z stage, DCT-IV in x, phi = x/(2 nxp nzp) into rhs_p with the
// z duplicate plane (projection.f90:479-482; no x duplicate, :446-448) and
// optionally P = phi/(dt*c) with its ghosts: Neumann copies at the two x ends
// (projection.f90:413-422), wall copies in y and periodic z.
template <int R>
__device__ __forceinline__ void xinv_dct_body(const XPar<double2>& q, int bx, int j, unsigned char* smem_raw) {
    using C = double2;
    const Geo& g = q.g;
    const auto spec = gptr(q.spec);
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int M = q.fd.n;
    const int nxp = g.nxp;
    const int nz2 = g.nzp / R;
    const int k0 = bx * q.T;
    const int nt = q.T;
    const int nl = R * nt;
    const int ntw = q.ntw;
    C* A = smem;
    C* twl = smem + nl * M;
    C* prl = twl + ntw;
    C* pol = prl + M;
    C* tzl = pol + M;
    double* Ar = reinterpret_cast<double*>(A);
    for (int e = threadIdx.x; e < ntw + 2 * M + nz2 * (R - 1); e += kXThreads) twl[e] = gld<C>(gptr(q.tw) + e);
    __syncthreads();

    // gather the R sub-line values of each (kx, column), undo the first z
    // stage, keep the real part in its packed slot
    const int npairs = nxp * nt;
    constexpr int HP = R <= 4 ? 2 : 1;
    for (int base = 0; base < npairs; base += HP * kXThreads) {
        C av[HP][R];
        int pk[HP];
#pragma unroll
        for (int h = 0; h < HP; ++h) {
            const int p0 = base + (int)threadIdx.x + h * kXThreads;
            const int pidx = min(p0, npairs - 1);
            pk[h] = p0 < npairs ? pidx : -1;
            const int kx = (int)q.div_t.div((uint32_t)pidx);
            const int c = pidx - kx * nt;
            const auto src = spec + ((long)kx * R * g.nym + j) * nz2 + k0 + c;
#pragma unroll
            for (int u = 0; u < R; ++u) av[h][u] = gld<C>(src + (long)u * g.nym * nz2);
        }
#pragma unroll
        for (int h = 0; h < HP; ++h) {
            if (pk[h] < 0) continue;
            const int kx = (int)q.div_t.div((uint32_t)pk[h]);
            const int c = pk[h] - kx * nt;
            const int kp = k0 + c;
#pragma unroll
            for (int u = 1; u < R; ++u) av[h][u] = cmulc(av[h][u], tzl[kp * (R - 1) + (u - 1)]);
            dft_reg<1, R>(av[h]);
            const int slot = dct_slot(kx, nxp);
#pragma unroll
            for (int t = 0; t < R; ++t) Ar[(t * nt + c) * nxp + slot] = av[h][t].x;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nl * M; e += kXThreads) {
        const int n = e - (int)q.div_half.div((uint32_t)e) * M;
        A[e] = cmul(A[e], prl[n]);
    }
    __syncthreads();
    lds_fft_inplace<-1, 2, C, false>(A, nl, M, q.fd, twl);
    // post-twiddle into registers, then the values in natural order
    double xe[kDctSlots], xo[kDctSlots];
#pragma unroll
    for (int s = 0; s < kDctSlots; ++s) {
        const int e = threadIdx.x + s * kXThreads;
        if (e < nl * M) {
            const int n = e - (int)q.div_half.div((uint32_t)e) * M;
            const C cv = cmul(A[e], pol[n]);
            xe[s] = cv.x;
            xo[s] = -cv.y;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kDctSlots; ++s) {
        const int e = threadIdx.x + s * kXThreads;
        if (e < nl * M) {
            const int l = (int)q.div_half.div((uint32_t)e);
            const int n = e - l * M;
            Ar[l * nxp + 2 * n] = xe[s];
            Ar[l * nxp + nxp - 1 - 2 * n] = xo[s];
        }
    }
    __syncthreads();
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const long pp_j = (long)g.nxg, pp_k = (long)g.nxg * g.nyg;
    const auto rr = gptr(q.rhs) + (long)j * pr_j;
    const auto P = gptr(q.P);
    const double scale = q.scale, dtc = q.dtc;
    const int save_p = q.save_p;
    // two neighbouring points (i, i+1), i even, per lane: 16-byte LDS reads and
    // stores of phi and of P's main copy
    for (int e2 = threadIdx.x; e2 < (nl * nxp) >> 1; e2 += kXThreads) {
        const int e = 2 * e2;
        const int l = (int)q.div_nxp.div((uint32_t)e);
        const int i = e - l * nxp;
        const int t = (int)q.div_t.div((uint32_t)l);
        const int k = k0 + (l - t * nt) + t * nz2;
        const C z = A[(l * nxp + i) >> 1];
        const double phi0 = z.x * scale, phi1 = z.y * scale;
        // rhs_p(i+2, j+2, k+2) and the z duplicate (k = nzp -> nzp+2)
        st2(rr + i + k * pr_k, phi0, phi1, true, true);
        if (k == 0) st2(rr + i + (long)g.nzp * pr_k, phi0, phi1, true, true);
        if (save_p) {
            // every copy of these cells in P (reference 1-based indices; 0 = none):
            // x: i+2, 1 if i = 0 (the first point), nxg if i = nxp-1 (the second)
            // y: j+2, 1 if j = 0, nyg if j = nym-1
            // z: k+2, 1 if k = nzp-1, nzg-1 if k = 0, nzg if k = 1 (mod nzp)
            const double pv0 = phi0 / dtc, pv1 = phi1 / dtc;
            const int y1 = j == 0 ? 1 : 0, y2 = j == g.nym - 1 ? g.nyg : 0;
            const int z1 = k == g.nzp - 1 ? 1 : 0, z2 = k == 0 ? g.nzg - 1 : 0, z3 = k == 1 % g.nzp ? g.nzg : 0;
#pragma unroll
            for (int c3 = 0; c3 < 4; ++c3) {
                const int zz = c3 == 0 ? k + 2 : (c3 == 1 ? z1 : (c3 == 2 ? z2 : z3));
                if (zz == 0) continue;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const int yy = b == 0 ? j + 2 : (b == 1 ? y1 : y2);
                    if (yy == 0) continue;
                    const auto Pr = P + (long)(yy - 1) * pp_j + (long)(zz - 1) * pp_k;
                    st2(Pr + (i + 1), pv0, pv1, true, true);  // x = i+2, i+3
                    if (i == 0) Pr[0] = pv0;                  // x = 1
                    if (i + 1 == nxp - 1) Pr[g.nxg - 1] = pv1;  // x = nxg
                }
            }
        }
    }
}

template <int R>
__global__ void __launch_bounds__(kXThreads) k_idct_x(XPar<double2> q) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    int bx, j;
    xtile(L, gridDim.x, gridDim.y, q.strip, bx, j);
    xinv_dct_body<R>(q, bx, j, smem_raw);
}

template <int R>
hipError_t launch_xdct_fwd_r(hipStream_t s, int mode, const XPar<double2>& q, dim3 grid, size_t lds) {
    if (mode == 0) hipLaunchKernelGGL((k_div_dct_x<0, R>), grid, dim3(kXThreads), lds, s, q);
    else if (mode == 1) hipLaunchKernelGGL((k_div_dct_x<1, R>), grid, dim3(kXThreads), lds, s, q);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int R>
hipError_t launch_xdct_inv_r(hipStream_t s, const XPar<double2>& q, dim3 grid, size_t lds) {
    hipLaunchKernelGGL((k_idct_x<R>), grid, dim3(kXThreads), lds, s, q);
    return hipGetLastError();
}

}  // namespace dfp

#define DFP_INSTANTIATE_XDCT(RV)                                                                       \
    namespace dfp {                                                                                    \
    template hipError_t launch_xdct_fwd_r<RV>(hipStream_t, int, const XPar<double2>&, dim3, size_t); \
    template hipError_t launch_xdct_inv_r<RV>(hipStream_t, const XPar<double2>&, dim3, size_t);      \
    }
