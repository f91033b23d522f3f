// Spectral stages of the projection on one MI355X (FP64):
//   k_div_rfft_x   divergence of (U,V,W) (or rhs_p) fused with the real FFT in x
//   k_fft_z        complex FFT along z on contiguous rows of the spectrum
//   k_thomas_y     tridiagonal solve along y, one (kx,kz) mode per lane
//   k_irfft_x      inverse real FFT in x, scaling, phi (and P) in the
//                  reference layout with the periodic duplicates
//
// Spectrum layout in HBM: spec[(kx*nym + j)*nzp + kz], kx = 0..nxp/2 (the
// R2C half of the x modes), j = 0..nym-1, kz = 0..nzp-1; complex FP64.
// z rows (fixed kx, j) are contiguous, so the z-FFT reads whole rows and the
// Thomas sweep reads one coalesced row segment per y step.
#include "fft_lds.h"

namespace dfp {

// ---------------------------------------------------------------------------
// divergence + x real-to-complex FFT
// Block (k-tile, j): nl <= tk consecutive z-lines at one y.  The real line of
// length nxp is packed as nxp/2 complex points z[n] = x[2n] + i x[2n+1],
// transformed with a length-nxp/2 complex FFT and unpacked into the nxp/2+1
// modes  X[k] = E[k] + w^k O[k],  E = (Z[k]+conj Z[M-k])/2,
// O = (Z[k]-conj Z[M-k])/(2i),  w = exp(-2 pi i / nxp).
// ---------------------------------------------------------------------------
template <int FROM_RHS>
__global__ void __launch_bounds__(256) k_div_rfft_x(const double* __restrict__ U, const double* __restrict__ V,
                                                    const double* __restrict__ W, const double* __restrict__ rhs,
                                                    double2* __restrict__ spec, Geo g, FftDesc fd,
                                                    const double2* __restrict__ tw, const double2* __restrict__ rtw,
                                                    int tk) {
    extern __shared__ double2 smem[];
    const int M = fd.n;
    const int nxp = g.nxp;
    const int j = blockIdx.y;
    const int k0 = blockIdx.x * tk;
    const int nl = min(tk, g.nzp - k0);
    double2* A = smem;
    double2* B = smem + tk * M;
    double* Ar = reinterpret_cast<double*>(A);

    // reference indices of cell (i,j,k): (i+2, j+2, k+2)
    const long pu_j = (long)g.nx, pu_k = (long)g.nx * g.nyg;
    const long pv_j = (long)g.nxg, pv_k = (long)g.nxg * g.ny;
    const long pw_j = (long)g.nxg, pw_k = (long)g.nxg * g.nyg;
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    for (int e = threadIdx.x; e < nl * nxp; e += blockDim.x) {
        const int l = e / nxp;
        const int i = e - l * nxp;
        const int k = k0 + l;
        double v;
        if (FROM_RHS) {
            v = rhs[i + j * pr_j + k * pr_k];
        } else {
#pragma clang fp contract(off)
            // projection.f90:55-57 with 1-based (i+2, j+2, k+2)
            const long iu = (i + 1) + (j + 1) * pu_j + (k + 1) * pu_k;   // U(i+2,j+2,k+2)
            const long iv = (i + 1) + (j + 1) * pv_j + (k + 1) * pv_k;   // V(i+2,j+2,k+2)
            const long iw = (i + 1) + (j + 1) * pw_j + (k + 1) * pw_k;   // W(i+2,j+2,k+2)
            const double du = (U[iu] - U[iu - 1]) * g.inv_dx;
            const double dv = (V[iv] - V[iv - pv_j]) * g.inv_dyf[j];
            const double dw = (W[iw] - W[iw - pw_k]) * g.inv_dzf[k];
            v = du + dv + dw;
        }
        Ar[l * nxp + i] = v;
    }
    __syncthreads();
    const double2* R = lds_fft_lines<-1>(A, B, nl, M, fd, tw);
    for (int e = threadIdx.x; e < (M + 1) * nl; e += blockDim.x) {
        const int kx = e / nl;
        const int l = e - kx * nl;
        const double2 zk = R[l * M + (kx == M ? 0 : kx)];
        const double2 zm = R[l * M + (kx == 0 ? 0 : M - kx)];
        const double2 E = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
        // O = (zk - conj(zm)) / (2i)
        const double2 D = make_double2(zk.x - zm.x, zk.y + zm.y);
        const double2 O = make_double2(0.5 * D.y, -0.5 * D.x);
        const double2 X = cadd(E, cmul(rtw[kx], O));
        spec[((long)kx * g.nym + j) * g.nzp + k0 + l] = X;
    }
}

hipError_t launch_div_rfft_x(dfp_plan* p, const double* U, const double* V, const double* W, const double* rhs,
                             int from_rhs) {
    const Geo& g = p->geo;
    const int tk = p->tk_x;
    dim3 grid(ceil_div(g.nzp, tk), g.nym);
    const size_t lds = sizeof(double2) * 2 * (size_t)tk * p->fx.n;
    const int slot = prof_begin(p, from_rhs ? "rfft_x_rhs" : "div_rfft_x");
    if (from_rhs)
        hipLaunchKernelGGL(k_div_rfft_x<1>, grid, dim3(256), lds, p->stream, U, V, W, rhs, p->d_spec, g, p->fx,
                           p->tw_x, p->tw_r2c, tk);
    else
        hipLaunchKernelGGL(k_div_rfft_x<0>, grid, dim3(256), lds, p->stream, U, V, W, rhs, p->d_spec, g, p->fx,
                           p->tw_x, p->tw_r2c, tk);
    prof_end(p, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// z FFT on rows spec[row*nzp .. row*nzp+nzp-1], row = kx*nym + j
// ---------------------------------------------------------------------------
template <int SIGN>
__global__ void __launch_bounds__(256) k_fft_z(double2* __restrict__ spec, long nrows, FftDesc fd,
                                               const double2* __restrict__ tw, int rb) {
    extern __shared__ double2 smem[];
    const int n = fd.n;
    const long r0 = (long)blockIdx.x * rb;
    const int nr = (int)min((long)rb, nrows - r0);
    double2* A = smem;
    double2* B = smem + rb * n;
    const double2* src = spec + r0 * n;
    for (int e = threadIdx.x; e < nr * n; e += blockDim.x) A[e] = src[e];
    __syncthreads();
    const double2* R = lds_fft_lines<SIGN>(A, B, nr, n, fd, tw);
    double2* dstp = spec + r0 * n;
    for (int e = threadIdx.x; e < nr * n; e += blockDim.x) dstp[e] = R[e];
}

hipError_t launch_fft_z(dfp_plan* p, int sign) {
    const Geo& g = p->geo;
    const long nrows = (long)g.nxh * g.nym;
    const int rb = p->rb_z;
    dim3 grid(ceil_div(nrows, rb));
    const size_t lds = sizeof(double2) * 2 * (size_t)rb * p->fz.n;
    const int slot = prof_begin(p, sign < 0 ? "fft_z_fwd" : "fft_z_inv");
    if (sign < 0)
        hipLaunchKernelGGL(k_fft_z<-1>, grid, dim3(256), lds, p->stream, p->d_spec, nrows, p->fz, p->tw_z, rb);
    else
        hipLaunchKernelGGL(k_fft_z<1>, grid, dim3(256), lds, p->stream, p->d_spec, nrows, p->fz, p->tw_z, rb);
    prof_end(p, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tridiagonal solve along y for every kept (kx,kz) mode, the system of
// projection.f90:277-297: DL(j)=Dyy(j+1,j), DU(j)=Dyy(j,j+1),
// D(j)=Dyy(j,j)+(kxx+kzz), and D(2) *= 3/2 for the (0,0) mode (:292).
// Elimination without row interchanges (the matrices are diagonally
// dominant; LAPACK zgtsv interchanges rows only in a few low modes, which
// changes the result at the 1e-15 level), in zgtsv's operation order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_thomas_y(double2* __restrict__ spec, double* __restrict__ piv, Ops o,
                                                  int nxh, int nym, int nzp) {
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (long)nxh * nzp) return;
    const int kx = (int)(m / nzp);
    const int kz = (int)(m - (long)kx * nzp);
    const double lam = o.kxx[kx] + o.kzz[kz];
    const long base = (long)kx * nym * nzp + kz;
    double dp = o.dd[0] + lam;
    if (kx == 0 && kz == 0) dp = 3.0 / 2.0 * dp;
    double2 b = spec[base];
    piv[base] = dp;
    for (int j = 1; j < nym; ++j) {
        const double mult = o.dl[j - 1] / dp;
        dp = (o.dd[j] + lam) - mult * o.du[j - 1];
        double2 bj = spec[base + (long)j * nzp];
        bj.x = bj.x - mult * b.x;
        bj.y = bj.y - mult * b.y;
        spec[base + (long)j * nzp] = bj;
        piv[base + (long)j * nzp] = dp;
        b = bj;
    }
    double2 x = make_double2(b.x / dp, b.y / dp);
    spec[base + (long)(nym - 1) * nzp] = x;
    for (int j = nym - 2; j >= 0; --j) {
        const long a = base + (long)j * nzp;
        const double2 bj = spec[a];
        const double d = piv[a];
        x.x = (bj.x - o.du[j] * x.x) / d;
        x.y = (bj.y - o.du[j] * x.y) / d;
        spec[a] = x;
    }
}

hipError_t launch_thomas_y(dfp_plan* p) {
    const Geo& g = p->geo;
    const long modes = (long)g.nxh * g.nzp;
    const int slot = prof_begin(p, "thomas_y");
    hipLaunchKernelGGL(k_thomas_y, dim3(ceil_div(modes, 256)), dim3(256), 0, p->stream, p->d_spec, p->d_piv, p->ops,
                       g.nxh, g.nym, g.nzp);
    prof_end(p, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// inverse x real FFT, phi = x / (nxp*nzp) into rhs_p with the periodic
// duplicates of projection.f90:453-456, :479-482, and optionally
// P = phi/(dt*c) with the ghosts of projection.f90:396-410 and
// boundary_conditions.f90:283-285, :597-599.
// Packing: Z[k] = (X[k] + conj X[M-k]) + i exp(+2 pi i k/nxp)(X[k] - conj X[M-k])
// gives z = IFFT_M(Z) = x_unnormalised[2n] + i x_unnormalised[2n+1].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_irfft_x(const double2* __restrict__ spec, double* __restrict__ rhs,
                                                 double* __restrict__ P, Geo g, FftDesc fd,
                                                 const double2* __restrict__ tw, const double2* __restrict__ rtw,
                                                 int tk, double scale, double dtc, int save_p) {
    extern __shared__ double2 smem[];
    const int M = fd.n;
    const int nxp = g.nxp;
    const int j = blockIdx.y;
    const int k0 = blockIdx.x * tk;
    const int nl = min(tk, g.nzp - k0);
    double2* A = smem;                  // tk*(M+1): the nxp/2+1 modes
    double2* B = smem + tk * (M + 1);   // tk*(M+1)
    for (int e = threadIdx.x; e < (M + 1) * nl; e += blockDim.x) {
        const int kx = e / nl;
        const int l = e - kx * nl;
        A[l * (M + 1) + kx] = spec[((long)kx * g.nym + j) * g.nzp + k0 + l];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < M * nl; e += blockDim.x) {
        const int l = e / M;
        const int k = e - l * M;
        const double2 xk = A[l * (M + 1) + k];
        const double2 xm = A[l * (M + 1) + M - k];
        const double2 S = make_double2(xk.x + xm.x, xk.y - xm.y);   // X[k] + conj X[M-k]
        const double2 D = make_double2(xk.x - xm.x, xk.y + xm.y);   // X[k] - conj X[M-k]
        const double2 T = cmulc(D, rtw[k]);                          // exp(+2 pi i k/nxp) * D
        B[l * M + k] = make_double2(S.x - T.y, S.y + T.x);           // S + i T
    }
    __syncthreads();
    const double2* R = lds_fft_lines<1>(B, A, nl, M, fd, tw);
    const double* Rr = reinterpret_cast<const double*>(R);
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const long pp_j = (long)g.nxg, pp_k = (long)g.nxg * g.nyg;
    for (int e = threadIdx.x; e < nl * nxp; e += blockDim.x) {
        const int l = e / nxp;
        const int i = e - l * nxp;
        const int k = k0 + l;
        const double phi = Rr[l * nxp + i] * scale;
        // rhs_p(i+2, j+2, k+2) and the duplicates (i = nxp -> nxg-1, k = nzp -> nzp+2)
        double* rr = rhs + (long)j * pr_j;
        rr[i + k * pr_k] = phi;
        if (i == 0) rr[nxp + k * pr_k] = phi;
        if (k == 0) {
            rr[i + (long)g.nzp * pr_k] = phi;
            if (i == 0) rr[nxp + (long)g.nzp * pr_k] = phi;
        }
        if (save_p) {
            const double pv = phi / dtc;
            // reference 1-based indices of every copy of this cell in P
            int xs[3], ys[3], zs[3];
            int nxs = 0, nys = 0, nzs = 0;
            xs[nxs++] = i + 2;
            if (i == nxp - 1) xs[nxs++] = 1;
            if (i == 0) xs[nxs++] = g.nxg - 1;
            if (i == 1) xs[nxs++] = g.nxg;
            ys[nys++] = j + 2;
            if (j == 0) ys[nys++] = 1;
            if (j == g.nym - 1) ys[nys++] = g.nyg;
            zs[nzs++] = k + 2;
            if (k == g.nzp - 1) zs[nzs++] = 1;
            if (k == 0) zs[nzs++] = g.nzg - 1;
            if (k == 1 % g.nzp) zs[nzs++] = g.nzg;  // P(nzg) = P(3), the duplicate cell when nzp = 1
            for (int c = 0; c < nzs; ++c)
                for (int b = 0; b < nys; ++b)
                    for (int a = 0; a < nxs; ++a)
                        P[(xs[a] - 1) + (long)(ys[b] - 1) * pp_j + (long)(zs[c] - 1) * pp_k] = pv;
        }
    }
}

hipError_t launch_irfft_x(dfp_plan* p, double* rhs, double* P, double dtc, int save_p) {
    const Geo& g = p->geo;
    const int tk = p->tk_x;
    dim3 grid(ceil_div(g.nzp, tk), g.nym);
    const size_t lds = sizeof(double2) * 2 * (size_t)tk * (p->fx.n + 1);
    const double scale = 1.0 / ((double)g.nxp * (double)g.nzp);
    const int slot = prof_begin(p, "irfft_x");
    hipLaunchKernelGGL(k_irfft_x, grid, dim3(256), lds, p->stream, p->d_spec, rhs, P, g, p->fx, p->tw_x, p->tw_r2c, tk,
                       scale, dtc, save_p);
    prof_end(p, slot);
    return hipGetLastError();
}

}  // namespace dfp
