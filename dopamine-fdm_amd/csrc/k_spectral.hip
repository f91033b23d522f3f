// Spectral stages of the projection on one MI355X (FP64), x direction:
//   k_div_rfft_x   divergence of (U,V,W) (or rhs_p) -> real FFT in x -> first
//                  radix-R stage of the z FFT (decimation in frequency)
//   k_irfft_x      last radix-R stage of the inverse z FFT (decimation in
//                  time) -> inverse real FFT in x -> phi (and P) in the
//                  reference layout with the periodic duplicates
// (the remaining z stages and the y solve are in k_zsolve.hip)
//
// z split.  With nzp = R*nz2, the DIF first stage maps the z line x[k] to R
// independent sub-lines y_u[k'] = (sum_t x[k'+t*nz2] w_R^{t u}) w_nzp^{k' u},
// u = 0..R-1, k' = 0..nz2-1, whose length-nz2 DFTs are the modes
// X[R*m + u].  Each (kx, u) sub-plane is then transformed and solved by its
// own block.  The inverse recombines x[k'+t*nz2] = sum_u w_R^{-t u}
// w_nzp^{-k' u} z_u[k'] from the inverse sub-line transforms z_u.
//
// Spectrum layout in HBM: spec[((kx*R + u)*nym + j)*nz2 + m], complex FP64,
// kx = 0..nxp/2 (the R2C half of the x modes), j = 0..nym-1.
//
// x transforms.  A block holds R*nt lines of one y row: columns
// k' = k0..k0+nt-1 and, for each, the R z positions k' + t*nz2 (line
// l = t*nt + c).  A real line of length nxp is packed as nxp/2 complex
// points z[n] = x[2n] + i x[2n+1] and transformed with a length-nxp/2
// complex FFT; X[k] = E[k] + w^k O[k], E = (Z[k]+conj Z[M-k])/2,
// O = (Z[k]-conj Z[M-k])/(2i), w = exp(-2 pi i/nxp) gives the nxp/2+1 modes.
// The x FFT stage twiddles, the packing twiddles w^k and the first z stage
// twiddles sit back to back in the plan's twiddle array (tw, rtw, tzs) and
// are staged into LDS with one copy.
#include "fft_lds.h"

namespace dfp {

constexpr int kXThreads = 512;

// small DFT of R <= 4 points held in a[0..R-1], in place, sign SIGN
template <int SIGN>
__device__ __forceinline__ void dft_small(double2 (&a)[4], int R) {
    if (R == 2) {
        const double2 s = cadd(a[0], a[1]), d = csub(a[0], a[1]);
        a[0] = s;
        a[1] = d;
    } else if (R == 3) {
        const double s3 = 0.86602540378443864676372317075294;
        const double2 t1 = cadd(a[1], a[2]);
        const double2 t2 = make_double2(a[0].x - 0.5 * t1.x, a[0].y - 0.5 * t1.y);
        const double2 dd = csub(a[1], a[2]);
        const double2 jd = mul_si<SIGN>(make_double2(s3 * dd.x, s3 * dd.y));
        a[0] = cadd(a[0], t1);
        a[1] = cadd(t2, jd);
        a[2] = csub(t2, jd);
    } else if (R == 4) {
        const double2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const double2 s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
        const double2 jd13 = mul_si<SIGN>(d13);
        a[0] = cadd(s02, s13);
        a[1] = cadd(d02, jd13);
        a[2] = csub(s02, s13);
        a[3] = csub(d02, jd13);
    }
}

// ---------------------------------------------------------------------------
// divergence + x R2C + first z stage
// ---------------------------------------------------------------------------
template <int FROM_RHS>
__global__ void __launch_bounds__(kXThreads) k_div_rfft_x(const double* __restrict__ U, const double* __restrict__ V,
                                                          const double* __restrict__ W, const double* __restrict__ rhs,
                                                          double2* __restrict__ spec, Geo g, FftDesc fd,
                                                          const double2* __restrict__ tw, int ntw,
                                                          const double2* __restrict__ rtw,
                                                          const double2* __restrict__ tzs, int R, int T) {
    extern __shared__ double2 smem[];
    const int M = fd.n;
    const int nxp = g.nxp;
    const int nz2 = g.nzp / R;
    const int j = blockIdx.y;
    const int k0 = blockIdx.x * T;
    const int nt = min(T, nz2 - k0);
    const int nl = R * nt;
    double2* A = smem;
    double2* twl = smem + R * T * M;      // FFT stage twiddles
    double2* rtl = twl + ntw;             // R2C unpack twiddles, M+1
    double2* tzl = rtl + (M + 1);         // first z stage twiddles, nz2*(R-1)
    double* Ar = reinterpret_cast<double*>(A);
    // the three tables are contiguous in the plan's twiddle array
    for (int e = threadIdx.x; e < ntw + (M + 1) + nz2 * (R - 1); e += kXThreads) twl[e] = tw[e];

    // reference indices of cell (i,j,k): (i+2, j+2, k+2)
    const long pu_j = (long)g.nx, pu_k = (long)g.nx * g.nyg;
    const long pv_j = (long)g.nxg, pv_k = (long)g.nxg * g.ny;
    const long pw_j = (long)g.nxg, pw_k = (long)g.nxg * g.nyg;
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const int total = nl * nxp;
    const double inv_dy = g.inv_dyf[j];
    // four elements per thread per batch; every load is issued before any is used
    for (int base = 0; base < total; base += 4 * kXThreads) {
        double v[4];
        int dst[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e0 = base + (int)threadIdx.x + q * kXThreads;
            const int e = min(e0, total - 1);
            const int l = e / nxp;
            const int i = e - l * nxp;
            const int t = l / nt;
            const int k = k0 + (l - t * nt) + t * nz2;
            dst[q] = e0 < total ? e : -1;
            if (FROM_RHS) {
                v[q] = rhs[i + j * pr_j + k * pr_k];
            } else {
                const long iu = (i + 1) + (j + 1) * pu_j + (k + 1) * pu_k;  // U(i+2,j+2,k+2)
                const long iv = (i + 1) + (j + 1) * pv_j + (k + 1) * pv_k;  // V(i+2,j+2,k+2)
                const long iw = (i + 1) + (j + 1) * pw_j + (k + 1) * pw_k;  // W(i+2,j+2,k+2)
                const double u1 = U[iu], u0 = U[iu - 1];
                const double v1 = V[iv], v0 = V[iv - pv_j];
                const double w1 = W[iw], w0 = W[iw - pw_k];
                const double idz = g.inv_dzf[k];
                {
#pragma clang fp contract(off)
                    // projection.f90:55-57
                    const double du = (u1 - u0) * g.inv_dx;
                    const double dv = (v1 - v0) * inv_dy;
                    const double dw = (w1 - w0) * idz;
                    v[q] = du + dv + dw;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (dst[q] >= 0) Ar[dst[q]] = v[q];
    }
    __syncthreads();
    lds_fft_inplace<-1, 2>(A, nl, M, fd, twl);

    // unpack the R2C modes of the R lines of each column, apply the first z stage
    for (int pidx = threadIdx.x; pidx < (M + 1) * nt; pidx += kXThreads) {
        const int kx = pidx / nt;
        const int c = pidx - kx * nt;
        double2 a[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t < R) {
                const double2* Z = A + (t * nt + c) * M;
                const double2 zk = Z[kx == M ? 0 : kx];
                const double2 zm = Z[kx == 0 ? 0 : M - kx];
                const double2 E = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
                const double2 D = make_double2(zk.x - zm.x, zk.y + zm.y);  // zk - conj(zm)
                const double2 O = make_double2(0.5 * D.y, -0.5 * D.x);     // D / (2i)
                a[t] = cadd(E, cmul(rtl[kx], O));
            } else {
                a[t] = make_double2(0.0, 0.0);
            }
        }
        dft_small<-1>(a, R);
        const int kp = k0 + c;
        double2* out = spec + ((long)kx * R * g.nym + j) * nz2 + kp;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u < R) {
                const double2 y = u == 0 ? a[0] : cmul(a[u], tzl[kp * (R - 1) + (u - 1)]);
                out[(long)u * g.nym * nz2] = y;
            }
        }
    }
}

hipError_t launch_div_rfft_x(dfp_plan* p, const double* U, const double* V, const double* W, const double* rhs,
                             int from_rhs) {
    const Geo& g = p->geo;
    const int R = p->zr, T = p->tk_x, nz2 = g.nzp / R;
    dim3 grid(ceil_div(nz2, T), g.nym);
    const int ntw = twiddle_count(p->fx);
    const size_t lds = sizeof(double2) * ((size_t)R * T * p->fx.n + ntw + p->fx.n + 1 + (size_t)nz2 * (R - 1));
    const int slot = prof_begin(p, from_rhs ? "rfft_x_rhs" : "div_rfft_x");
    if (from_rhs)
        hipLaunchKernelGGL(k_div_rfft_x<1>, grid, dim3(kXThreads), lds, p->stream, U, V, W, rhs, p->d_spec, g, p->fx,
                           p->tw_x, ntw, p->tw_r2c, p->tw_zs, R, T);
    else
        hipLaunchKernelGGL(k_div_rfft_x<0>, grid, dim3(kXThreads), lds, p->stream, U, V, W, rhs, p->d_spec, g, p->fx,
                           p->tw_x, ntw, p->tw_r2c, p->tw_zs, R, T);
    prof_end(p, slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// last inverse z stage + inverse x real FFT, phi = x / (nxp*nzp) into rhs_p
// with the periodic duplicates of projection.f90:453-456, :479-482, and
// optionally P = phi/(dt*c) with the ghosts of projection.f90:396-410 and
// boundary_conditions.f90:283-285, :597-599.
// C2R packing: Z[k] = (X[k] + conj X[M-k]) + i exp(+2 pi i k/nxp)(X[k] - conj X[M-k])
// gives z = IFFT_M(Z) = x_unnormalised[2n] + i x_unnormalised[2n+1].
// Lines are stored with stride M+1 (the nxp/2+1 modes, then Z in place).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kXThreads) k_irfft_x(const double2* __restrict__ spec, double* __restrict__ rhs,
                                                       double* __restrict__ P, Geo g, FftDesc fd,
                                                       const double2* __restrict__ tw, int ntw,
                                                       const double2* __restrict__ rtw,
                                                       const double2* __restrict__ tzs, int R, int T, double scale,
                                                       double dtc, int save_p) {
    extern __shared__ double2 smem[];
    const int M = fd.n;
    const int ld = M + 1;
    const int nxp = g.nxp;
    const int nz2 = g.nzp / R;
    const int j = blockIdx.y;
    const int k0 = blockIdx.x * T;
    const int nt = min(T, nz2 - k0);
    const int nl = R * nt;
    double2* A = smem;
    double2* twl = smem + R * T * ld;     // FFT stage twiddles
    double2* rtl = twl + ntw;             // C2R packing twiddles, M+1
    double2* tzl = rtl + (M + 1);         // last z stage twiddles, nz2*(R-1)
    // the three tables are contiguous in the plan's twiddle array
    for (int e = threadIdx.x; e < ntw + (M + 1) + nz2 * (R - 1); e += kXThreads) twl[e] = tw[e];
    __syncthreads();

    // gather the R sub-line values of each (kx, column), undo the first z stage
    const int npairs = (M + 1) * nt;
    for (int base = 0; base < npairs; base += 2 * kXThreads) {
        double2 a[2][4];
        int pk[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int p0 = base + (int)threadIdx.x + h * kXThreads;
            const int pidx = min(p0, npairs - 1);
            pk[h] = p0 < npairs ? pidx : -1;
            const int kx = pidx / nt;
            const int c = pidx - kx * nt;
            const double2* src = spec + ((long)kx * R * g.nym + j) * nz2 + k0 + c;
#pragma unroll
            for (int u = 0; u < 4; ++u) a[h][u] = u < R ? src[(long)u * g.nym * nz2] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (pk[h] < 0) continue;
            const int kx = pk[h] / nt;
            const int c = pk[h] - kx * nt;
            const int kp = k0 + c;
#pragma unroll
            for (int u = 1; u < 4; ++u)
                if (u < R) a[h][u] = cmulc(a[h][u], tzl[kp * (R - 1) + (u - 1)]);
            dft_small<1>(a[h], R);
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < R) A[(t * nt + c) * ld + kx] = a[h][t];
        }
    }
    __syncthreads();
    // C2R packing in place: the thread of pair (k, M-k) writes Z[k] and Z[M-k]
    const int half = M / 2 + 1;
    for (int e = threadIdx.x; e < nl * half; e += kXThreads) {
        const int l = e / half;
        const int k = e - l * half;
        double2* X = A + l * ld;
        const int km = M - k;
        const double2 xk = X[k];
        const double2 xm = X[km];
        if (k < M) {
            const double2 S = make_double2(xk.x + xm.x, xk.y - xm.y);  // X[k] + conj X[M-k]
            const double2 D = make_double2(xk.x - xm.x, xk.y + xm.y);  // X[k] - conj X[M-k]
            const double2 Tt = cmulc(D, rtl[k]);
            X[k] = make_double2(S.x - Tt.y, S.y + Tt.x);
        }
        if (km != k && km < M) {
            const double2 S = make_double2(xm.x + xk.x, xm.y - xk.y);
            const double2 D = make_double2(xm.x - xk.x, xm.y + xk.y);
            const double2 Tt = cmulc(D, rtl[km]);
            X[km] = make_double2(S.x - Tt.y, S.y + Tt.x);
        }
    }
    __syncthreads();
    lds_fft_inplace<1, 2>(A, nl, ld, fd, twl);
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const long pp_j = (long)g.nxg, pp_k = (long)g.nxg * g.nyg;
    double* rr = rhs + (long)j * pr_j;
    for (int e = threadIdx.x; e < nl * nxp; e += kXThreads) {
        const int l = e / nxp;
        const int i = e - l * nxp;
        const int t = l / nt;
        const int k = k0 + (l - t * nt) + t * nz2;
        const double phi = reinterpret_cast<const double*>(A + l * ld)[i] * scale;
        // rhs_p(i+2, j+2, k+2) and the duplicates (i = nxp -> nxg-1, k = nzp -> nzp+2)
        rr[i + k * pr_k] = phi;
        if (i == 0) rr[nxp + k * pr_k] = phi;
        if (k == 0) {
            rr[i + (long)g.nzp * pr_k] = phi;
            if (i == 0) rr[nxp + (long)g.nzp * pr_k] = phi;
        }
        if (save_p) {
            const double pv = phi / dtc;
            // reference 1-based indices of every copy of this cell in P
            int xs[4], ys[3], zs[4];
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
            for (int c3 = 0; c3 < nzs; ++c3)
                for (int b = 0; b < nys; ++b)
                    for (int a = 0; a < nxs; ++a)
                        P[(xs[a] - 1) + (long)(ys[b] - 1) * pp_j + (long)(zs[c3] - 1) * pp_k] = pv;
        }
    }
}

hipError_t launch_irfft_x(dfp_plan* p, double* rhs, double* P, double dtc, int save_p) {
    const Geo& g = p->geo;
    const int R = p->zr, T = p->tk_x, nz2 = g.nzp / R;
    dim3 grid(ceil_div(nz2, T), g.nym);
    const int ntw = twiddle_count(p->fx);
    const size_t lds = sizeof(double2) * ((size_t)R * T * (p->fx.n + 1) + ntw + p->fx.n + 1 + (size_t)nz2 * (R - 1));
    const double scale = 1.0 / ((double)g.nxp * (double)g.nzp);
    const int slot = prof_begin(p, "irfft_x");
    hipLaunchKernelGGL(k_irfft_x, grid, dim3(kXThreads), lds, p->stream, p->d_spec, rhs, P, g, p->fx, p->tw_x, ntw,
                       p->tw_r2c, p->tw_zs, R, T, scale, dtc, save_p);
    prof_end(p, slot);
    return hipGetLastError();
}

}  // namespace dfp
