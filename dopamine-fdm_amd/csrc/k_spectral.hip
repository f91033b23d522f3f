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
// twiddles sit back to back in the plan's twiddle array and are staged into
// LDS with one copy (the FP32 form has its own rounded copy of the table).
//
// Every kernel is a block body over one (column block, y row) of one case;
// k_* run it for one case (blockIdx = block coordinates), kb_* for a batch of
// cases, each workgroup walking the work items w = blockIdx.x,
// blockIdx.x + gridDim.x, ... of the concatenated per-case block lists.
#include "fft_lds.h"

namespace dfp {

constexpr int kXThreads = 512;

// small DFT of R <= 4 points held in a[0..R-1], in place, sign SIGN
template <int SIGN, typename C>
__device__ __forceinline__ void dft_small(C (&a)[4], int R) {
    using Rl = typename CplxT<C>::real;
    if (R == 2) {
        const C s = cadd(a[0], a[1]), d = csub(a[0], a[1]);
        a[0] = s;
        a[1] = d;
    } else if (R == 3) {
        const Rl s3 = (Rl)0.86602540378443864676372317075294;
        const C t1 = cadd(a[1], a[2]);
        const C t2 = CplxT<C>::make(a[0].x - (Rl)0.5 * t1.x, a[0].y - (Rl)0.5 * t1.y);
        const C dd = csub(a[1], a[2]);
        const C jd = mul_si<SIGN>(CplxT<C>::make(s3 * dd.x, s3 * dd.y));
        a[0] = cadd(a[0], t1);
        a[1] = cadd(t2, jd);
        a[2] = csub(t2, jd);
    } else if (R == 4) {
        const C s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const C s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
        const C jd13 = mul_si<SIGN>(d13);
        a[0] = cadd(s02, s13);
        a[1] = cadd(d02, jd13);
        a[2] = csub(s02, s13);
        a[3] = csub(d02, jd13);
    }
}

// ---------------------------------------------------------------------------
// divergence + x R2C + first z stage
// ---------------------------------------------------------------------------
// MODE 0: divergence of (U,V,W) (projection.f90:55-57); MODE 1: rhs_p as
// given; MODE 2: the FP64 residual b - L_pin phi of the FP32 form, where
// L_pin is the 7-point operator whose eigenvalues the spectral solve
// inverts: 3-point stencils in x and z with spacings dx, dz (eigenvalues
// kxx, kzz), Dyy in y, and the (0,0) pin of projection.f90:292, which in
// physical space adds 0.5*Dyy(2,2)*mean_xz(phi(:,2,:)) on the first row.
// C: double2 (FP64 spectrum) or float2 (FP32 spectrum; the field is still
// formed in FP64 and rounded once).  a.b_out receives the FP64 field in the
// rhs_p layout; a.rowpart the per-block sums of the field (FP64, fixed
// order), rowpart[j*nbx + bx].
template <int MODE, typename C>
__device__ __forceinline__ void xfwd_body(const XPar<C>& q, int bx, int j, unsigned char* smem_raw) {
    using Rl = typename CplxT<C>::real;
    __shared__ double wsum[kXThreads / 64];
    const Geo& g = q.g;
    const XArgs& a = q.a;
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int M = q.fd.n;
    const int nxp = g.nxp;
    const int R = q.R;
    const int nz2 = g.nzp / R;
    const int k0 = bx * q.T;
    const int nt = q.T;  // T divides nz2
    const int nl = R * nt;
    const int ntw = q.ntw;
    C* A = smem;
    C* twl = smem + nl * M;         // FFT stage twiddles
    C* rtl = twl + ntw;             // R2C unpack twiddles, M+1
    C* tzl = rtl + (M + 1);         // first z stage twiddles, nz2*(R-1)
    Rl* Ar = reinterpret_cast<Rl*>(A);
    // the three tables are contiguous in the plan's twiddle array
    for (int e = threadIdx.x; e < ntw + (M + 1) + nz2 * (R - 1); e += kXThreads) twl[e] = q.tw[e];

    // reference indices of cell (i,j,k): (i+2, j+2, k+2)
    const long pu_j = (long)g.nx, pu_k = (long)g.nx * g.nyg;
    const long pv_j = (long)g.nxg, pv_k = (long)g.nxg * g.ny;
    const long pw_j = (long)g.nxg, pw_k = (long)g.nxg * g.nyg;
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const int total = nl * nxp;
    const double inv_dy = g.inv_dyf[j];
    double mean0 = 0.0;
    if (MODE == 2 && j == 0) {
        for (int t = 0; t < q.nbx; ++t) mean0 += a.part0[t];
        mean0 = mean0 / ((double)nxp * (double)g.nzp);
    }
    double rsum = 0.0;
    // four elements per thread per batch; every load is issued before any is used
    for (int base = 0; base < total; base += 4 * kXThreads) {
        double v[4];
        int dst[4];
        long cell[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e0 = base + (int)threadIdx.x + u * kXThreads;
            const int e = min(e0, total - 1);
            const int l = (int)q.div_nxp.div((uint32_t)e);
            const int i = e - l * nxp;
            const int t = (int)q.div_t.div((uint32_t)l);
            const int k = k0 + (l - t * nt) + t * nz2;
            dst[u] = e0 < total ? e : -1;
            cell[u] = i + j * pr_j + k * pr_k;
            if (MODE == 1) {
                v[u] = a.rhs[cell[u]];
            } else if (MODE == 2) {
                // rhs holds phi; index nxp (x) and nzp (z) hold the periodic duplicates of 0
                const double* ph = a.rhs + cell[u];
                const double p0 = ph[0];
                const double pxm = ph[i == 0 ? nxp - 1 : -1];
                const double pxp = ph[1];
                const double pzm = ph[k == 0 ? (long)(g.nzp - 1) * pr_k : -pr_k];
                const double pzp = ph[pr_k];
                const double pym = j > 0 ? ph[-pr_j] : 0.0;
                const double pyp = j < g.nym - 1 ? ph[pr_j] : 0.0;
                const double lo = j > 0 ? q.o.dl[j - 1] : 0.0;
                const double up = j < g.nym - 1 ? q.o.du[j] : 0.0;
                double lap = (pxp - 2.0 * p0 + pxm) * a.idx2 + (lo * pym + q.o.dd[j] * p0 + up * pyp) +
                             (pzp - 2.0 * p0 + pzm) * a.idz2;
                if (j == 0) lap += 0.5 * q.o.dd[0] * mean0;
                v[u] = a.b[cell[u]] - lap;
            } else {
                const long iu = (i + 1) + (j + 1) * pu_j + (k + 1) * pu_k;  // U(i+2,j+2,k+2)
                const long iv = (i + 1) + (j + 1) * pv_j + (k + 1) * pv_k;  // V(i+2,j+2,k+2)
                const long iw = (i + 1) + (j + 1) * pw_j + (k + 1) * pw_k;  // W(i+2,j+2,k+2)
                const double u1 = a.U[iu], u0 = a.U[iu - 1];
                const double v1 = a.V[iv], v0 = a.V[iv - pv_j];
                const double w1 = a.W[iw], w0 = a.W[iw - pw_k];
                const double idz = g.inv_dzf[k];
                {
#pragma clang fp contract(off)
                    // projection.f90:55-57
                    const double du = (u1 - u0) * g.inv_dx;
                    const double dv = (v1 - v0) * inv_dy;
                    const double dw = (w1 - w0) * idz;
                    v[u] = du + dv + dw;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (dst[u] >= 0) {
                Ar[dst[u]] = (Rl)v[u];
                if (MODE != 2 && a.b_out) a.b_out[cell[u]] = v[u];
                rsum += v[u];
            }
        }
    }
    if (a.rowpart) {
        const double sb = block_sum<kXThreads>(rsum, wsum);
        if (threadIdx.x == 0) a.rowpart[(long)j * q.nbx + bx] = sb;
    }
    __syncthreads();
    lds_fft_inplace<-1, 2>(A, nl, M, q.fd, twl);

    // unpack the R2C modes of the R lines of each column, apply the first z stage
    for (int pidx = threadIdx.x; pidx < (M + 1) * nt; pidx += kXThreads) {
        const int kx = (int)q.div_t.div((uint32_t)pidx);
        const int c = pidx - kx * nt;
        C av[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t < R) {
                const C* Z = A + (t * nt + c) * M;
                const C zk = Z[kx == M ? 0 : kx];
                const C zm = Z[kx == 0 ? 0 : M - kx];
                const C E = CplxT<C>::make((Rl)0.5 * (zk.x + zm.x), (Rl)0.5 * (zk.y - zm.y));
                const C D = CplxT<C>::make(zk.x - zm.x, zk.y + zm.y);  // zk - conj(zm)
                const C O = CplxT<C>::make((Rl)0.5 * D.y, (Rl)-0.5 * D.x);  // D / (2i)
                av[t] = cadd(E, cmul(rtl[kx], O));
            } else {
                av[t] = CplxT<C>::make(0, 0);
            }
        }
        dft_small<-1>(av, R);
        const int kp = k0 + c;
        C* out = q.spec + ((long)kx * R * g.nym + j) * nz2 + kp;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u < R) {
                const C y = u == 0 ? av[0] : cmul(av[u], tzl[kp * (R - 1) + (u - 1)]);
                out[(long)u * g.nym * nz2] = y;
            }
        }
    }
}

template <int MODE, typename C>
__global__ void __launch_bounds__(kXThreads) k_div_rfft_x(XPar<C> q) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int j = L / gridDim.x;
    xfwd_body<MODE, C>(q, L - j * gridDim.x, j, smem_raw);
}

// case of work item w: the last c with off[c] <= w (off[0] = 0, off[nc] = total)
__device__ __forceinline__ int find_case(const int* __restrict__ off, int nc, int w) {
    int lo = 0, hi = nc;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int MODE, typename C>
__global__ void __launch_bounds__(kXThreads) kb_div_rfft_x(const XPar<C>* __restrict__ cs, const int* __restrict__ off,
                                                           const int* __restrict__ pnc) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nc = *pnc;
    const int total = off[nc];
    const int me = xcd_remap(blockIdx.x, gridDim.x);
    for (int base = 0; base < total; base += gridDim.x) {
        const int w = base + me;
        if (w >= total) break;
        const int c = find_case(off, nc, w);
        const XPar<C>& q = cs[c];
        const int li = w - off[c];
        const int j = li / q.nbx;
        xfwd_body<MODE, C>(q, li - j * q.nbx, j, smem_raw);
        __syncthreads();
    }
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
// C: spectrum precision.  ACCUM: phi(i,j,k) = phi_old + x*scale + pbar[j]
// (the refinement sweeps of the FP32 form, pbar = the FP64 (0,0)-mode
// correction of row j, may be null); otherwise phi = x*scale (+ pbar[j]).
// q.part0, if given, receives the per-block sums of the new phi on the first
// row (j = 0), part0[bx] (FP64, fixed order).
template <typename C, int ACCUM>
__device__ __forceinline__ void xinv_body(const XPar<C>& q, int bx, int j, unsigned char* smem_raw) {
    using Rl = typename CplxT<C>::real;
    __shared__ double wsum[kXThreads / 64];
    const Geo& g = q.g;
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int M = q.fd.n;
    const int ld = M + 1;
    const int nxp = g.nxp;
    const int R = q.R;
    const int nz2 = g.nzp / R;
    const int k0 = bx * q.T;
    const int nt = q.T;  // T divides nz2
    const int nl = R * nt;
    const int ntw = q.ntw;
    C* A = smem;
    C* twl = smem + nl * ld;        // FFT stage twiddles
    C* rtl = twl + ntw;             // C2R packing twiddles, M+1
    C* tzl = rtl + (M + 1);         // last z stage twiddles, nz2*(R-1)
    // the three tables are contiguous in the plan's twiddle array
    for (int e = threadIdx.x; e < ntw + (M + 1) + nz2 * (R - 1); e += kXThreads) twl[e] = q.tw[e];
    __syncthreads();

    // gather the R sub-line values of each (kx, column), undo the first z stage
    const int npairs = (M + 1) * nt;
    for (int base = 0; base < npairs; base += 2 * kXThreads) {
        C av[2][4];
        int pk[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int p0 = base + (int)threadIdx.x + h * kXThreads;
            const int pidx = min(p0, npairs - 1);
            pk[h] = p0 < npairs ? pidx : -1;
            const int kx = (int)q.div_t.div((uint32_t)pidx);
            const int c = pidx - kx * nt;
            const C* src = q.spec + ((long)kx * R * g.nym + j) * nz2 + k0 + c;
#pragma unroll
            for (int u = 0; u < 4; ++u) av[h][u] = u < R ? src[(long)u * g.nym * nz2] : CplxT<C>::make(0, 0);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (pk[h] < 0) continue;
            const int kx = (int)q.div_t.div((uint32_t)pk[h]);
            const int c = pk[h] - kx * nt;
            const int kp = k0 + c;
#pragma unroll
            for (int u = 1; u < 4; ++u)
                if (u < R) av[h][u] = cmulc(av[h][u], tzl[kp * (R - 1) + (u - 1)]);
            dft_small<1>(av[h], R);
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < R) A[(t * nt + c) * ld + kx] = av[h][t];
        }
    }
    __syncthreads();
    // C2R packing in place: the thread of pair (k, M-k) writes Z[k] and Z[M-k]
    const int half = M / 2 + 1;
    for (int e = threadIdx.x; e < nl * half; e += kXThreads) {
        const int l = (int)q.div_half.div((uint32_t)e);
        const int k = e - l * half;
        C* X = A + l * ld;
        const int km = M - k;
        const C xk = X[k];
        const C xm = X[km];
        if (k < M) {
            const C S = CplxT<C>::make(xk.x + xm.x, xk.y - xm.y);  // X[k] + conj X[M-k]
            const C D = CplxT<C>::make(xk.x - xm.x, xk.y + xm.y);  // X[k] - conj X[M-k]
            const C Tt = cmulc(D, rtl[k]);
            X[k] = CplxT<C>::make(S.x - Tt.y, S.y + Tt.x);
        }
        if (km != k && km < M) {
            const C S = CplxT<C>::make(xm.x + xk.x, xm.y - xk.y);
            const C D = CplxT<C>::make(xm.x - xk.x, xm.y + xk.y);
            const C Tt = cmulc(D, rtl[km]);
            X[km] = CplxT<C>::make(S.x - Tt.y, S.y + Tt.x);
        }
    }
    __syncthreads();
    lds_fft_inplace<1, 2>(A, nl, ld, q.fd, twl);
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const long pp_j = (long)g.nxg, pp_k = (long)g.nxg * g.nyg;
    double* rr = q.rhs + (long)j * pr_j;
    double* P = q.P;
    const double scale = q.scale, dtc = q.dtc;
    const int save_p = q.save_p;
    double psum = 0.0;
    for (int e = threadIdx.x; e < nl * nxp; e += kXThreads) {
        const int l = (int)q.div_nxp.div((uint32_t)e);
        const int i = e - l * nxp;
        const int t = (int)q.div_t.div((uint32_t)l);
        const int k = k0 + (l - t * nt) + t * nz2;
        double phi = (double)reinterpret_cast<const Rl*>(A + l * ld)[i] * scale;
        if (q.pbar) phi += q.pbar[j];
        if (ACCUM) phi += rr[i + k * pr_k];
        psum += phi;
        // rhs_p(i+2, j+2, k+2) and the duplicates (i = nxp -> nxg-1, k = nzp -> nzp+2)
        rr[i + k * pr_k] = phi;
        if (i == 0) rr[nxp + k * pr_k] = phi;
        if (k == 0) {
            rr[i + (long)g.nzp * pr_k] = phi;
            if (i == 0) rr[nxp + (long)g.nzp * pr_k] = phi;
        }
        if (save_p) {
            // every copy of this cell in P (reference 1-based indices; 0 = none):
            // x: i+2, 1 if i = nxp-1, nxg-1 if i = 0, nxg if i = 1
            // y: j+2, 1 if j = 0, nyg if j = nym-1
            // z: k+2, 1 if k = nzp-1, nzg-1 if k = 0, nzg if k = 1 (mod nzp)
            const double pv = phi / dtc;
            const int x1 = i == nxp - 1 ? 1 : 0, x2 = i == 0 ? g.nxg - 1 : 0, x3 = i == 1 ? g.nxg : 0;
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
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const int xx = a == 0 ? i + 2 : (a == 1 ? x1 : (a == 2 ? x2 : x3));
                        if (xx == 0) continue;
                        P[(xx - 1) + (long)(yy - 1) * pp_j + (long)(zz - 1) * pp_k] = pv;
                    }
                }
            }
        }
    }
    if (q.part0 && j == 0) {
        const double sblk = block_sum<kXThreads>(psum, wsum);
        if (threadIdx.x == 0) q.part0[bx] = sblk;
    }
}

template <typename C, int ACCUM>
__global__ void __launch_bounds__(kXThreads) k_irfft_x(XPar<C> q) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int j = L / gridDim.x;
    xinv_body<C, ACCUM>(q, L - j * gridDim.x, j, smem_raw);
}

template <typename C, int ACCUM>
__global__ void __launch_bounds__(kXThreads) kb_irfft_x(const XPar<C>* __restrict__ cs, const int* __restrict__ off,
                                                        const int* __restrict__ pnc) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nc = *pnc;
    const int total = off[nc];
    const int me = xcd_remap(blockIdx.x, gridDim.x);
    for (int base = 0; base < total; base += gridDim.x) {
        const int w = base + me;
        if (w >= total) break;
        const int c = find_case(off, nc, w);
        const XPar<C>& q = cs[c];
        const int li = w - off[c];
        const int j = li / q.nbx;
        xinv_body<C, ACCUM>(q, li - j * q.nbx, j, smem_raw);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host side: parameters, LDS sizes, launches
// ---------------------------------------------------------------------------
template <typename C>
XPar<C> make_xpar(const dfp_plan* p) {
    XPar<C> q{};
    const Geo& g = p->geo;
    q.g = g;
    q.o = p->ops;
    q.fd = p->fx;
    q.R = p->zr;
    q.T = p->tk_x;
    q.nbx = (g.nzp / p->zr) / p->tk_x;
    q.ntw = twiddle_count(p->fx);
    q.div_nxp = FastDiv(g.nxp);
    q.div_t = FastDiv(p->tk_x);
    q.div_half = FastDiv(p->fx.n / 2 + 1);
    q.scale = 1.0 / ((double)g.nxp * (double)g.nzp);
    q.dtc = 1.0;
    if (sizeof(C) == sizeof(double2)) {
        q.tw = reinterpret_cast<const C*>(p->tw_x);
        q.spec = reinterpret_cast<C*>(p->d_spec);
    } else {
        q.tw = reinterpret_cast<const C*>(p->tw_x32);
        q.spec = reinterpret_cast<C*>(p->d_spec);
    }
    return q;
}
template XPar<double2> make_xpar<double2>(const dfp_plan*);
template XPar<float2> make_xpar<float2>(const dfp_plan*);

// dynamic LDS of the forward (inv = 0) or inverse (inv = 1) x kernel of a plan
size_t x_lds_bytes(const dfp_plan* p, int inv, int f32) {
    const int R = p->zr, T = p->tk_x, nz2 = p->geo.nzp / R, M = p->fx.n;
    const size_t cs = f32 ? sizeof(float2) : sizeof(double2);
    return cs * ((size_t)R * T * (M + inv) + twiddle_count(p->fx) + M + 1 + (size_t)nz2 * (R - 1));
}

static const char* xname(int mode, bool f32) {
    static const char* names[2][3] = {{"div_rfft_x", "rfft_x_rhs", "res_rfft_x"},
                                      {"div_rfft_x32", "rfft_x_rhs32", "res_rfft_x32"}};
    return names[f32 ? 1 : 0][mode];
}

template <typename C>
static hipError_t launch_xfwd(dfp_plan* p, int mode, const XPar<C>& q) {
    const dim3 grid(q.nbx, q.g.nym);
    const size_t lds = x_lds_bytes(p, 0, sizeof(C) == 8);
    const int slot = prof_begin(p, xname(mode, sizeof(C) == 8));
    if (mode == 0) hipLaunchKernelGGL((k_div_rfft_x<0, C>), grid, dim3(kXThreads), lds, p->stream, q);
    else if (mode == 1) hipLaunchKernelGGL((k_div_rfft_x<1, C>), grid, dim3(kXThreads), lds, p->stream, q);
    else hipLaunchKernelGGL((k_div_rfft_x<2, C>), grid, dim3(kXThreads), lds, p->stream, q);
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_div_rfft_x(dfp_plan* p, const double* U, const double* V, const double* W, const double* rhs,
                             int from_rhs) {
    XPar<double2> q = make_xpar<double2>(p);
    q.a.U = U; q.a.V = V; q.a.W = W; q.a.rhs = rhs;
    return launch_xfwd(p, from_rhs ? 1 : 0, q);
}

hipError_t launch_rfft_x32(dfp_plan* p, int mode, const XArgs& a) {
    XPar<float2> q = make_xpar<float2>(p);
    q.a = a;
    return launch_xfwd(p, mode, q);
}

template <typename C, int ACCUM>
static hipError_t launch_xinv(dfp_plan* p, const XPar<C>& q) {
    const dim3 grid(q.nbx, q.g.nym);
    const size_t lds = x_lds_bytes(p, 1, sizeof(C) == 8);
    const int slot = prof_begin(p, sizeof(C) == 8 ? "irfft_x32" : "irfft_x");
    hipLaunchKernelGGL((k_irfft_x<C, ACCUM>), grid, dim3(kXThreads), lds, p->stream, q);
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_irfft_x(dfp_plan* p, double* rhs, double* P, double dtc, int save_p) {
    XPar<double2> q = make_xpar<double2>(p);
    q.rhs = rhs; q.P = P; q.dtc = dtc; q.save_p = save_p;
    return launch_xinv<double2, 0>(p, q);
}

hipError_t launch_irfft_x32(dfp_plan* p, double* rhs, int accumulate, const double* pbar, double* part0, double* P,
                            double dtc, int save_p) {
    XPar<float2> q = make_xpar<float2>(p);
    q.rhs = rhs; q.pbar = pbar; q.part0 = part0;
    q.P = P; q.dtc = save_p ? dtc : 1.0; q.save_p = save_p;
    if (accumulate) return launch_xinv<float2, 1>(p, q);
    return launch_xinv<float2, 0>(p, q);
}

// batched launches: cs, off and the case count nc in device memory, grid =
// persistent workgroups
template <typename C>
hipError_t launch_xfwd_batch(hipStream_t s, int mode, const XPar<C>* cs, const int* off, const int* nc, int grid,
                             size_t lds) {
    if (mode == 0) hipLaunchKernelGGL((kb_div_rfft_x<0, C>), dim3(grid), dim3(kXThreads), lds, s, cs, off, nc);
    else if (mode == 1) hipLaunchKernelGGL((kb_div_rfft_x<1, C>), dim3(grid), dim3(kXThreads), lds, s, cs, off, nc);
    else hipLaunchKernelGGL((kb_div_rfft_x<2, C>), dim3(grid), dim3(kXThreads), lds, s, cs, off, nc);
    return hipGetLastError();
}
template hipError_t launch_xfwd_batch<double2>(hipStream_t, int, const XPar<double2>*, const int*, const int*, int,
                                               size_t);
template hipError_t launch_xfwd_batch<float2>(hipStream_t, int, const XPar<float2>*, const int*, const int*, int,
                                              size_t);

template <typename C>
hipError_t launch_xinv_batch(hipStream_t s, int accum, const XPar<C>* cs, const int* off, const int* nc, int grid,
                             size_t lds) {
    if (accum) hipLaunchKernelGGL((kb_irfft_x<C, 1>), dim3(grid), dim3(kXThreads), lds, s, cs, off, nc);
    else hipLaunchKernelGGL((kb_irfft_x<C, 0>), dim3(grid), dim3(kXThreads), lds, s, cs, off, nc);
    return hipGetLastError();
}
template hipError_t launch_xinv_batch<double2>(hipStream_t, int, const XPar<double2>*, const int*, const int*, int,
                                               size_t);
template hipError_t launch_xinv_batch<float2>(hipStream_t, int, const XPar<float2>*, const int*, const int*, int,
                                              size_t);

}  // namespace dfp
