// Spectral stages of the projection on one MI355X, x direction:
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
#pragma once

#include "fft_lds.h"

namespace dfp {

constexpr int kXThreads = 512;

// Tile (column block bx, y row j) of the L-th workgroup of a case: strips of
// S column blocks, rows j in order within a strip.  Consecutive workgroups
// (which xcd_remap keeps on one XCD and its L2) then read the rows beside
// their own (the v* and phi halos of the forward kernel) and the column
// blocks beside them (the w* and phi halos; the other half of each 128-byte
// spectrum line) shortly after another workgroup did.  S = nbx is plain
// row-major order.
__device__ __forceinline__ void xtile(int L, int nbx, int nym, int S, int& bx, int& j) {
    const int sw = S * nym;
    const int s = L / sw;
    const int r = L - s * sw;
    const int w = min(S, nbx - s * S);  // column blocks of this strip
    j = r / w;
    bx = s * S + (r - j * w);
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
template <int MODE, typename C, int R, int KX>
__device__ __forceinline__ void xfwd_body(const XPar<C>& q, int bx, int j, unsigned char* smem_raw) {
    using Rl = typename CplxT<C>::real;
    __shared__ double wsum[kXThreads / 64];
    const Geo& g = q.g;
    const XArgs& a = q.a;
    const auto aU = gptr(a.U);
    const auto aV = gptr(a.V);
    const auto aW = gptr(a.W);
    const auto arhs = gptr(a.rhs);
    const auto ab = gptr(a.b);
    const auto abo = gptr(a.b_out);
    const auto inv_dyf = gptr(g.inv_dyf);
    const auto inv_dzf = gptr(g.inv_dzf);
    const auto odl = gptr(q.o.dl);
    const auto odd = gptr(q.o.dd);
    const auto odu = gptr(q.o.du);
    const auto spec = gptr(q.spec);
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int M = q.fd.n;
    // lines padded to M+1 points: the unpack below reads the same mode of
    // neighbouring columns in neighbouring lanes, which would otherwise all
    // fall in one LDS bank
    const int ld = M + 1;
    const int nxp = g.nxp;
    const int nz2 = g.nzp / R;
    const int k0 = bx * q.T;
    const int nt = q.T;  // T divides nz2
    const int nl = R * nt;
    const int ntw = q.ntw;
    C* A = smem;
    C* twl = smem + nl * ld;        // FFT stage twiddles
    C* rtl = twl + ntw;             // R2C unpack twiddles, M+1
    C* tzl = rtl + (M + 1);         // first z stage twiddles, nz2*(R-1)
    // the three tables are contiguous in the plan's twiddle array
    for (int e = threadIdx.x; e < ntw + (M + 1) + nz2 * (R - 1); e += kXThreads) twl[e] = gld<C>(gptr(q.tw) + e);

    // reference indices of cell (i,j,k): (i+2, j+2, k+2)
    const long pu_j = (long)g.nx, pu_k = (long)g.nx * g.nyg;
    const long pv_j = (long)g.nxg, pv_k = (long)g.nxg * g.ny;
    const long pw_j = (long)g.nxg, pw_k = (long)g.nxg * g.nyg;
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const int total = nl * nxp;
    const double inv_dy = inv_dyf[j];
    double mean0 = 0.0;
    if (MODE == 2 && j == 0) {
        for (int t = 0; t < q.npart0; ++t) mean0 += gptr(a.part0)[t];
        mean0 = mean0 / ((double)nxp * (double)g.nzp);
    }
    double rsum = 0.0;
    // two pairs of neighbouring points (i, i+1), i even, per thread per batch
    // (16-byte field loads); every load is issued before any is used
    const int total2 = total >> 1;  // nxp is even
    for (int base = 0; base < total2; base += 2 * kXThreads) {
        double v[2][2];
        int dst[2];
        long cell[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e0 = base + (int)threadIdx.x + u * kXThreads;
            const int e = 2 * min(e0, total2 - 1);  // first point of the pair
            const int l = (int)q.div_nxp.div((uint32_t)e);
            const int i = e - l * nxp;
            const int t = (int)q.div_t.div((uint32_t)l);
            const int k = k0 + (l - t * nt) + t * nz2;
            dst[u] = e0 < total2 ? l * ld + (i >> 1) : -1;  // complex slot of the pair
            cell[u] = i + j * pr_j + k * pr_k;
            if (MODE == 1) {
                const dbl2u r2 = ld2(arhs + cell[u]);
                v[u][0] = r2.x;
                v[u][1] = r2.y;
            } else if (MODE == 2) {
                // no contraction, so the single and batched kernels form the same value
#pragma clang fp contract(off)
                // rhs holds phi; index nxp (x) and nzp (z) hold the periodic duplicates of 0
                const auto ph = arhs + cell[u];
                const dbl2u p0 = ld2(ph);                           // phi(i), phi(i+1)
                const double pxm = ph[i == 0 ? nxp - 1 : -1];       // phi(i-1)
                const double pxp = ph[2];                           // phi(i+2)
                const dbl2u pzm = ld2(ph + (k == 0 ? (long)(g.nzp - 1) * pr_k : -pr_k));
                const dbl2u pzp = ld2(ph + pr_k);
                const dbl2u pym = j > 0 ? ld2(ph - pr_j) : dbl2u{0.0, 0.0};
                const dbl2u pyp = j < g.nym - 1 ? ld2(ph + pr_j) : dbl2u{0.0, 0.0};
                const dbl2u bb = ld2(ab + cell[u]);
                const double lo = j > 0 ? odl[j - 1] : 0.0;
                const double up = j < g.nym - 1 ? odu[j] : 0.0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const double c0 = h ? p0.y : p0.x;
                    const double xm = h ? p0.x : pxm;
                    const double xp = h ? pxp : p0.y;
                    double lap = (xp - 2.0 * c0 + xm) * a.idx2 +
                                 (lo * (h ? pym.y : pym.x) + odd[j] * c0 + up * (h ? pyp.y : pyp.x)) +
                                 ((h ? pzp.y : pzp.x) - 2.0 * c0 + (h ? pzm.y : pzm.x)) * a.idz2;
                    if (j == 0) lap += 0.5 * odd[0] * mean0;
                    v[u][h] = (h ? bb.y : bb.x) - lap;
                }
            } else {
                const long iu = (i + 1) + (j + 1) * pu_j + (k + 1) * pu_k;  // U(i+2,j+2,k+2)
                const long iv = (i + 1) + (j + 1) * pv_j + (k + 1) * pv_k;  // V(i+2,j+2,k+2)
                const long iw = (i + 1) + (j + 1) * pw_j + (k + 1) * pw_k;  // W(i+2,j+2,k+2)
                const dbl2u ua = ld2(aU + iu - 1);  // U(i+1), U(i+2)
                const double ub = aU[iu + 1];       // U(i+3)
                const dbl2u v1 = ld2(aV + iv), v0 = ld2(aV + iv - pv_j);
                const dbl2u w1 = ld2(aW + iw), w0 = ld2(aW + iw - pw_k);
                const double idz = inv_dzf[k];
                {
#pragma clang fp contract(off)
                    // projection.f90:55-57
                    v[u][0] = (ua.y - ua.x) * g.inv_dx + (v1.x - v0.x) * inv_dy + (w1.x - w0.x) * idz;
                    v[u][1] = (ub - ua.y) * g.inv_dx + (v1.y - v0.y) * inv_dy + (w1.y - w0.y) * idz;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (dst[u] >= 0) {
                A[dst[u]] = CplxT<C>::make((Rl)v[u][0], (Rl)v[u][1]);
                if (MODE != 2 && abo) st2(abo + cell[u], v[u][0], v[u][1], true, true);
                rsum += v[u][0];
                rsum += v[u][1];
            }
        }
    }
    if (a.rowpart) {
        const double sb = block_sum<kXThreads>(rsum, wsum);
        if (threadIdx.x == 0) gptr(a.rowpart)[(long)j * q.nbx + bx] = sb;
    }
    __syncthreads();
    // tiles of up to 2048 * KX points (radix-2 stages paired): 64-byte FP32 and 128-byte FP64 spectrum segments
    lds_fft_inplace<-1, KX, C, true>(A, nl, ld, q.fd, twl);

    // unpack the R2C modes of the R lines of each column, apply the first z stage
    for (int pidx = threadIdx.x; pidx < (M + 1) * nt; pidx += kXThreads) {
        const int kx = (int)q.div_t.div((uint32_t)pidx);
        const int c = pidx - kx * nt;
        C av[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const C* Z = A + (t * nt + c) * ld;
            const C zk = Z[kx == M ? 0 : kx];
            const C zm = Z[kx == 0 ? 0 : M - kx];
            const C E = CplxT<C>::make((Rl)0.5 * (zk.x + zm.x), (Rl)0.5 * (zk.y - zm.y));
            const C D = CplxT<C>::make(zk.x - zm.x, zk.y + zm.y);  // zk - conj(zm)
            const C O = CplxT<C>::make((Rl)0.5 * D.y, (Rl)-0.5 * D.x);  // D / (2i)
            av[t] = cadd(E, cmul(rtl[kx], O));
        }
        dft_reg<-1, R>(av);
        const int kp = k0 + c;
        const auto out = spec + ((long)kx * R * g.nym + j) * nz2 + kp;
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const C y = u == 0 ? av[0] : cmul(av[u], tzl[kp * (R - 1) + (u - 1)]);
            gst<C>(out + (long)u * g.nym * nz2, y);
        }
    }
}

template <int MODE, typename C, int R, int KX>
__global__ void __launch_bounds__(kXThreads) k_div_rfft_x(XPar<C> q) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    int bx, j;
    xtile(L, gridDim.x, gridDim.y, q.strip, bx, j);
    xfwd_body<MODE, C, R, KX>(q, bx, j, smem_raw);
}

// case of work item w: the last c with off[c] <= w (off[0] = 0, off[nc] = total)
__device__ __forceinline__ int xfind_case(const int* __restrict__ off, int nc, int w) {
    int lo = 0, hi = nc;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

// KX butterfly slots per thread (4: FP32 tiles too large for two);
// with two, at most four waves per SIMD (128 VGPRs): two workgroups per CU
template <int MODE, typename C, int R, int KX>
__global__ void __launch_bounds__(kXThreads) __attribute__((amdgpu_waves_per_eu(KX == 2 ? 4 : 1)))
kb_div_rfft_x(const XPar<C>* __restrict__ cs, const int* __restrict__ off, const int* __restrict__ idx,
              const int* __restrict__ pnc) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nc = *pnc;
    const int total = off[nc];
    const int me = xcd_remap(blockIdx.x, gridDim.x);
    for (int base = 0; base < total; base += gridDim.x) {
        const int w = base + me;
        if (w >= total) break;
        const int c = xfind_case(off, nc, w);
        const XPar<C>& q = cs[idx[c]];
        int bx, j;
        xtile(w - off[c], q.nbx, q.g.nym, q.strip, bx, j);
        xfwd_body<MODE, C, R, KX>(q, bx, j, smem_raw);
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
template <typename C, int ACCUM, int R>
__device__ __forceinline__ void xinv_body(const XPar<C>& q, int bx, int j, unsigned char* smem_raw) {
    __shared__ double wsum[kXThreads / 64];
    const Geo& g = q.g;
    const auto spec = gptr(q.spec);
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int M = q.fd.n;
    const int ld = M + 1;
    const int nxp = g.nxp;
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
    for (int e = threadIdx.x; e < ntw + (M + 1) + nz2 * (R - 1); e += kXThreads) twl[e] = gld<C>(gptr(q.tw) + e);
    __syncthreads();

    // gather the R sub-line values of each (kx, column), undo the first z stage;
    // a periodic y's duplicate row nym-1 reads row 0 (projection.f90:350-354)
    const int js = (g.ybc == 0 && j == g.nym - 1) ? 0 : j;
    const int npairs = (M + 1) * nt;
    constexpr int HP = R <= 4 ? 2 : 1;  // pairs per thread in flight
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
            const auto src = spec + ((long)kx * R * g.nym + js) * nz2 + k0 + c;
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
#pragma unroll
            for (int t = 0; t < R; ++t) A[(t * nt + c) * ld + kx] = av[h][t];
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
    lds_fft_inplace<1, 2, C, sizeof(C) == sizeof(float2)>(A, nl, ld, q.fd, twl);
    const long pr_j = (long)(g.nxg - 1), pr_k = (long)(g.nxg - 1) * g.nym;
    const long pp_j = (long)g.nxg, pp_k = (long)g.nxg * g.nyg;
    const auto rr = gptr(q.rhs) + (long)j * pr_j;
    const auto P = gptr(q.P);
    const auto pbar = gptr(q.pbar);
    const double scale = q.scale, dtc = q.dtc;
    const int save_p = q.save_p;
    double psum = 0.0;
    // two neighbouring points (i, i+1), i even, per lane: one complex slot of
    // the line, 16-byte stores of phi and of P's main copy
    for (int e2 = threadIdx.x; e2 < (nl * nxp) >> 1; e2 += kXThreads) {
        const int e = 2 * e2;
        const int l = (int)q.div_nxp.div((uint32_t)e);
        const int i = e - l * nxp;
        const int t = (int)q.div_t.div((uint32_t)l);
        const int k = k0 + (l - t * nt) + t * nz2;
        const C z = A[l * ld + (i >> 1)];
        double phi0 = (double)z.x * scale, phi1 = (double)z.y * scale;
        if (pbar) {
            phi0 += pbar[j];
            phi1 += pbar[j];
        }
        if (ACCUM) {
            const dbl2u o = ld2(rr + i + k * pr_k);
            phi0 += o.x;
            phi1 += o.y;
        }
        psum += phi0;
        psum += phi1;
        // rhs_p(i+2, j+2, k+2) and the duplicates (i = nxp -> nxg-1, k = nzp -> nzp+2;
        // z walls have no z duplicate, projection.f90:472-497)
        st2(rr + i + k * pr_k, phi0, phi1, true, true);
        if (i == 0) rr[nxp + k * pr_k] = phi0;
        if (k == 0 && g.zbc == 0) {
            st2(rr + i + (long)g.nzp * pr_k, phi0, phi1, true, true);
            if (i == 0) rr[nxp + (long)g.nzp * pr_k] = phi0;
        }
        if (save_p) {
            // every copy of these cells in P (reference 1-based indices; 0 = none):
            // x: i+2, 1 if i = nxp-1, nxg-1 if i = 0, nxg if i = 1 (i even: the
            //    first two for the second point, the third for the first)
            // y: j+2, 1 if j = 0, nyg if j = nym-1
            // z: k+2, 1 if k = nzp-1, nzg-1 if k = 0, nzg if k = 1 (mod nzp);
            //    z walls (Neumann, projection.f90:423-436): 1 if k = 0, nzg if k = nzp-1
            const double pv0 = phi0 / dtc, pv1 = phi1 / dtc;
            const int y1 = j == 0 ? 1 : 0, y2 = j == g.nym - 1 ? g.nyg : 0;
            const bool zw = g.zbc == 1;
            const int z1 = zw ? (k == 0 ? 1 : 0) : (k == g.nzp - 1 ? 1 : 0);
            const int z2 = zw ? (k == g.nzp - 1 ? g.nzg : 0) : (k == 0 ? g.nzg - 1 : 0);
            const int z3 = zw ? 0 : (k == 1 % g.nzp ? g.nzg : 0);
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
                    if (i == 0) Pr[g.nxg - 2] = pv0;          // x = nxg-1
                    if (i + 1 == nxp - 1) Pr[0] = pv1;        // x = 1
                    if (i + 1 == 1) Pr[g.nxg - 1] = pv1;      // x = nxg
                }
            }
        }
    }
    if (q.part0 && j == 0) {
        const double sblk = block_sum<kXThreads>(psum, wsum);
        if (threadIdx.x == 0) gptr(q.part0)[bx] = sblk;
    }
}

template <typename C, int ACCUM, int R>
__global__ void __launch_bounds__(kXThreads) k_irfft_x(XPar<C> q) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int L = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    int bx, j;
    xtile(L, gridDim.x, gridDim.y, q.strip, bx, j);
    xinv_body<C, ACCUM, R>(q, bx, j, smem_raw);
}

template <typename C, int ACCUM, int R>
__global__ void __launch_bounds__(kXThreads) kb_irfft_x(const XPar<C>* __restrict__ cs, const int* __restrict__ off,
                                                        const int* __restrict__ idx, const int* __restrict__ pnc) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nc = *pnc;
    const int total = off[nc];
    const int me = xcd_remap(blockIdx.x, gridDim.x);
    for (int base = 0; base < total; base += gridDim.x) {
        const int w = base + me;
        if (w >= total) break;
        const int c = xfind_case(off, nc, w);
        const XPar<C>& q = cs[idx[c]];
        int bx, j;
        xtile(w - off[c], q.nbx, q.g.nym, q.strip, bx, j);
        xinv_body<C, ACCUM, R>(q, bx, j, smem_raw);
        __syncthreads();
    }
}


// launches of the kernels above for one z-split radix R (instantiated per R
// in k_xfft_r<R>.hip, so the radices compile in parallel)
template <int R, typename C>
hipError_t launch_xfwd_r(hipStream_t s, int mode, const XPar<C>& q, dim3 grid, size_t lds) {
    auto go = [&](auto k) { hipLaunchKernelGGL(k, grid, dim3(kXThreads), lds, s, q); };
    if constexpr (sizeof(C) == sizeof(float2)) {
        if (xfwd_wide(q.fd, R, q.T)) {
            if (mode == 0) go(k_div_rfft_x<0, C, R, 4>);
            else if (mode == 1) go(k_div_rfft_x<1, C, R, 4>);
            else go(k_div_rfft_x<2, C, R, 4>);
            return hipGetLastError();
        }
    }
    if (mode == 0) {
        go(k_div_rfft_x<0, C, R, 2>);
    } else if (mode == 1) {
        go(k_div_rfft_x<1, C, R, 2>);
    } else {
        if constexpr (sizeof(C) == sizeof(float2)) {
            go(k_div_rfft_x<2, C, R, 2>);
        } else {
            return hipErrorInvalidValue;  // the residual form exists for the FP32 spectrum only
        }
    }
    return hipGetLastError();
}

template <int R, typename C>
hipError_t launch_xinv_r(hipStream_t s, int accum, const XPar<C>& q, dim3 grid, size_t lds) {
    if (!accum) {
        hipLaunchKernelGGL((k_irfft_x<C, 0, R>), grid, dim3(kXThreads), lds, s, q);
    } else {
        if constexpr (sizeof(C) == sizeof(float2)) {
            hipLaunchKernelGGL((k_irfft_x<C, 1, R>), grid, dim3(kXThreads), lds, s, q);
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

// mode: 0, 1, 2 as for the single launch, + 8 when a case of the list has
// forward tiles too large for two butterfly slots per thread (FP32 only)
template <int R, typename C>
hipError_t launch_xfwd_batch_r(hipStream_t s, int mode, const XPar<C>* cs, const int* off, const int* idx,
                               const int* nc, int grid, size_t lds) {
    auto go = [&](auto k) {
        hipLaunchKernelGGL(k, dim3(resident_grid(k, grid, kXThreads, lds)), dim3(kXThreads), lds, s, cs, off, idx, nc);
    };
    const bool wide = (mode & 8) != 0;
    mode &= 7;
    if constexpr (sizeof(C) == sizeof(float2)) {
        if (wide) {
            if (mode == 0) go(kb_div_rfft_x<0, C, R, 4>);
            else if (mode == 1) go(kb_div_rfft_x<1, C, R, 4>);
            else go(kb_div_rfft_x<2, C, R, 4>);
            return hipGetLastError();
        }
    }
    if (wide) return hipErrorInvalidValue;
    if (mode == 0) {
        go(kb_div_rfft_x<0, C, R, 2>);
    } else if (mode == 1) {
        go(kb_div_rfft_x<1, C, R, 2>);
    } else {
        if constexpr (sizeof(C) == sizeof(float2)) {
            go(kb_div_rfft_x<2, C, R, 2>);
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

template <int R, typename C>
hipError_t launch_xinv_batch_r(hipStream_t s, int accum, const XPar<C>* cs, const int* off, const int* idx,
                               const int* nc, int grid, size_t lds) {
    if (!accum) {
        const auto k = kb_irfft_x<C, 0, R>;
        hipLaunchKernelGGL(k, dim3(resident_grid(k, grid, kXThreads, lds)), dim3(kXThreads), lds, s, cs, off, idx, nc);
    } else {
        if constexpr (sizeof(C) == sizeof(float2)) {
            const auto k = kb_irfft_x<C, 1, R>;
            hipLaunchKernelGGL(k, dim3(resident_grid(k, grid, kXThreads, lds)), dim3(kXThreads), lds, s, cs, off, idx,
                               nc);
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace dfp

#define DFP_INSTANTIATE_XFFT(RV)                                                                                   \
    namespace dfp {                                                                                                \
    template hipError_t launch_xfwd_r<RV, double2>(hipStream_t, int, const XPar<double2>&, dim3, size_t);         \
    template hipError_t launch_xfwd_r<RV, float2>(hipStream_t, int, const XPar<float2>&, dim3, size_t);           \
    template hipError_t launch_xinv_r<RV, double2>(hipStream_t, int, const XPar<double2>&, dim3, size_t);         \
    template hipError_t launch_xinv_r<RV, float2>(hipStream_t, int, const XPar<float2>&, dim3, size_t);           \
    template hipError_t launch_xfwd_batch_r<RV, double2>(hipStream_t, int, const XPar<double2>*, const int*,      \
                                                         const int*, const int*, int, size_t);                                 \
    template hipError_t launch_xfwd_batch_r<RV, float2>(hipStream_t, int, const XPar<float2>*, const int*,        \
                                                        const int*, const int*, int, size_t);                                  \
    template hipError_t launch_xinv_batch_r<RV, double2>(hipStream_t, int, const XPar<double2>*, const int*,      \
                                                         const int*, const int*, int, size_t);                                 \
    template hipError_t launch_xinv_batch_r<RV, float2>(hipStream_t, int, const XPar<float2>*, const int*,        \
                                                        const int*, const int*, int, size_t);                                  \
    }
