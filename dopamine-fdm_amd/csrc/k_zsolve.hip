// Fused z-transform + wall-normal tridiagonal solve, one block per
// (kx, u) sub-plane of the split spectrum (see k_spectral.hip: the x kernel
// already applied the first radix-R z stage, so each sub-plane holds
// independent rows of length nz2 = nzp/R whose DFTs are the modes kz = R*m+u).
//
// The y system of every mode (kx,kz) is the one of projection.f90:277-297:
//   DL(j) = Dyy(j+1,j), DU(j) = Dyy(j,j+1), D(j) = Dyy(j,j) + (kxx + kzz),
//   D(2) *= 3/2 for the (0,0) mode (:292).
// Its LU factors depend only on kxx(kx) + kzz(kz), and kzz(kz) = kzz(nzp-kz)
// exactly, so the plan holds the inverse pivots 1/D'(j) of the elimination
// once per (kx, min(kz, nzp-kz)):  lu[(kx*nym + j)*nzs + lu_pos(kz, R, nz2)],
// nzs = nzp/2 + 1, the columns ordered so that each (kx, u) sub-plane reads
// one contiguous run per row (lu_pos in dfp_internal.h).  The elimination follows zgtsv's no-interchange branch
// (mult = DL/D', D' <- D - mult*DU); LAPACK interchanges rows only in a few
// low modes, which moves the solution at the 1e-15 level.
//
// k_zsolve, one workgroup per sub-plane (or P planes, below), two halves:
//   elimination: rows j = 0..nym-1 in groups of G: FFT of the rows (two
//   register passes with one LDS exchange for n = N1*N2, else the staged LDS
//   transform), then per mode b'(j) = b(j) - DL(j-1)/D'(j-1) * b'(j-1),
//   the rows written back to the spectrum;
//   back substitution: groups from the top, rows j = nym-1..0:
//   x(j) = (b'(j) - DU(j) x(j+1)) / D'(j), then the inverse FFT of the rows.
// One lane per mode runs the recurrence of its column of the group.  Each
// half reads and writes the sub-plane once (64 H bytes for both in FP64,
// 32 H in the FP32 form, whose z FFTs run in FP32 while the elimination
// stays FP64 on values rounded to FP32 only where stored).  While a group
// is transformed and swept, the next group's rows, pivots and coefficients
// are already loading into registers.
#include <algorithm>

#include "fft_lds.h"

namespace dfp {

// pin: the periodic x has the singular (0,0) mode, pinned by D(2) *= 3/2
// (projection.f90:292); the DCT-IV modes of the inflow/outflow x are all
// nonzero and nothing is pinned
__global__ void __launch_bounds__(256) k_lu_table(double* __restrict__ lu, Ops o, int nxh, int nym, int nzp, int R,
                                                  int pin) {
    const int nzs = nzp / 2 + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nxh * nzs) return;
    const int kx = (int)(t / nzs);
    const int kzs = (int)(t - (long)kx * nzs);
    const double lam = o.kxx[kx] + o.kzz[kzs];
    double dp = o.dd[0] + lam;
    if (pin && kx == 0 && kzs == 0) dp = 3.0 / 2.0 * dp;
    double* col = lu + (long)kx * nym * nzs + lu_pos(kzs, R, nzp / R);
    col[0] = 1.0 / dp;
    for (int j = 1; j < nym; ++j) {
        const double mult = o.dl[j - 1] / dp;
        dp = (o.dd[j] + lam) - mult * o.du[j - 1];
        col[(long)j * nzs] = 1.0 / dp;
    }
}

hipError_t launch_lu_table(dfp_plan* p) {
    const Geo& g = p->geo;
    const long n = (long)g.nxh * (g.nzp / 2 + 1);
    hipLaunchKernelGGL(k_lu_table, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, p->d_lu, p->ops, g.nxh, g.nym,
                       g.nzp, p->zr, g.xbc == 0 ? 1 : 0);
    return hipGetLastError();
}

constexpr int kZsThreads = 256;
constexpr int kZsPf = 8;     // group points per thread held in flight
constexpr int kZsAhead = 2;  // rows read ahead in the elimination sweeps

// Shared layout of one group (G rows of length n):
//   A[r*n + m]     the rows (the two-pass kernels hold element (i1, k2) of row
//                  r at zsw(r, i1, k2), a permutation within the row)
//   lus[r*n + m]   inverse pivot of mode m at row j0 + r
//   dus[r]         DL(j-1) (forward) or DU(j) (backward) of row j = j0 + r
// The prefetch registers are named variables (pr0..pr7, pl0..pl7, pd) so
// they are never placed in scratch memory; loads are unconditional (indices
// clamped into the group) so they stay in flight until staged.
#define ZS_LOAD1(K)                                                               \
    {                                                                             \
        const int e = min((int)threadIdx.x + (K) * kZsThreads, lim_);             \
        pr##K = gld<C>(src_ + e);                                                 \
        const int r = (int)div_n.div((uint32_t)e);                                \
        const int mm = e - r * n;                                                 \
        pl##K = lsrc_[(long)r * nzs + lucol[u * n + mm]];                         \
    }
#define ZS_PREFETCH(J0, GQ, COEF, COFF)                                           \
    do {                                                                          \
        const auto src_ = plane + (long)(J0) * n;                                 \
        const auto lsrc_ = lup + (long)(J0) * nzs;                                \
        const int lim_ = (GQ) * n - 1;                                            \
        ZS_LOAD1(0) ZS_LOAD1(1) ZS_LOAD1(2) ZS_LOAD1(3)                           \
        ZS_LOAD1(4) ZS_LOAD1(5) ZS_LOAD1(6) ZS_LOAD1(7)                           \
        const int jc = (J0) + min((int)threadIdx.x, (GQ) - 1) + (COFF);           \
        pd = COEF[max(0, min(jc, nym - 2))];                                      \
    } while (0)
#define ZS_STORE1(K)                                                              \
    {                                                                             \
        const int e = (int)threadIdx.x + (K) * kZsThreads;                        \
        if (e < np_) {                                                            \
            A[e] = pr##K;                                                         \
            lus[e] = pl##K;                                                       \
        }                                                                         \
    }
#define ZS_STAGE(GR)                                                              \
    do {                                                                          \
        const int np_ = (GR) * n;                                                 \
        ZS_STORE1(0) ZS_STORE1(1) ZS_STORE1(2) ZS_STORE1(3)                       \
        ZS_STORE1(4) ZS_STORE1(5) ZS_STORE1(6) ZS_STORE1(7)                       \
        if ((int)threadIdx.x < (GR)) dus[threadIdx.x] = pd;                       \
    } while (0)
#define ZS_REGS                                                                   \
    C pr0, pr1, pr2, pr3, pr4, pr5, pr6, pr7;                                     \
    double pl0, pl1, pl2, pl3, pl4, pl5, pl6, pl7;                                \
    double pd

// forward elimination of one group for mode m: b'(j) = b(j) - DL(j-1)/D'(j-1) * b'(j-1);
// at(r) is the LDS index of the mode in group row r, put(r, b') keeps the result.
// The carry (b'(j0-1), 1/D'(j0-1)) starts at zero, which leaves row 0 as it is.
// Rows, pivots and coefficients are read kZsAhead rows ahead (clamped, so the
// loads are unconditional) into a ring of registers whose slots are fixed by
// the unrolled loop: no copies wait on the loads in flight.
template <typename C, typename At, typename Put>
__device__ __forceinline__ void fwd_sweep(const C* A, const double* lus, const double* dus, int gr, int n, int m,
                                          double& bpx, double& bpy, double& ipp, At at, Put put) {
    using Rl = typename CplxT<C>::real;
    C ra[kZsAhead];
    double rl[kZsAhead], rd[kZsAhead];
#pragma unroll
    for (int q = 0; q < kZsAhead; ++q) {
        const int rr = min(q, gr - 1);
        ra[q] = A[at(rr)];
        rl[q] = lus[rr * n + m];
        rd[q] = dus[rr];
    }
    auto row = [&](int r, C& sa, double& sl, double& sd) {
        const C a = sa;
        const double lr = sl, dr = sd;
        const int rn = min(r + kZsAhead, gr - 1);
        sa = A[at(rn)];
        sl = lus[rn * n + m];
        sd = dus[rn];
        const double mult = dr * ipp;  // DL(j-1) / D'(j-1)
        const double bx = (double)a.x - mult * bpx;
        const double by = (double)a.y - mult * bpy;
        put(r, CplxT<C>::make((Rl)bx, (Rl)by));
        bpx = bx;
        bpy = by;
        ipp = lr;
    };
    int r0 = 0;
    for (; r0 + kZsAhead <= gr; r0 += kZsAhead) {
#pragma unroll
        for (int q = 0; q < kZsAhead; ++q) row(r0 + q, ra[q], rl[q], rd[q]);
    }
#pragma unroll
    for (int q = 0; q < kZsAhead; ++q)
        if (r0 + q < gr) row(r0 + q, ra[q], rl[q], rd[q]);
}

// back substitution of one group for mode m, rows from the top of the group:
// x(j) = (b'(j) - DU(j) x(j+1)) / D'(j), with x(nym) = 0 as the first carry;
// at(r) and the read-ahead as for fwd_sweep
template <typename C, typename At>
__device__ __forceinline__ void bwd_sweep(C* A, const double* lus, const double* dus, int gr, int n, int m,
                                          bool is00, double& xnx, double& xny, At at) {
    using Rl = typename CplxT<C>::real;
    C ra[kZsAhead];
    double rl[kZsAhead], rd[kZsAhead];
#pragma unroll
    for (int q = 0; q < kZsAhead; ++q) {
        const int rr = max(gr - 1 - q, 0);
        ra[q] = A[at(rr)];
        rl[q] = lus[rr * n + m];
        rd[q] = dus[rr];
    }
    auto row = [&](int t, C& sa, double& sl, double& sd) {
        const int r = gr - 1 - t;
        const C bb = sa;
        const double ip = sl, dr = sd;
        const int rn = max(gr - 1 - (t + kZsAhead), 0);
        sa = A[at(rn)];
        sl = lus[rn * n + m];
        sd = dus[rn];
        double xx = ((double)bb.x - dr * xnx) * ip;
        double xy = ((double)bb.y - dr * xny) * ip;
        if (is00) xx = xy = 0.0;
        A[at(r)] = CplxT<C>::make((Rl)xx, (Rl)xy);
        xnx = xx;
        xny = xy;
    };
    int t0 = 0;
    for (; t0 + kZsAhead <= gr; t0 += kZsAhead) {
#pragma unroll
        for (int q = 0; q < kZsAhead; ++q) row(t0 + q, ra[q], rl[q], rd[q]);
    }
#pragma unroll
    for (int q = 0; q < kZsAhead; ++q)
        if (t0 + q < gr) row(t0 + q, ra[q], rl[q], rd[q]);
}

// C: spectrum precision (double2, or float2 for the FP32 form).  The
// pivots, coefficients and the elimination itself are FP64 in both forms;
// the FP32 form rounds only what it stores.  b = (kx, u) sub-plane index.
template <typename C>
__device__ __forceinline__ void zfwd_body(const ZPar<C>& zp, int b, unsigned char* smem_raw) {
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int n = zp.fz.n;
    const int nym = zp.nym, nzp = zp.nzp, R = zp.R, G = zp.G, ntw = zp.ntw;
    const FastDiv div_n = zp.div_n;
    const auto dl = gptr(zp.dl);
    const int nzs = nzp / 2 + 1;
    const int kx = b / R;
    const int u = b - kx * R;
    C* A = smem;
    C* twl = A + G * n;
    double* lus = reinterpret_cast<double*>(twl + ntw);
    double* dus = lus + G * n;
    const auto plane = gptr(zp.spec) + (long)b * nym * n;
    const auto lup = gptr(zp.lu) + (long)kx * nym * nzs;
    const auto lucol = gptr(zp.lucol);
    for (int e = threadIdx.x; e < ntw; e += kZsThreads) twl[e] = gld<C>(gptr(zp.tw) + e);
    ZS_REGS;
    ZS_PREFETCH(0, min(G, nym), dl, -1);
    double bpx = 0.0, bpy = 0.0, ipp = 0.0;
    const int m = threadIdx.x;  // mode of this thread (m < n)
    for (int j0 = 0; j0 < nym; j0 += G) {
        const int gr = min(G, nym - j0);
        __syncthreads();
        ZS_STAGE(gr);
        __syncthreads();
        if (j0 + G < nym) ZS_PREFETCH(j0 + G, min(G, nym - j0 - G), dl, -1);
        lds_fft_inplace<-1, 2>(A, gr, n, zp.fz, twl);
        auto at = [&](int r) { return r * n + m; };
        if (m < n) fwd_sweep<C>(A, lus, dus, gr, n, m, bpx, bpy, ipp, at, [&](int r, C v) { A[at(r)] = v; });
        __syncthreads();
        const auto dst = plane + (long)j0 * n;
        for (int e = threadIdx.x; e < gr * n; e += kZsThreads) gst<C>(dst + e, A[e]);
    }
}

// zero00: the (0,0) mode is left out (set to zero); the FP32 form solves it
// in FP64 separately
template <typename C>
__device__ __forceinline__ void zbwd_body(const ZPar<C>& zp, int b, unsigned char* smem_raw) {
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int n = zp.fz.n;
    const int nym = zp.nym, nzp = zp.nzp, R = zp.R, G = zp.G, ntw = zp.ntw;
    const FastDiv div_n = zp.div_n;
    const auto du = gptr(zp.du);
    const int nzs = nzp / 2 + 1;
    const int kx = b / R;
    const int u = b - kx * R;
    C* A = smem;
    C* twl = A + G * n;
    double* lus = reinterpret_cast<double*>(twl + ntw);
    double* dus = lus + G * n;
    const auto plane = gptr(zp.spec) + (long)b * nym * n;
    const auto lup = gptr(zp.lu) + (long)kx * nym * nzs;
    const auto lucol = gptr(zp.lucol);
    for (int e = threadIdx.x; e < ntw; e += kZsThreads) twl[e] = gld<C>(gptr(zp.tw) + e);
    // groups from the top: [j0, j0+gr), j0 = max(0, nym - (gi+1)*G)
    const int ngroups = (nym + G - 1) / G;
    ZS_REGS;
    {
        const int j0q = max(0, nym - G);
        ZS_PREFETCH(j0q, nym - j0q, du, 0);
    }
    double xnx = 0.0, xny = 0.0;
    const int m = threadIdx.x;
    const bool is00 = zp.zero00 && kx == 0 && u == 0 && m == 0;
    for (int gi = 0; gi < ngroups; ++gi) {
        const int j0 = max(0, nym - (gi + 1) * G);
        const int gr = nym - gi * G - j0;
        __syncthreads();
        ZS_STAGE(gr);
        __syncthreads();
        if (gi + 1 < ngroups) {
            const int j0q = max(0, nym - (gi + 2) * G);
            ZS_PREFETCH(j0q, j0 - j0q, du, 0);
        }
        if (m < n) bwd_sweep<C>(A, lus, dus, gr, n, m, is00, xnx, xny, [&](int r) { return r * n + m; });
        __syncthreads();
        lds_fft_inplace<1, 2>(A, gr, n, zp.fz, twl);
        const auto dst = plane + (long)j0 * n;
        for (int e = threadIdx.x; e < gr * n; e += kZsThreads) gst<C>(dst + e, A[e]);
    }
}

// Two-pass z transforms for n = N1*N2 (N2 = 4 or 8).  Forward: the thread
// of item (row r, i1) loads x[i1 + N1*i2], i2 < N2, straight from HBM into
// registers (the next group's items load while this group is swept),
// transforms them (length N2), applies w_n^{i1*k2} and writes Y[r][i1*N2 + k2]
// to LDS; the item (r, k2) then transforms Y[r][k2 + N2*i1] over i1 (length
// N1) in place, which leaves X[r][k2 + N2*k1] in natural order.  Inverse: the
// same two steps in the other order, the second one storing x[i1 + N1*i2]
// from registers to HBM.  Both replace the log(n) LDS passes of the staged
// transform by one exchange through LDS.
// LDS position of element (i1, k2) of row r in the two-pass layouts (i1*N2 + k2
// unswizzled): the N2 elements of each i1 are rotated by i1 and, for N2 = 8
// and rows of a multiple of 16 points, odd rows swap 8-element halves, so the
// 16 lanes of an LDS pass (FP64) hit distinct banks in both passes and the
// sweep; the same map serves the Y and X layouts, so both passes stay in place
template <int N1, int N2>
__device__ __forceinline__ int zsw(int r, int i1, int k2) {
    constexpr int N = N1 * N2;
    int p = i1 * N2 + (k2 + i1) % N2;
    if constexpr (N2 == 8 && N % 16 == 0) p ^= (r & 1) << 3;
    return r * N + p;
}

// Rows of 24 or 32 points: a block takes P = 2 consecutive sub-planes, so the
// sweeping wave has a mode in (nearly) every lane and a launch needs half the
// blocks.  Group row r of plane p is the virtual row v = r*P + p (n points,
// LDS layout of the two-pass transforms); lane t = p*n + m sweeps mode m of
// plane p.  A partial last block loads its last plane again and stores
// nothing of it.  Shorter rows keep one plane per block: with P = 64/n their
// blocks hold more groups of fewer rows, and a launch covers fewer CUs; at
// 8, 12 and 16 points that measured slower, while 24 and 32 gained.
__host__ __device__ constexpr int zplanes(int n) { return n > 16 && n <= 32 ? 64 / n : 1; }

// Point k of a thread's share of a group of gq rows, each row holding N
// points of each of P planes: row r, plane p, position m.  Points past the
// group are clamped to its last (loaded, never staged).
template <int N, int P>
__device__ __forceinline__ void zpoint(int k, int gq, int& r, int& p, int& m) {
    const int e = min((int)threadIdx.x + k * kZsThreads, gq * P * N - 1);
    const int v = e / N;
    m = e - v * N;
    r = v / P;
    p = v - r * P;
}

// COLS: form the pivot column of a one-plane row that divides the threads in
// registers (the single-case kernels) instead of reading the plan's column
// table (the batched kernels, where the extra registers cost a wave)
template <typename C, int N1, int N2, bool COLS>
__device__ __forceinline__ void zfwd2_body(const ZPar<C>& zp, int blk, unsigned char* smem_raw) {
    constexpr int N = N1 * N2;
    constexpr int P = zplanes(N);
    constexpr int KI = kZsPf / N2;  // first-step items per thread and group
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int nym = zp.nym, nzp = zp.nzp, R = zp.R, G = zp.G;
    const auto dl = gptr(zp.dl);
    const int nzs = nzp / 2 + 1;
    const int b0 = blk * P;
    C* A = smem;
    C* twl = A + G * P * N;
    double* lus = reinterpret_cast<double*>(twl + N);
    double* dus = lus + G * P * N;
    const auto spec = gptr(zp.spec);
    const auto lu = gptr(zp.lu);
    auto plane_of = [&](int p) { return P == 1 ? b0 : min(b0 + p, zp.nplanes - 1); };
    for (int e = threadIdx.x; e < N; e += kZsThreads) twl[e] = gld<C>(gptr(zp.tw2) + e);
    C pr[KI][N2];
    double pl[kZsPf];
    double pd;
    // offset of row 0 of the pivots of position mm of plane p
    // one plane whose rows divide the threads: every item of a thread has the
    // mode threadIdx.x % N, whose column is formed in registers (clamped items
    // of a short last group read that column of their row, unused); else the
    // plan's column table
    auto lu_col = [&](int p, int mm) {
        const int bp = plane_of(p), kx = bp / R;
        if constexpr (COLS && P == 1 && kZsThreads % N == 0)  // every item has mode threadIdx.x % N
            return (long)kx * nym * nzs + lu_pos_mu((int)threadIdx.x % N, bp - kx * R, R, N);
        return (long)kx * nym * nzs + gptr(zp.lucol)[(bp - kx * R) * N + mm];
    };
    auto prefetch = [&](int j0, int gq) {
#pragma unroll
        for (int k = 0; k < KI; ++k) {
            int r, p, i1;
            zpoint<N1, P>(k, gq, r, p, i1);
            const auto src = spec + ((long)plane_of(p) * nym + j0 + r) * N + i1;
#pragma unroll
            for (int i2 = 0; i2 < N2; ++i2) pr[k][i2] = gld<C>(src + N1 * i2);
        }
#pragma unroll
        for (int k = 0; k < kZsPf; ++k) {
            int r, p, mm;
            zpoint<N, P>(k, gq, r, p, mm);
            pl[k] = lu[lu_col(p, mm) + (long)(j0 + r) * nzs];
        }
        const int jc = j0 + min((int)threadIdx.x, gq - 1) - 1;  // DL(j-1) of row j
        pd = dl[max(0, min(jc, nym - 2))];
    };
    prefetch(0, min(G, nym));
    double bpx = 0.0, bpy = 0.0, ipp = 0.0;
    const int tp = (int)threadIdx.x / N, m = (int)threadIdx.x - tp * N;  // sweeping lane: plane, mode
    const bool sweeps = tp < P;
    const bool stores = sweeps && (P == 1 || b0 + tp < zp.nplanes);
    for (int j0 = 0; j0 < nym; j0 += G) {
        const int gr = min(G, nym - j0);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KI; ++k) {
            const int it = (int)threadIdx.x + k * kZsThreads;
            if (it < gr * P * N1) {
                const int v = it / N1, i1 = it - v * N1;
                C w[N2];
#pragma unroll
                for (int i2 = 0; i2 < N2; ++i2) w[i2] = pr[k][i2];
                dft_reg<-1, N2>(w);
                A[zsw<N1, N2>(v, i1, 0)] = w[0];
#pragma unroll
                for (int k2 = 1; k2 < N2; ++k2) A[zsw<N1, N2>(v, i1, k2)] = twiddle<-1>(w[k2], twl[i1 * k2]);
            }
        }
#pragma unroll
        for (int k = 0; k < kZsPf; ++k) {
            const int e = (int)threadIdx.x + k * kZsThreads;
            if (e < gr * P * N) lus[e] = pl[k];
        }
        if ((int)threadIdx.x < gr) dus[threadIdx.x] = pd;
        __syncthreads();
        if (j0 + G < nym) prefetch(j0 + G, min(G, nym - j0 - G));
        for (int it = threadIdx.x; it < gr * P * N2; it += kZsThreads) {
            const int v = it / N2, k2 = it - v * N2;
            C w[N1];
#pragma unroll
            for (int i1 = 0; i1 < N1; ++i1) w[i1] = A[zsw<N1, N2>(v, i1, k2)];
            dft_reg<-1, N1>(w);
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) A[zsw<N1, N2>(v, k1, k2)] = w[k1];
        }
        __syncthreads();
        // the lanes of the sweep store their rows: N contiguous modes per row
        if (sweeps) {
            const auto dst = spec + ((long)plane_of(tp) * nym + j0) * N + m;
            fwd_sweep<C>(
                A, lus + tp * N, dus, gr, P * N, m, bpx, bpy, ipp,
                [&](int r) { return zsw<N1, N2>(r * P + tp, m / N2, m % N2); },
                [&](int r, C w) {
                    if (stores) gst<C>(dst + r * N, w);
                });
        }
    }
}

// COLS: form the pivot column of a one-plane row that divides the threads in
// registers (the single-case kernels) instead of reading the plan's column
// table (the batched kernels, where the extra registers cost a wave)
template <typename C, int N1, int N2, bool COLS>
__device__ __forceinline__ void zbwd2_body(const ZPar<C>& zp, int blk, unsigned char* smem_raw) {
    constexpr int N = N1 * N2;
    constexpr int P = zplanes(N);
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int nym = zp.nym, nzp = zp.nzp, R = zp.R, G = zp.G;
    const auto du = gptr(zp.du);
    const int nzs = nzp / 2 + 1;
    const int b0 = blk * P;
    C* A = smem;
    C* twl = A + G * P * N;
    double* lus = reinterpret_cast<double*>(twl + N);
    double* dus = lus + G * P * N;
    const auto spec = gptr(zp.spec);
    const auto lu = gptr(zp.lu);
    auto plane_of = [&](int p) { return P == 1 ? b0 : min(b0 + p, zp.nplanes - 1); };
    for (int e = threadIdx.x; e < N; e += kZsThreads) twl[e] = gld<C>(gptr(zp.tw2) + e);
    const int ngroups = (nym + G - 1) / G;
    C pr[kZsPf];
    double pl[kZsPf];
    double pd;
    // one plane whose rows divide the threads: every item of a thread has the
    // mode threadIdx.x % N, whose column is formed in registers (clamped items
    // of a short last group read that column of their row, unused); else the
    // plan's column table
    auto lu_col = [&](int p, int mm) {
        const int bp = plane_of(p), kx = bp / R;
        if constexpr (COLS && P == 1 && kZsThreads % N == 0)  // every item has mode threadIdx.x % N
            return (long)kx * nym * nzs + lu_pos_mu((int)threadIdx.x % N, bp - kx * R, R, N);
        return (long)kx * nym * nzs + gptr(zp.lucol)[(bp - kx * R) * N + mm];
    };
    // rows j0..j0+gq-1 of the block's planes
    auto prefetch = [&](int j0, int gq) {
#pragma unroll
        for (int k = 0; k < kZsPf; ++k) {
            int r, p, mm;
            zpoint<N, P>(k, gq, r, p, mm);
            pr[k] = gld<C>(spec + ((long)plane_of(p) * nym + j0 + r) * N + mm);
            pl[k] = lu[lu_col(p, mm) + (long)(j0 + r) * nzs];
        }
        const int jc = j0 + min((int)threadIdx.x, gq - 1);  // DU(j) of row j
        pd = du[max(0, min(jc, nym - 2))];
    };
    {
        const int j0q = max(0, nym - G);
        prefetch(j0q, nym - j0q);
    }
    double xnx = 0.0, xny = 0.0;
    const int tp = (int)threadIdx.x / N, m = (int)threadIdx.x - tp * N;  // sweeping lane: plane, mode
    const bool sweeps = tp < P;
    const int bt = plane_of(tp);
    const bool is00 = zp.zero00 && sweeps && bt == 0 && m == 0;  // kx = 0, u = 0
    for (int gi = 0; gi < ngroups; ++gi) {
        const int j0 = max(0, nym - (gi + 1) * G);
        const int gr = nym - gi * G - j0;
        __syncthreads();
        // stage rows (natural order in HBM) into the swizzled layout
#pragma unroll
        for (int k = 0; k < kZsPf; ++k) {
            const int e = (int)threadIdx.x + k * kZsThreads;
            if (e < gr * P * N) {
                const int v = e / N, mm = e - v * N;
                A[zsw<N1, N2>(v, mm / N2, mm % N2)] = pr[k];
                lus[e] = pl[k];
            }
        }
        if ((int)threadIdx.x < gr) dus[threadIdx.x] = pd;
        __syncthreads();
        if (gi + 1 < ngroups) {
            const int j0q = max(0, nym - (gi + 2) * G);
            prefetch(j0q, j0 - j0q);
        }
        if (sweeps)
            bwd_sweep<C>(A, lus + tp * N, dus, gr, P * N, m, is00, xnx, xny,
                         [&](int r) { return zsw<N1, N2>(r * P + tp, m / N2, m % N2); });
        __syncthreads();
        for (int it = threadIdx.x; it < gr * P * N2; it += kZsThreads) {
            const int v = it / N2, k2 = it - v * N2;
            C w[N1];
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) w[k1] = A[zsw<N1, N2>(v, k1, k2)];
            dft_reg<1, N1>(w);
            A[zsw<N1, N2>(v, 0, k2)] = w[0];
#pragma unroll
            for (int i1 = 1; i1 < N1; ++i1) A[zsw<N1, N2>(v, i1, k2)] = twiddle<1>(w[i1], twl[i1 * k2]);
        }
        __syncthreads();
        for (int it = threadIdx.x; it < gr * P * N1; it += kZsThreads) {
            const int v = it / N1, i1 = it - v * N1;
            const int r = v / P, p = v - r * P;
            C w[N2];
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) w[k2] = A[zsw<N1, N2>(v, i1, k2)];
            dft_reg<1, N2>(w);
            if (P == 1 || b0 + p < zp.nplanes) {
                const auto dst = spec + ((long)(b0 + p) * nym + j0 + r) * N + i1;
#pragma unroll
                for (int i2 = 0; i2 < N2; ++i2) gst<C>(dst + N1 * i2, w[i2]);
            }
        }
    }
}

// the two-pass factorizations (N1, N2) with N1 * N2 = nzp / R; others use the
// staged transform
#define DFP_ZV_LIST(X) X(2, 4) X(3, 4) X(2, 8) X(3, 8) X(4, 8) X(9, 4) X(6, 8) X(8, 8) X(9, 8) X(12, 8) X(16, 8)

bool z_two_pass(int n, int& n1, int& n2) {
#define DFP_ZV_TRY(a, c) \
    if (a * c == n) { n1 = a; n2 = c; return true; }
    DFP_ZV_LIST(DFP_ZV_TRY)
#undef DFP_ZV_TRY
    n1 = n2 = 0;
    return false;
}

// Both halves in one launch: the workgroup that eliminates a sub-plane
// (writing its rows back) then substitutes back through the same sub-plane,
// top group first, i.e. the rows it stored last.  Those reads follow their
// writes by the duration of one workgroup rather than one launch, so they
// are served from the caches in front of HBM, and the rows overwritten by
// the back substitution are written to HBM once.  The barrier between the
// halves orders the stores of every lane before the loads of every other
// lane of the workgroup.  N1 = N2 = 0: the staged transform.
template <typename C, int N1, int N2>
__global__ void __launch_bounds__(kZsThreads) k_zsolve(ZPar<C> zp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if constexpr (N1 == 0) {
        zfwd_body<C>(zp, b, smem_raw);
        __syncthreads();
        zbwd_body<C>(zp, b, smem_raw);
    } else {
        zfwd2_body<C, N1, N2, true>(zp, b, smem_raw);
        __syncthreads();
        zbwd2_body<C, N1, N2, true>(zp, b, smem_raw);
    }
}

__device__ __forceinline__ int zfind_case(const int* __restrict__ off, int nc, int w) {
    int lo = 0, hi = nc;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename C, int N1, int N2>
__global__ void __launch_bounds__(kZsThreads) kb_zsolve(const ZPar<C>* __restrict__ cs, const int* __restrict__ off,
                                                        const int* __restrict__ idx, const int* __restrict__ pnc) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nc = *pnc;
    const int total = off[nc];
    const int me = xcd_remap(blockIdx.x, gridDim.x);
    for (int base = 0; base < total; base += gridDim.x) {
        const int w = base + me;
        if (w >= total) break;
        const int c = zfind_case(off, nc, w);
        // both halves of the sub-plane, as k_zsolve
        if constexpr (N1 == 0) {
            zfwd_body<C>(cs[idx[c]], w - off[c], smem_raw);
            __syncthreads();
            zbwd_body<C>(cs[idx[c]], w - off[c], smem_raw);
        } else {
            zfwd2_body<C, N1, N2, false>(cs[idx[c]], w - off[c], smem_raw);
            __syncthreads();
            zbwd2_body<C, N1, N2, false>(cs[idx[c]], w - off[c], smem_raw);
        }
        __syncthreads();
    }
}

template <typename C>
ZPar<C> make_zpar(const dfp_plan* p, int f32_zero00) {
    ZPar<C> z{};
    const Geo& g = p->geo;
    z.spec = reinterpret_cast<C*>(p->d_spec);
    z.lu = p->d_lu;
    z.lucol = p->d_lucol;
    z.dl = p->ops.dl;
    z.du = p->ops.du;
    z.fz = p->fz;
    z.tw = sizeof(C) == sizeof(double2) ? reinterpret_cast<const C*>(p->tw_z) : reinterpret_cast<const C*>(p->tw_z32);
    z.nym = g.nym;
    z.nzp = g.nzp;
    z.R = p->zr;
    z.ntw = twiddle_count(p->fz);
    z.n1 = p->zfused ? 0 : p->zn1;
    z.n2 = p->zfused ? 0 : p->zn2;
    z.tw2 = sizeof(C) == sizeof(double2) ? reinterpret_cast<const C*>(p->tw_z2) : reinterpret_cast<const C*>(p->tw_z2_32);
    if (z.n1) z.ntw = p->fz.n;
    z.G = p->g_zs;
    z.zero00 = f32_zero00;
    z.nplanes = g.nxh * p->zr;
    z.nblocks = ceil_div(z.nplanes, z.n1 ? zplanes(p->fz.n) : 1);
    z.div_n = FastDiv(p->fz.n);
    return z;
}
template ZPar<double2> make_zpar<double2>(const dfp_plan*, int);
template ZPar<float2> make_zpar<float2>(const dfp_plan*, int);

size_t z_lds_bytes(const dfp_plan* p, int f32) {
    const int G = p->g_zs, n = p->fz.n, ntw = p->zn1 ? n : twiddle_count(p->fz);
    const size_t pts = (size_t)G * n * (p->zn1 ? zplanes(n) : 1);
    return (f32 ? sizeof(float2) : sizeof(double2)) * (pts + ntw) + sizeof(double) * (pts + G);
}

int z_variant(const dfp_plan* p) { return p->zfused ? -1 : p->zn1 * 32 + p->zn2; }

template <typename C>
static hipError_t launch_zsolve_t(dfp_plan* p, const ZPar<C>& zp) {
    const bool f32 = sizeof(C) == 8;
    const size_t lds = z_lds_bytes(p, f32);
    const int slot = prof_begin(p, f32 ? "zsolve32" : "zsolve");
    const dim3 grid(zp.nblocks), blk(kZsThreads);
    switch (z_variant(p)) {
#define DFP_ZV_CASE(a, c)                                                          \
    case a * 32 + c:                                                               \
        hipLaunchKernelGGL((k_zsolve<C, a, c>), grid, blk, lds, p->stream, zp);    \
        break;
        DFP_ZV_LIST(DFP_ZV_CASE)
#undef DFP_ZV_CASE
        default:
            hipLaunchKernelGGL((k_zsolve<C, 0, 0>), grid, blk, lds, p->stream, zp);
    }
    prof_end(p, slot);
    return hipGetLastError();
}

hipError_t launch_zsolve(dfp_plan* p) { return launch_zsolve_t<double2>(p, make_zpar<double2>(p, 0)); }

hipError_t launch_zsolve32(dfp_plan* p) { return launch_zsolve_t<float2>(p, make_zpar<float2>(p, 1)); }

// variant: z_variant of the cases of the list (all alike)
template <typename C>
hipError_t launch_zsolve_batch(hipStream_t s, int variant, const ZPar<C>* cs, const int* off, const int* idx,
                               const int* nc, int grid, size_t lds) {
    const dim3 blk(kZsThreads);
    auto go = [&](auto k) {
        hipLaunchKernelGGL(k, dim3(resident_grid(k, grid, kZsThreads, lds)), blk, lds, s, cs, off, idx, nc);
    };
    switch (variant) {
#define DFP_ZV_CASE(a, c)          \
    case a * 32 + c:               \
        go(kb_zsolve<C, a, c>);    \
        break;
        DFP_ZV_LIST(DFP_ZV_CASE)
#undef DFP_ZV_CASE
        default:
            go(kb_zsolve<C, 0, 0>);
    }
    return hipGetLastError();
}
template hipError_t launch_zsolve_batch<double2>(hipStream_t, int, const ZPar<double2>*, const int*, const int*,
                                                 const int*, int, size_t);
template hipError_t launch_zsolve_batch<float2>(hipStream_t, int, const ZPar<float2>*, const int*, const int*,
                                                const int*, int, size_t);

// rows per group: the group (its rows in each of the block's planes) fits
// the prefetch registers (kZsPf points per thread), the in-place FFT capacity
// (2 butterflies per thread; not a limit of the two-pass transforms) and,
// with the twiddles and pivots, the device's LDS per block; 0 if one row does
// not fit
int zsolve_rows_per_group(const FftDesc& fz, int nym, size_t lds_limit) {
    const int n = fz.n;
    if (n > kZsThreads) return 0;  // one mode per thread
    int n1, n2;
    const bool two = z_two_pass(n, n1, n2);
    const size_t ntw = two ? (size_t)n : (size_t)twiddle_count(fz);
    const int pn = n * (two ? zplanes(n) : 1);  // points of one group row
    int G = (two ? kZsPf * kZsThreads : std::min(kZsPf * kZsThreads, inplace_capacity(fz, kZsThreads, 2))) / pn;
    G = std::max(0, std::min(G, nym));
    auto bytes = [&](int g) {
        return sizeof(double2) * ((size_t)g * pn + ntw) + sizeof(double) * ((size_t)g * pn + g);
    };
    while (G > 0 && bytes(G) > lds_limit) --G;
    return G;
}

}  // namespace dfp
