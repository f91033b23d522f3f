// Fused z-transform + wall-normal solve, whole sub-plane in LDS: one
// workgroup per (kx, u) sub-plane of the split spectrum (nym rows of
// n = nzp/R modes, the rows' first radix-R z stage already applied by the x
// kernel) loads the plane once, transforms every row along z, eliminates and
// back-substitutes every mode along y (projection.f90:277-297, zgtsv's
// no-interchange branch with the plan's inverse pivots), transforms back and
// stores the plane once: 32 H bytes in FP64 where the split pair of kernels
// (k_zsolve.hip) moves 64 H.  It needs nym * n complex values in LDS, so the
// plan picks a z-split radix R that makes the plane fit (R = 16 at
// 256x192x256: 49 KB, three workgroups per CU).
//
// The arithmetic is that of k_zsolve_fwd/bwd operation for operation (the
// eliminated rows b' are held in LDS at the spectrum's precision, as the
// split kernels hold them in HBM), so for the same R both paths agree bit
// for bit.
#include "fft_lds.h"

namespace dfp {

constexpr int kZfThreads = 256;
constexpr int kZfAhead = 8;  // pivot rows read ahead in the sweeps (global memory)
constexpr int kZfLds = 2;    // plane rows and coefficients read ahead (LDS); divides kZfAhead

// rows of the plane through the in-place LDS FFT, zp.G rows per call (the
// per-call capacity of two butterflies per thread)
template <int SIGN, typename C>
__device__ __forceinline__ void zf_rows_fft(C* A, const ZPar<C>& zp, const C* twl) {
    const int n = zp.fz.n;
    for (int l0 = 0; l0 < zp.nym; l0 += zp.G) lds_fft_inplace<SIGN, 2>(A + (long)l0 * n, min(zp.G, zp.nym - l0), n, zp.fz, twl);
}

template <typename C>
__device__ __forceinline__ void zfused_body(const ZPar<C>& zp, int b, unsigned char* smem_raw) {
    using Rl = typename CplxT<C>::real;
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int n = zp.fz.n;
    const int nym = zp.nym, nzp = zp.nzp, R = zp.R, ntw = zp.ntw;
    const int nzs = nzp / 2 + 1;
    const int kx = b / R;
    const int u = b - kx * R;
    C* A = smem;                                             // nym * n
    C* twl = A + (long)nym * n;                              // ntw
    double* cl = reinterpret_cast<double*>(twl + ntw);       // cl[j] = DL(j-1), cl[0] = 0
    double* cu = cl + nym;                                   // cu[j] = DU(j), cu[nym-1] = 0
    const auto plane = gptr(zp.spec) + (long)b * nym * n;
    const int total = nym * n;
    // load the plane: four values per thread in flight per batch (named
    // registers, clamped unconditional loads)
    for (int base = 0; base < total; base += 4 * kZfThreads) {
        const int e0 = base + (int)threadIdx.x;
        const C v0 = gld<C>(plane + min(e0, total - 1));
        const C v1 = gld<C>(plane + min(e0 + kZfThreads, total - 1));
        const C v2 = gld<C>(plane + min(e0 + 2 * kZfThreads, total - 1));
        const C v3 = gld<C>(plane + min(e0 + 3 * kZfThreads, total - 1));
        if (e0 < total) A[e0] = v0;
        if (e0 + kZfThreads < total) A[e0 + kZfThreads] = v1;
        if (e0 + 2 * kZfThreads < total) A[e0 + 2 * kZfThreads] = v2;
        if (e0 + 3 * kZfThreads < total) A[e0 + 3 * kZfThreads] = v3;
    }
    const auto tw = gptr(zp.tw);
    const auto dl = gptr(zp.dl);
    const auto du = gptr(zp.du);
    for (int e = threadIdx.x; e < ntw; e += kZfThreads) twl[e] = gld<C>(tw + e);
    for (int j = threadIdx.x; j < nym; j += kZfThreads) {
        cl[j] = j > 0 ? dl[j - 1] : 0.0;
        cu[j] = j < nym - 1 ? du[j] : 0.0;
    }
    __syncthreads();
    zf_rows_fft<-1>(A, zp, twl);

    const int m = threadIdx.x;  // mode of this thread (m < n)
    if (m < n) {
        const int kz = R * m + u;
        const auto lcol = gptr(zp.lu) + (long)kx * nym * nzs + min(kz, nzp - kz);
        // forward elimination: b'(j) = b(j) - DL(j-1)/D'(j-1) * b'(j-1).
        // Pivots (global memory) are read kZfAhead rows ahead, rows and
        // coefficients (LDS) kZfLds rows ahead; loads are unconditional with
        // clamped indices (the clamped reads are never used).
        C ra[kZfLds];
        double rc[kZfLds], rl[kZfAhead];
#pragma unroll
        for (int q = 0; q < kZfAhead; ++q) rl[q] = lcol[(long)min(q, nym - 1) * nzs];
#pragma unroll
        for (int q = 0; q < kZfLds; ++q) {
            const int jr = min(q, nym - 1);
            ra[q] = A[jr * n + m];
            rc[q] = cl[jr];
        }
        double bpx = 0.0, bpy = 0.0, ipp = 0.0;
        for (int j0 = 0; j0 < nym; j0 += kZfAhead) {
#pragma unroll
            for (int q = 0; q < kZfAhead; ++q) {
                const int j = j0 + q;
                if (j < nym) {
                    const int qa = q % kZfLds;
                    const C a = ra[qa];
                    const double ip = rl[q], dlj = rc[qa];
                    rl[q] = lcol[(long)min(j + kZfAhead, nym - 1) * nzs];
                    const int jn = min(j + kZfLds, nym - 1);
                    ra[qa] = A[jn * n + m];
                    rc[qa] = cl[jn];
                    double bx = (double)a.x, by = (double)a.y;
                    if (j > 0) {
                        const double mult = dlj * ipp;
                        bx = bx - mult * bpx;
                        by = by - mult * bpy;
                    }
                    A[j * n + m] = CplxT<C>::make((Rl)bx, (Rl)by);
                    bpx = bx;
                    bpy = by;
                    ipp = ip;
                }
            }
        }
        // back substitution: x(nym-1) = b'/D', x(j) = (b'(j) - DU(j) x(j+1)) / D'(j)
#pragma unroll
        for (int q = 0; q < kZfAhead; ++q) rl[q] = lcol[(long)max(nym - 1 - q, 0) * nzs];
#pragma unroll
        for (int q = 0; q < kZfLds; ++q) {
            const int jr = max(nym - 1 - q, 0);
            ra[q] = A[jr * n + m];
            rc[q] = cu[jr];
        }
        const bool is00 = zp.zero00 && kx == 0 && u == 0 && m == 0;
        double xnx = 0.0, xny = 0.0;
        for (int t0 = 0; t0 < nym; t0 += kZfAhead) {
#pragma unroll
            for (int q = 0; q < kZfAhead; ++q) {
                const int t = t0 + q;
                if (t < nym) {
                    const int j = nym - 1 - t;
                    const int qa = q % kZfLds;
                    const C bb = ra[qa];
                    const double ip = rl[q], duj = rc[qa];
                    rl[q] = lcol[(long)max(j - kZfAhead, 0) * nzs];
                    const int jn = max(j - kZfLds, 0);
                    ra[qa] = A[jn * n + m];
                    rc[qa] = cu[jn];
                    double xx, xy;
                    if (j == nym - 1) {
                        xx = (double)bb.x * ip;
                        xy = (double)bb.y * ip;
                    } else {
                        xx = ((double)bb.x - duj * xnx) * ip;
                        xy = ((double)bb.y - duj * xny) * ip;
                    }
                    if (is00) xx = xy = 0.0;
                    A[j * n + m] = CplxT<C>::make((Rl)xx, (Rl)xy);
                    xnx = xx;
                    xny = xy;
                }
            }
        }
    }
    __syncthreads();
    zf_rows_fft<1>(A, zp, twl);
    for (int e = threadIdx.x; e < total; e += kZfThreads) gst<C>(plane + e, A[e]);
}

template <typename C>
__global__ void __launch_bounds__(kZfThreads) k_zfused(ZPar<C> zp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    zfused_body<C>(zp, xcd_remap(blockIdx.x, gridDim.x), smem_raw);
}

template <typename C>
__global__ void __launch_bounds__(kZfThreads) kb_zfused(const ZPar<C>* __restrict__ cs, const int* __restrict__ off,
                                                        const int* __restrict__ idx, const int* __restrict__ pnc) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
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
        zfused_body<C>(cs[idx[lo]], w - off[lo], smem_raw);
        __syncthreads();
    }
}

// LDS of the fused kernel for a plan (spectrum precision f32 or FP64)
size_t zf_lds_bytes(int nym, int n, int ntw, int f32) {
    const size_t cs = f32 ? sizeof(float2) : sizeof(double2);
    return cs * ((size_t)nym * n + ntw) + sizeof(double) * 2 * (size_t)nym;
}

size_t zf_lds_bytes(const dfp_plan* p, int f32) {
    return zf_lds_bytes(p->geo.nym, p->fz.n, twiddle_count(p->fz), f32);
}

// FFT rows per call of the fused kernel: the in-place capacity of two
// butterflies per thread
int zf_rows_per_call(const FftDesc& fz) {
    return std::max(1, inplace_capacity(fz, kZfThreads, 2) / std::max(1, fz.n));
}

template <typename C>
hipError_t launch_zfused(dfp_plan* p, const ZPar<C>& zp) {
    const bool f32 = sizeof(C) == 8;
    const size_t lds = zf_lds_bytes(p, f32);
    const int slot = prof_begin(p, f32 ? "zfused32" : "zfused");
    hipLaunchKernelGGL(k_zfused<C>, dim3(zp.nblocks), dim3(kZfThreads), lds, p->stream, zp);
    prof_end(p, slot);
    return hipGetLastError();
}
template hipError_t launch_zfused<double2>(dfp_plan*, const ZPar<double2>&);
template hipError_t launch_zfused<float2>(dfp_plan*, const ZPar<float2>&);

template <typename C>
hipError_t launch_zfused_batch(hipStream_t s, const ZPar<C>* cs, const int* off, const int* idx, const int* nc,
                               int grid, size_t lds) {
    hipLaunchKernelGGL(kb_zfused<C>, dim3(grid), dim3(kZfThreads), lds, s, cs, off, idx, nc);
    return hipGetLastError();
}
template hipError_t launch_zfused_batch<double2>(hipStream_t, const ZPar<double2>*, const int*, const int*,
                                                 const int*, int, size_t);
template hipError_t launch_zfused_batch<float2>(hipStream_t, const ZPar<float2>*, const int*, const int*,
                                                const int*, int, size_t);

}  // namespace dfp
