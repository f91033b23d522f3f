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
// The y solves run chunk-parallel over all lanes of the workgroup (see the
// sweep below), so they round differently from the split kernels' serial
// sweeps: the two paths agree to round-off, not bit for bit.
#include "fft_lds.h"

namespace dfp {

constexpr int kZfThreads = 256;
constexpr int kZfPyThreads = 512;  // periodic-y kernel: no serial sweeps, more waves to hide the plane loads

// rows of the plane through the in-place LDS FFT, zp.G rows per call (the
// per-call capacity of two butterflies per thread)
template <int SIGN, typename C>
__device__ __forceinline__ void zf_rows_fft(C* A, const ZPar<C>& zp, const C* twl) {
    const int n = zp.fz.n;
    for (int l0 = 0; l0 < zp.nym; l0 += zp.G) lds_fft_inplace<SIGN, 2>(A + (long)l0 * n, min(zp.G, zp.nym - l0), n, zp.fz, twl);
}

// the plane (total values) into LDS by NT threads: four values per thread in
// flight per batch (named registers, clamped unconditional loads)
template <int NT, typename C>
__device__ __forceinline__ void zf_load_plane(C* A, const Glob<C>* plane, int total) {
    for (int base = 0; base < total; base += 4 * NT) {
        const int e0 = base + (int)threadIdx.x;
        const C v0 = gld<C>(plane + min(e0, total - 1));
        const C v1 = gld<C>(plane + min(e0 + NT, total - 1));
        const C v2 = gld<C>(plane + min(e0 + 2 * NT, total - 1));
        const C v3 = gld<C>(plane + min(e0 + 3 * NT, total - 1));
        if (e0 < total) A[e0] = v0;
        if (e0 + NT < total) A[e0 + NT] = v1;
        if (e0 + 2 * NT < total) A[e0 + 2 * NT] = v2;
        if (e0 + 3 * NT < total) A[e0 + 3 * NT] = v3;
    }
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
    double* sg = cu + nym;                                   // chunk summaries: kZfThreads x (re, im, gain)
    double* sc = sg + 3 * kZfThreads;                        // chunk carries: kZfThreads x (re, im)
    const auto plane = gptr(zp.spec) + (long)b * nym * n;
    const int total = nym * n;
    zf_load_plane<kZfThreads>(A, plane, total);
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

    // y solves, chunk-parallel: lane t takes mode m = t % n and the rows
    // [s, e) of chunk c = t / n (P = kZfThreads / n chunks of L rows).  The
    // elimination y(j) = b(j) - mu(j) y(j-1), mu(j) = DL(j-1)/D'(j-1), and the
    // back substitution x(j) = alpha(j) - beta(j) x(j+1), alpha = y/D',
    // beta = DU(j)/D', are first-order linear recurrences: each chunk runs its
    // own from a zero start while accumulating the gain of its start value
    // (phase 1), one lane per mode chains the chunk carries (phase 2), and each
    // chunk adds gain x carry to its rows (phase 3).  Pivots 1/D'(j) come from
    // the plan's table (pin of projection.f90:292 included).
    const int P = max(1, kZfThreads / n);
    const int L = (nym + P - 1) / P;
    const int t = threadIdx.x;
    const int m = t % n;
    const int c = t / n;
    const bool act = c < P && t < n * P;
    const int s = min(c * L, nym), e = min(s + L, nym);
    const auto lcol = gptr(zp.lu) + (long)kx * nym * nzs + lu_pos_mu(m, u, R, n);
    // phase 1, elimination: local yh(j) from yh(s-1) = 0, gain g(j) = prod(-mu)
    if (act && s < e) {
        double yx = 0.0, yy = 0.0, gn = 1.0;
        double ipp = s > 0 ? lcol[(long)(s - 1) * nzs] : 0.0;
        for (int j = s; j < e; ++j) {
            const double mu = j > 0 ? cl[j] * ipp : 0.0;
            ipp = lcol[(long)j * nzs];
            const C a = A[j * n + m];
            yx = (double)a.x - mu * yx;
            yy = (double)a.y - mu * yy;
            gn = -mu * gn;
            A[j * n + m] = CplxT<C>::make((Rl)yx, (Rl)yy);
        }
        sg[3 * t] = yx;
        sg[3 * t + 1] = yy;
        sg[3 * t + 2] = gn;
    }
    __syncthreads();
    // phase 2: y(s_c - 1) for every chunk, chained from chunk 0 (y(-1) = 0)
    if (t < n) {
        double cx = 0.0, cy = 0.0;
        for (int q = 0; q < P; ++q) {
            const int tq = q * n + t;
            sc[2 * tq] = cx;
            sc[2 * tq + 1] = cy;
            if (min(q * L, nym) < nym) {
                const double gq = sg[3 * tq + 2];
                const double nx = sg[3 * tq] + gq * cx, ny = sg[3 * tq + 1] + gq * cy;
                cx = nx;
                cy = ny;
            }
        }
    }
    __syncthreads();
    // phase 3: y(j) = yh(j) + g(j) y(s-1); then the local back substitution
    // xh(j) from xh(e) = 0 with gain h(j) = prod(-beta)
    if (act && s < e) {
        const double cx = sc[2 * t], cy = sc[2 * t + 1];
        double gn = 1.0;
        double ipp = s > 0 ? lcol[(long)(s - 1) * nzs] : 0.0;
        for (int j = s; j < e; ++j) {
            const double mu = j > 0 ? cl[j] * ipp : 0.0;
            ipp = lcol[(long)j * nzs];
            gn = -mu * gn;
            const C a = A[j * n + m];
            A[j * n + m] = CplxT<C>::make((Rl)((double)a.x + gn * cx), (Rl)((double)a.y + gn * cy));
        }
        double xx = 0.0, xy = 0.0, hn = 1.0;
        for (int j = e - 1; j >= s; --j) {
            const double ip = lcol[(long)j * nzs];
            const double beta = cu[j] * ip;  // cu[nym-1] = 0
            const C a = A[j * n + m];
            xx = (double)a.x * ip - beta * xx;
            xy = (double)a.y * ip - beta * xy;
            hn = -beta * hn;
            A[j * n + m] = CplxT<C>::make((Rl)xx, (Rl)xy);
        }
        sg[3 * t] = xx;
        sg[3 * t + 1] = xy;
        sg[3 * t + 2] = hn;
    }
    __syncthreads();
    // phase 4: x(e_c) for every chunk, chained from the last (x(nym) = 0)
    if (t < n) {
        double cx = 0.0, cy = 0.0;
        for (int q = P - 1; q >= 0; --q) {
            const int tq = q * n + t;
            sc[2 * tq] = cx;
            sc[2 * tq + 1] = cy;
            if (min(q * L, nym) < nym) {
                const double hq = sg[3 * tq + 2];
                const double nx = sg[3 * tq] + hq * cx, ny = sg[3 * tq + 1] + hq * cy;
                cx = nx;
                cy = ny;
            }
        }
    }
    __syncthreads();
    // phase 5: x(j) = xh(j) + h(j) x(e)
    if (act && s < e) {
        const double cx = sc[2 * t], cy = sc[2 * t + 1];
        const bool is00 = zp.zero00 && kx == 0 && u == 0 && m == 0;
        double hn = 1.0;
        for (int j = e - 1; j >= s; --j) {
            const double beta = cu[j] * lcol[(long)j * nzs];
            hn = -beta * hn;
            const C a = A[j * n + m];
            double xx = (double)a.x + hn * cx, xy = (double)a.y + hn * cy;
            if (is00) xx = xy = 0.0;
            A[j * n + m] = CplxT<C>::make((Rl)xx, (Rl)xy);
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

// Periodic y (y_bc_type = 0 on a uniform y grid, projection.f90:246-273):
// the wall solve's place is taken by a y FFT over the nyp = nym - 1
// independent rows of the sub-plane, a division of each (kx, ky, kz) mode by
// kxx + kyy + kzz with the (0,0,0) mode set to zero, and the inverse y FFT
// scaled by 1/nyp (folded into the division).  Row nym - 1, the duplicate
// the reference leaves out of its transform, then takes row 0
// (projection.f90:350-354): the inverse z and x transforms reproduce row 0 in
// it bit for bit.  (The reference's copy also spans rhs_p's ghost positions
// i = nxg and k = nzg; it zeroes rhs_p at allocation, initialization.f90:220,
// and only ever copies those positions among themselves, so they stay zero,
// and the library leaves them as the caller's array holds them.)  One
// workgroup per (kx, u) sub-plane, as above.
__device__ __forceinline__ void zf_py_body(const ZyPar<double2>& q, int b, unsigned char* smem_raw) {
    using C = double2;
    const ZPar<C>& zp = q.z;
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int n = zp.fz.n;
    const int nym = zp.nym, nyp = nym - 1, R = zp.R, ntw = zp.ntw;
    const int kx = b / R;
    const int u = b - kx * R;
    C* A = smem;                       // nym * n
    C* twl = A + (long)nym * n;        // ntw: z stages
    C* twy = twl + ntw;                // q.ntwy: y stages
    const auto plane = gptr(zp.spec) + (long)b * nym * n;
    const int total = nym * n;
    zf_load_plane<kZfPyThreads>(A, plane, total);
    const auto tw = gptr(zp.tw);
    const auto tyg = gptr(q.twy);
    for (int e = threadIdx.x; e < ntw; e += kZfPyThreads) twl[e] = gld<C>(tw + e);
    for (int e = threadIdx.x; e < q.ntwy; e += kZfPyThreads) twy[e] = gld<C>(tyg + e);
    __syncthreads();
    zf_rows_fft<-1>(A, zp, twl);
    for (int c0 = 0; c0 < n; c0 += q.gy) lds_fft_cols<-1, 4>(A, c0, q.gy, n, q.fy, twy, q.div_gy);
    // element (ky, m) of the sub-plane is mode (kx, ky, kz = R m + u)
    const double lx = q.kxx[kx];
    for (int e = threadIdx.x; e < nyp * n; e += kZfPyThreads) {
        const int ky = (int)zp.div_n.div((uint32_t)e);
        const int m = e - ky * n;
        const int kz = R * m + u;
        const double lam = lx + q.kyy[ky] + q.kzz[kz];
        const C a = A[e];
        if (kx == 0 && ky == 0 && kz == 0) {
            A[e] = make_double2(0.0, 0.0);
        } else {
            const double s = q.inv_ny / lam;
            A[e] = make_double2(a.x * s, a.y * s);
        }
    }
    __syncthreads();
    for (int c0 = 0; c0 < n; c0 += q.gy) lds_fft_cols<1, 4>(A, c0, q.gy, n, q.fy, twy, q.div_gy);
    for (int m = threadIdx.x; m < n; m += kZfPyThreads) A[(long)nyp * n + m] = A[m];
    __syncthreads();
    zf_rows_fft<1>(A, zp, twl);
    for (int e = threadIdx.x; e < total; e += kZfPyThreads) gst<C>(plane + e, A[e]);
}

__global__ void __launch_bounds__(kZfPyThreads) k_zfused_py(ZyPar<double2> q) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    zf_py_body(q, xcd_remap(blockIdx.x, gridDim.x), smem_raw);
}

size_t zf_py_lds_bytes(int nym, int n, int ntwz, int ntwy) {
    return sizeof(double2) * ((size_t)nym * n + ntwz + ntwy);
}

int zf_py_columns(const FftDesc& fy, int n) {
    const int cap = inplace_capacity(fy, kZfPyThreads, 4);  // butterfly slots x smallest radix
    for (int g = n; g >= 1; --g)
        if (n % g == 0 && (long)g * fy.n <= cap) return g;
    return 0;
}

hipError_t launch_zfused_py(dfp_plan* p) {
    ZyPar<double2> q{};
    q.z = make_zpar<double2>(p, 0);
    q.fy = p->fy;
    q.twy = p->tw_y;
    q.kxx = p->ops.kxx;
    q.kyy = p->ops.kyy;
    q.kzz = p->ops.kzz;
    q.ntwy = twiddle_count(p->fy);
    q.gy = p->gy;
    q.div_gy = FastDiv((uint32_t)p->gy);
    q.inv_ny = 1.0 / (double)(p->geo.nym - 1);
    const size_t lds = zf_py_lds_bytes(p->geo.nym, p->fz.n, q.z.ntw, q.ntwy);
    const int slot = prof_begin(p, "zfused_py");
    hipLaunchKernelGGL(k_zfused_py, dim3(q.z.nblocks), dim3(kZfPyThreads), lds, p->stream, q);
    prof_end(p, slot);
    return hipGetLastError();
}

// LDS of the fused kernel for a plan (spectrum precision f32 or FP64)
size_t zf_lds_bytes(int nym, int n, int ntw, int f32) {
    const size_t cs = f32 ? sizeof(float2) : sizeof(double2);
    return cs * ((size_t)nym * n + ntw) + sizeof(double) * (2 * (size_t)nym + 5 * kZfThreads);
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
    const auto k = kb_zfused<C>;
    hipLaunchKernelGGL(k, dim3(resident_grid(k, grid, kZfThreads, lds)), dim3(kZfThreads), lds, s, cs, off, idx, nc);
    return hipGetLastError();
}
template hipError_t launch_zfused_batch<double2>(hipStream_t, const ZPar<double2>*, const int*, const int*,
                                                 const int*, int, size_t);
template hipError_t launch_zfused_batch<float2>(hipStream_t, const ZPar<float2>*, const int*, const int*,
                                                const int*, int, size_t);

}  // namespace dfp
