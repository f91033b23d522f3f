// Batched complex FFTs staged in LDS (FP64 or FP32), Stockham decimation in
// frequency with natural-order output, radix 2, 3 and 4, in place.
//
// Stage st of a length-n transform with radix r, current length ncur and
// stride s (= product of the radices already applied), m = ncur / r:
//   a_t = src[q + s*(p + t*m)],                t = 0..r-1
//   dst[q + s*(r*p + u)] = (sum_t a_t w_r^{t*u}) * w_ncur^{p*u}
// for every butterfly (p, q), p = 0..m-1, q = 0..s-1, where
// w_L = exp(SIGN * 2 pi i / L).  SIGN = -1 is the forward transform
// (e^{-i}, FFTW_FORWARD), SIGN = +1 the unnormalised backward one.
#pragma once

#include "dfp_internal.h"

namespace dfp {

// complex arithmetic on double2 (FP64 path) and float2 (FP32 spectral path)
template <typename C> struct CplxT;
template <> struct CplxT<double2> {
    using real = double;
    static __device__ __forceinline__ double2 make(double x, double y) { return make_double2(x, y); }
};
template <> struct CplxT<float2> {
    using real = float;
    static __device__ __forceinline__ float2 make(float x, float y) { return make_float2(x, y); }
};

// complex values through global-address-space pointers (see gptr): via the
// builtin vector type, since the HIP vector classes do not bind to
// address-space-qualified references
template <typename C> struct NativeT;
template <> struct NativeT<double2> { using type = double __attribute__((ext_vector_type(2))); };
template <> struct NativeT<float2> { using type = float __attribute__((ext_vector_type(2))); };
template <typename C>
__device__ __forceinline__ C gld(const Glob<C>* p) {
    using N = typename NativeT<C>::type;
    const N v = *(const Glob<N>*)p;
    return CplxT<C>::make(v.x, v.y);
}
template <typename C>
__device__ __forceinline__ void gst(Glob<C>* p, C v) {
    using N = typename NativeT<C>::type;
    N w;
    w.x = v.x;
    w.y = v.y;
    *(Glob<N>*)p = w;
}

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return CplxT<C>::make(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return CplxT<C>::make(a.x - b.x, a.y - b.y); }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) { return CplxT<C>::make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
template <typename C>
__device__ __forceinline__ C cmulc(C a, C b) { return CplxT<C>::make(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }

template <int SIGN, typename C>
__device__ __forceinline__ C twiddle(C a, C w) {
    return SIGN < 0 ? cmul(a, w) : cmulc(a, w);
}

// SIGN * i * a
template <int SIGN, typename C>
__device__ __forceinline__ C mul_si(C a) {
    return SIGN < 0 ? CplxT<C>::make(a.y, -a.x) : CplxT<C>::make(-a.y, a.x);
}

// ---------------------------------------------------------------------------
// In-register DFTs of R points, R in {1,2,3,4,6,8,9,12,16}, in place with
// natural-order input and output: X[u] = sum_t a[t] exp(SIGN 2 pi i t u / R).
// Composite R = R1*R2 by one Cooley-Tukey step (t = t1*R2 + t2,
// u = u1 + R1*u2): R2 DFTs of length R1, twiddles w_R^(t2*u1), R1 DFTs of
// length R2.  The twiddle constants are cos/sin(2 pi k / R) rounded once
// from 50-digit values (exactly symmetric).
// ---------------------------------------------------------------------------
constexpr double kCos6[6] = {1.000000000000000000000e+00, 5.000000000000000000000e-01, -5.000000000000000000000e-01, -1.000000000000000000000e+00, -5.000000000000000000000e-01, 5.000000000000000000000e-01};
constexpr double kSin6[6] = {0.0, 8.660254037844385965883e-01, 8.660254037844385965883e-01, 0.0, -8.660254037844385965883e-01, -8.660254037844385965883e-01};
constexpr double kCos8[8] = {1.000000000000000000000e+00, 7.071067811865475727373e-01, 0.0, -7.071067811865475727373e-01, -1.000000000000000000000e+00, -7.071067811865475727373e-01, 0.0, 7.071067811865475727373e-01};
constexpr double kSin8[8] = {0.0, 7.071067811865475727373e-01, 1.000000000000000000000e+00, 7.071067811865475727373e-01, 0.0, -7.071067811865475727373e-01, -1.000000000000000000000e+00, -7.071067811865475727373e-01};
constexpr double kCos9[9] = {1.000000000000000000000e+00, 7.660444431189780134517e-01, 1.736481776669303589422e-01, -5.000000000000000000000e-01, -9.396926207859084279050e-01, -9.396926207859084279050e-01, -5.000000000000000000000e-01, 1.736481776669303589422e-01, 7.660444431189780134517e-01};
constexpr double kSin9[9] = {0.0, 6.427876096865393629187e-01, 9.848077530122080203157e-01, 8.660254037844385965883e-01, 3.420201433256687129081e-01, -3.420201433256687129081e-01, -8.660254037844385965883e-01, -9.848077530122080203157e-01, -6.427876096865393629187e-01};
constexpr double kCos12[12] = {1.000000000000000000000e+00, 8.660254037844385965883e-01, 5.000000000000000000000e-01, 0.0, -5.000000000000000000000e-01, -8.660254037844385965883e-01, -1.000000000000000000000e+00, -8.660254037844385965883e-01, -5.000000000000000000000e-01, 0.0, 5.000000000000000000000e-01, 8.660254037844385965883e-01};
constexpr double kSin12[12] = {0.0, 5.000000000000000000000e-01, 8.660254037844385965883e-01, 1.000000000000000000000e+00, 8.660254037844385965883e-01, 5.000000000000000000000e-01, 0.0, -5.000000000000000000000e-01, -8.660254037844385965883e-01, -1.000000000000000000000e+00, -8.660254037844385965883e-01, -5.000000000000000000000e-01};
constexpr double kCos16[16] = {1.000000000000000000000e+00, 9.238795325112867384831e-01, 7.071067811865475727373e-01, 3.826834323650897817792e-01, 0.0, -3.826834323650897817792e-01, -7.071067811865475727373e-01, -9.238795325112867384831e-01, -1.000000000000000000000e+00, -9.238795325112867384831e-01, -7.071067811865475727373e-01, -3.826834323650897817792e-01, 0.0, 3.826834323650897817792e-01, 7.071067811865475727373e-01, 9.238795325112867384831e-01};
constexpr double kSin16[16] = {0.0, 3.826834323650897817792e-01, 7.071067811865475727373e-01, 9.238795325112867384831e-01, 1.000000000000000000000e+00, 9.238795325112867384831e-01, 7.071067811865475727373e-01, 3.826834323650897817792e-01, 0.0, -3.826834323650897817792e-01, -7.071067811865475727373e-01, -9.238795325112867384831e-01, -1.000000000000000000000e+00, -9.238795325112867384831e-01, -7.071067811865475727373e-01, -3.826834323650897817792e-01};

template <int R> struct SmallTw;
template <> struct SmallTw<6> { static __device__ __forceinline__ double c(int k) { return kCos6[k]; } static __device__ __forceinline__ double s(int k) { return kSin6[k]; } };
template <> struct SmallTw<8> { static __device__ __forceinline__ double c(int k) { return kCos8[k]; } static __device__ __forceinline__ double s(int k) { return kSin8[k]; } };
template <> struct SmallTw<9> { static __device__ __forceinline__ double c(int k) { return kCos9[k]; } static __device__ __forceinline__ double s(int k) { return kSin9[k]; } };
template <> struct SmallTw<12> { static __device__ __forceinline__ double c(int k) { return kCos12[k]; } static __device__ __forceinline__ double s(int k) { return kSin12[k]; } };
template <> struct SmallTw<16> { static __device__ __forceinline__ double c(int k) { return kCos16[k]; } static __device__ __forceinline__ double s(int k) { return kSin16[k]; } };

template <int SIGN, int R, typename C>
__device__ __forceinline__ void dft_reg(C* a);

template <int SIGN, int R1, int R2, typename C>
__device__ __forceinline__ void dft_ct(C* a) {
    using Rl = typename CplxT<C>::real;
    constexpr int R = R1 * R2;
    C b[R];
#pragma unroll
    for (int t2 = 0; t2 < R2; ++t2) {
        C v[R1];
#pragma unroll
        for (int t1 = 0; t1 < R1; ++t1) v[t1] = a[t1 * R2 + t2];
        dft_reg<SIGN, R1>(v);
#pragma unroll
        for (int u1 = 0; u1 < R1; ++u1) {
            const int k = (t2 * u1) % R;
            if (k == 0) {
                b[u1 * R2 + t2] = v[u1];
            } else {
                const C w = CplxT<C>::make((Rl)SmallTw<R>::c(k), (Rl)(SIGN * SmallTw<R>::s(k)));
                b[u1 * R2 + t2] = cmul(v[u1], w);
            }
        }
    }
#pragma unroll
    for (int u1 = 0; u1 < R1; ++u1) {
        C v[R2];
#pragma unroll
        for (int t2 = 0; t2 < R2; ++t2) v[t2] = b[u1 * R2 + t2];
        dft_reg<SIGN, R2>(v);
#pragma unroll
        for (int u2 = 0; u2 < R2; ++u2) a[u1 + R1 * u2] = v[u2];
    }
}

template <int SIGN, int R, typename C>
__device__ __forceinline__ void dft_reg(C* a) {
    using Rl = typename CplxT<C>::real;
    if constexpr (R == 2) {
        const C s = cadd(a[0], a[1]), d = csub(a[0], a[1]);
        a[0] = s;
        a[1] = d;
    } else if constexpr (R == 3) {
        const Rl s3 = (Rl)0.86602540378443864676372317075294;
        const C t1 = cadd(a[1], a[2]);
        const C t2 = CplxT<C>::make(a[0].x - (Rl)0.5 * t1.x, a[0].y - (Rl)0.5 * t1.y);
        const C dd = csub(a[1], a[2]);
        const C jd = mul_si<SIGN>(CplxT<C>::make(s3 * dd.x, s3 * dd.y));
        a[0] = cadd(a[0], t1);
        a[1] = cadd(t2, jd);
        a[2] = csub(t2, jd);
    } else if constexpr (R == 4) {
        const C s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const C s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
        const C jd13 = mul_si<SIGN>(d13);
        a[0] = cadd(s02, s13);
        a[1] = cadd(d02, jd13);
        a[2] = csub(s02, s13);
        a[3] = csub(d02, jd13);
    } else if constexpr (R == 6) {
        dft_ct<SIGN, 2, 3>(a);
    } else if constexpr (R == 8) {
        dft_ct<SIGN, 4, 2>(a);
    } else if constexpr (R == 9) {
        dft_ct<SIGN, 3, 3>(a);
    } else if constexpr (R == 12) {
        dft_ct<SIGN, 4, 3>(a);
    } else if constexpr (R == 16) {
        dft_ct<SIGN, 4, 4>(a);
    }
}

// runtime R (a plan constant, uniform) over a register array of RMAX entries
template <int SIGN, int RMAX, typename C>
__device__ __forceinline__ void dft_rt(C (&a)[RMAX], int R) {
    switch (R) {
        case 2: dft_reg<SIGN, 2>(a); break;
        case 3: dft_reg<SIGN, 3>(a); break;
        case 4: dft_reg<SIGN, 4>(a); break;
        default:
            if constexpr (RMAX >= 16) {
                switch (R) {
                    case 6: dft_reg<SIGN, 6>(a); break;
                    case 8: dft_reg<SIGN, 8>(a); break;
                    case 9: dft_reg<SIGN, 9>(a); break;
                    case 12: dft_reg<SIGN, 12>(a); break;
                    case 16: dft_reg<SIGN, 16>(a); break;
                    default: break;
                }
            }
            break;
    }
}

// Sum over the block of one value per thread in a fixed order (butterflies
// within each 64-lane wave, then the waves in order); every thread of the
// block must call it and receives the total.  wsum holds NT/64 doubles.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* wsum) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < NT / 64; ++q) s += wsum[q];
    __syncthreads();
    return s;
}

// Transforms nl lines in LDS (line l, element e at buf[l*ld + e]) in place:
// each thread holds the outputs of up to KMAX butterflies in registers
// between a read phase and a write phase, so one LDS buffer suffices.
// Requires nl * n / r <= KMAX * blockDim.x for every stage radix r; with
// PAIR2 a radix-2 stage with more butterflies than that puts two in each
// register slot (b and b + KMAX * blockDim.x), which doubles what its stages
// hold: nl * n / 2 <= 2 * KMAX * blockDim.x.
// Ends with a barrier.
template <int SIGN, int KMAX, typename C, bool PAIR2 = false>
__device__ void lds_fft_inplace(C* buf, int nl, int ld, const FftDesc& d, const C* tw) {
    using R = typename CplxT<C>::real;
    int ncur = d.n, s = 1;
    for (int st = 0; st < d.nst; ++st) {
        const int r = d.rad[st];
        const int m = ncur / r;
        const int nb = m * s;
        const int tot = nb * nl;
        const C* tws = tw + d.twoff[st];
        if (PAIR2 && r == 2 && tot > KMAX * (int)blockDim.x) {
            C v[KMAX][4];
            int dst[KMAX][2];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int b = threadIdx.x + (k + h * KMAX) * blockDim.x;
                    dst[k][h] = -1;
                    if (b < tot) {
                        const int l = (int)d.nbdiv[st].div((uint32_t)b);
                        const int bi = b - l * nb;
                        const int p = (int)d.sdiv[st].div((uint32_t)bi);
                        const int q = bi - p * s;
                        const C* sl = buf + l * ld + q + s * p;
                        dst[k][h] = l * ld + q + s * 2 * p;
                        const C a0 = sl[0], a1 = sl[s * m];
                        v[k][2 * h] = cadd(a0, a1);
                        v[k][2 * h + 1] = twiddle<SIGN>(csub(a0, a1), tws[p]);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (dst[k][h] >= 0) {
                        C* dl = buf + dst[k][h];
                        dl[0] = v[k][2 * h];
                        dl[s] = v[k][2 * h + 1];
                    }
                }
            }
            __syncthreads();
            s *= r;
            ncur = m;
            continue;
        }
        C v[KMAX][4];
        int dsto[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const int b = threadIdx.x + k * blockDim.x;
            dsto[k] = -1;
            if (b < tot) {
                const int l = (int)d.nbdiv[st].div((uint32_t)b);
                const int bi = b - l * nb;
                const int p = (int)d.sdiv[st].div((uint32_t)bi);
                const int q = bi - p * s;
                const C* sl = buf + l * ld + q + s * p;
                dsto[k] = l * ld + q + s * r * p;
                if (r == 4) {
                    const C a0 = sl[0], a1 = sl[s * m], a2 = sl[2 * s * m], a3 = sl[3 * s * m];
                    const C s02 = cadd(a0, a2), d02 = csub(a0, a2);
                    const C s13 = cadd(a1, a3), d13 = csub(a1, a3);
                    const C jd13 = mul_si<SIGN>(d13);
                    const C* w = tws + p * 3;
                    v[k][0] = cadd(s02, s13);
                    v[k][1] = twiddle<SIGN>(cadd(d02, jd13), w[0]);
                    v[k][2] = twiddle<SIGN>(csub(s02, s13), w[1]);
                    v[k][3] = twiddle<SIGN>(csub(d02, jd13), w[2]);
                } else if (r == 2) {
                    const C a0 = sl[0], a1 = sl[s * m];
                    v[k][0] = cadd(a0, a1);
                    v[k][1] = twiddle<SIGN>(csub(a0, a1), tws[p]);
                } else {
                    const R s3 = (R)0.86602540378443864676372317075294;
                    const C a0 = sl[0], a1 = sl[s * m], a2 = sl[2 * s * m];
                    const C t1 = cadd(a1, a2);
                    const C t2 = CplxT<C>::make(a0.x - (R)0.5 * t1.x, a0.y - (R)0.5 * t1.y);
                    const C dd = csub(a1, a2);
                    const C jd = mul_si<SIGN>(CplxT<C>::make(s3 * dd.x, s3 * dd.y));
                    const C* w = tws + p * 2;
                    v[k][0] = cadd(a0, t1);
                    v[k][1] = twiddle<SIGN>(cadd(t2, jd), w[0]);
                    v[k][2] = twiddle<SIGN>(csub(t2, jd), w[1]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if (dsto[k] >= 0) {
                C* dl = buf + dsto[k];
                dl[0] = v[k][0];
                dl[s] = v[k][1];
                if (r >= 3) dl[2 * s] = v[k][2];
                if (r == 4) dl[3 * s] = v[k][3];
            }
        }
        __syncthreads();
        s *= r;
        ncur = m;
    }
}

// One Stockham butterfly of radix r (2, 3 or 4): a_t = sl[t*step], outputs
// v[u] = (sum_t a_t w_r^{t*u}) * tws[p*(r-1) + u-1] (u >= 1), returned by
// value so that the callers' register arrays stay in registers
template <typename C>
struct Bfly4 {
    C v[4];
};
template <int SIGN, typename C>
__device__ __forceinline__ Bfly4<C> stockham_bfly(const C* sl, int step, int r, int p, const C* tws) {
    using R = typename CplxT<C>::real;
    Bfly4<C> o;
    C* v = o.v;
    if (r == 4) {
        const C a0 = sl[0], a1 = sl[step], a2 = sl[2 * step], a3 = sl[3 * step];
        const C s02 = cadd(a0, a2), d02 = csub(a0, a2);
        const C s13 = cadd(a1, a3), d13 = csub(a1, a3);
        const C jd13 = mul_si<SIGN>(d13);
        const C* w = tws + p * 3;
        v[0] = cadd(s02, s13);
        v[1] = twiddle<SIGN>(cadd(d02, jd13), w[0]);
        v[2] = twiddle<SIGN>(csub(s02, s13), w[1]);
        v[3] = twiddle<SIGN>(csub(d02, jd13), w[2]);
    } else if (r == 2) {
        const C a0 = sl[0], a1 = sl[step];
        v[0] = cadd(a0, a1);
        v[1] = twiddle<SIGN>(csub(a0, a1), tws[p]);
    } else {
        const R s3 = (R)0.86602540378443864676372317075294;
        const C a0 = sl[0], a1 = sl[step], a2 = sl[2 * step];
        const C t1 = cadd(a1, a2);
        const C t2 = CplxT<C>::make(a0.x - (R)0.5 * t1.x, a0.y - (R)0.5 * t1.y);
        const C dd = csub(a1, a2);
        const C jd = mul_si<SIGN>(CplxT<C>::make(s3 * dd.x, s3 * dd.y));
        const C* w = tws + p * 2;
        v[0] = cadd(a0, t1);
        v[1] = twiddle<SIGN>(cadd(t2, jd), w[0]);
        v[2] = twiddle<SIGN>(csub(t2, jd), w[1]);
    }
    return o;
}

// Transforms the columns c0 .. c0+nc-1 of a row-major LDS array (element e of
// column c at buf[e*ld + c]) in place, the same stages as lds_fft_inplace.
// Butterfly b takes column c0 + b % nc, so neighbouring lanes read
// neighbouring columns (no bank conflicts whatever ld is).  divc divides by
// nc.  Requires nc * n / r <= KMAX * blockDim.x for every stage radix r.
// Ends with a barrier.
template <int SIGN, int KMAX, typename C>
__device__ void lds_fft_cols(C* buf, int c0, int nc, int ld, const FftDesc& d, const C* tw, const FastDiv& divc) {
    int ncur = d.n, s = 1;
    for (int st = 0; st < d.nst; ++st) {
        const int r = d.rad[st];
        const int m = ncur / r;
        const int tot = m * s * nc;
        const C* tws = tw + d.twoff[st];
        C v[KMAX][4];
        int dsto[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const int b = threadIdx.x + k * blockDim.x;
            dsto[k] = -1;
            if (b < tot) {
                const int bi = (int)divc.div((uint32_t)b);
                const int c = c0 + b - bi * nc;
                const int p = (int)d.sdiv[st].div((uint32_t)bi);
                const int q = bi - p * s;
                dsto[k] = (q + s * r * p) * ld + c;
                const Bfly4<C> o = stockham_bfly<SIGN>(buf + (q + s * p) * ld + c, s * m * ld, r, p, tws);
                v[k][0] = o.v[0];
                v[k][1] = o.v[1];
                v[k][2] = o.v[2];
                v[k][3] = o.v[3];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if (dsto[k] >= 0) {
                C* dl = buf + dsto[k];
                const int step = s * ld;
                dl[0] = v[k][0];
                dl[step] = v[k][1];
                if (r >= 3) dl[2 * step] = v[k][2];
                if (r == 4) dl[3 * step] = v[k][3];
            }
        }
        __syncthreads();
        s *= r;
        ncur = m;
    }
}

// number of complex points one block can transform in place with KMAX
// butterflies per thread and the smallest radix of the plan (a radix-2
// stage counts as 4 with PAIR2)
inline int inplace_capacity(const FftDesc& d, int threads, int kmax, bool pair2 = false) {
    int rmin = 4;
    for (int st = 0; st < d.nst; ++st) {
        const int r = d.rad[st] == 2 && pair2 ? 4 : d.rad[st];
        rmin = r < rmin ? r : rmin;
    }
    return kmax * threads * rmin;
}

// total twiddle count of a plan (the table length of its stages)
inline int twiddle_count(const FftDesc& d) {
    int ncur = d.n, t = 0;
    for (int st = 0; st < d.nst; ++st) {
        t += (ncur / d.rad[st]) * (d.rad[st] - 1);
        ncur /= d.rad[st];
    }
    return t;
}

}  // namespace dfp
