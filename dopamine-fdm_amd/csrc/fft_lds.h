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
// between a read phase and a write phase, so one LDS buffer suffices.  Requires nl * n / r <= KMAX * blockDim.x for every stage radix r.
// Ends with a barrier.
template <int SIGN, int KMAX, typename C>
__device__ void lds_fft_inplace(C* buf, int nl, int ld, const FftDesc& d, const C* tw) {
    using R = typename CplxT<C>::real;
    int ncur = d.n, s = 1;
    for (int st = 0; st < d.nst; ++st) {
        const int r = d.rad[st];
        const int m = ncur / r;
        const int nb = m * s;
        const int tot = nb * nl;
        const C* tws = tw + d.twoff[st];
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

// number of complex points one block can transform in place with KMAX
// butterflies per thread and the smallest radix of the plan
inline int inplace_capacity(const FftDesc& d, int threads, int kmax) {
    int rmin = 4;
    for (int st = 0; st < d.nst; ++st) rmin = d.rad[st] < rmin ? d.rad[st] : rmin;
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
