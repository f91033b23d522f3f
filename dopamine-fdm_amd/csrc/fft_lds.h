// Batched complex FFTs staged in LDS (FP64), Stockham decimation in
// frequency with natural-order output, radix 2, 3 and 4.
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

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

template <int SIGN>
__device__ __forceinline__ double2 twiddle(double2 a, double2 w) {
    return SIGN < 0 ? cmul(a, w) : cmulc(a, w);
}

// SIGN * i * a
template <int SIGN>
__device__ __forceinline__ double2 mul_si(double2 a) {
    return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// Transforms nl lines in LDS; line l, element e at buf[l*ld + e].  bufA holds
// the input, bufB is scratch of the same size; the result ends in one of the
// two and is returned.  Every thread of the block must call this; it ends
// with a barrier.
template <int SIGN>
__device__ double2* lds_fft_lines(double2* bufA, double2* bufB, int nl, int ld, const FftDesc& d,
                                  const double2* __restrict__ tw) {
    double2* src = bufA;
    double2* dst = bufB;
    int ncur = d.n, s = 1;
    for (int st = 0; st < d.nst; ++st) {
        const int r = d.rad[st];
        const int m = ncur / r;
        const int nb = m * s;  // butterflies per line
        const double2* __restrict__ tws = tw + d.twoff[st];
        for (int b = threadIdx.x; b < nb * nl; b += blockDim.x) {
            const int l = b / nb;
            const int bi = b - l * nb;
            const int p = bi / s;
            const int q = bi - p * s;
            const double2* sl = src + l * ld + q + s * p;
            double2* dl = dst + l * ld + q + s * r * p;
            if (r == 4) {
                const double2 a0 = sl[0], a1 = sl[s * m], a2 = sl[2 * s * m], a3 = sl[3 * s * m];
                const double2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
                const double2 s13 = cadd(a1, a3), d13 = csub(a1, a3);
                const double2 jd13 = mul_si<SIGN>(d13);
                dl[0] = cadd(s02, s13);
                const double2* w = tws + p * 3;
                dl[s] = twiddle<SIGN>(cadd(d02, jd13), w[0]);
                dl[2 * s] = twiddle<SIGN>(csub(s02, s13), w[1]);
                dl[3 * s] = twiddle<SIGN>(csub(d02, jd13), w[2]);
            } else if (r == 2) {
                const double2 a0 = sl[0], a1 = sl[s * m];
                dl[0] = cadd(a0, a1);
                dl[s] = twiddle<SIGN>(csub(a0, a1), tws[p]);
            } else {  // r == 3
                const double s3 = 0.86602540378443864676372317075294;
                const double2 a0 = sl[0], a1 = sl[s * m], a2 = sl[2 * s * m];
                const double2 t1 = cadd(a1, a2);
                const double2 t2 = make_double2(a0.x - 0.5 * t1.x, a0.y - 0.5 * t1.y);
                const double2 dd = csub(a1, a2);
                const double2 jd = mul_si<SIGN>(make_double2(s3 * dd.x, s3 * dd.y));
                const double2* w = tws + p * 2;
                dl[0] = cadd(a0, t1);
                dl[s] = twiddle<SIGN>(cadd(t2, jd), w[0]);
                dl[2 * s] = twiddle<SIGN>(csub(t2, jd), w[1]);
            }
        }
        __syncthreads();
        double2* t = src;
        src = dst;
        dst = t;
        s *= r;
        ncur = m;
    }
    return src;
}

}  // namespace dfp

namespace dfp {

// In-place variant: each thread holds the outputs of up to KMAX butterflies
// in registers between a read phase and a write phase, so one LDS buffer
// suffices.  Requires nl * n / r <= KMAX * blockDim.x for every stage radix r.
// Ends with a barrier.
template <int SIGN, int KMAX>
__device__ void lds_fft_inplace(double2* buf, int nl, int ld, const FftDesc& d, const double2* tw) {
    int ncur = d.n, s = 1;
    for (int st = 0; st < d.nst; ++st) {
        const int r = d.rad[st];
        const int m = ncur / r;
        const int nb = m * s;
        const int tot = nb * nl;
        const double2* tws = tw + d.twoff[st];
        double2 v[KMAX][4];
        int dsto[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            const int b = threadIdx.x + k * blockDim.x;
            dsto[k] = -1;
            if (b < tot) {
                const int l = b / nb;
                const int bi = b - l * nb;
                const int p = bi / s;
                const int q = bi - p * s;
                const double2* sl = buf + l * ld + q + s * p;
                dsto[k] = l * ld + q + s * r * p;
                if (r == 4) {
                    const double2 a0 = sl[0], a1 = sl[s * m], a2 = sl[2 * s * m], a3 = sl[3 * s * m];
                    const double2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
                    const double2 s13 = cadd(a1, a3), d13 = csub(a1, a3);
                    const double2 jd13 = mul_si<SIGN>(d13);
                    const double2* w = tws + p * 3;
                    v[k][0] = cadd(s02, s13);
                    v[k][1] = twiddle<SIGN>(cadd(d02, jd13), w[0]);
                    v[k][2] = twiddle<SIGN>(csub(s02, s13), w[1]);
                    v[k][3] = twiddle<SIGN>(csub(d02, jd13), w[2]);
                } else if (r == 2) {
                    const double2 a0 = sl[0], a1 = sl[s * m];
                    v[k][0] = cadd(a0, a1);
                    v[k][1] = twiddle<SIGN>(csub(a0, a1), tws[p]);
                } else {
                    const double s3 = 0.86602540378443864676372317075294;
                    const double2 a0 = sl[0], a1 = sl[s * m], a2 = sl[2 * s * m];
                    const double2 t1 = cadd(a1, a2);
                    const double2 t2 = make_double2(a0.x - 0.5 * t1.x, a0.y - 0.5 * t1.y);
                    const double2 dd = csub(a1, a2);
                    const double2 jd = mul_si<SIGN>(make_double2(s3 * dd.x, s3 * dd.y));
                    const double2* w = tws + p * 2;
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
                double2* dl = buf + dsto[k];
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
