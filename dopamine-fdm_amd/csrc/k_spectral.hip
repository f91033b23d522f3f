// Host side of the x-direction spectral kernels (k_xfft_impl.h): parameter
// records, LDS sizes, and launches dispatched on the plan's z-split radix.
#include <cmath>

#include "fft_lds.h"

namespace dfp {


template <int R, typename C>
hipError_t launch_xfwd_r(hipStream_t s, int mode, const XPar<C>& q, dim3 grid, size_t lds);
template <int R, typename C>
hipError_t launch_xinv_r(hipStream_t s, int accum, const XPar<C>& q, dim3 grid, size_t lds);
template <int R, typename C>
hipError_t launch_xfwd_batch_r(hipStream_t s, int mode, const XPar<C>* cs, const int* off, const int* idx,
                               const int* nc, int grid, size_t lds);
template <int R, typename C>
hipError_t launch_xinv_batch_r(hipStream_t s, int accum, const XPar<C>* cs, const int* off, const int* idx,
                               const int* nc, int grid, size_t lds);
template <int R>
hipError_t launch_xdct_fwd_r(hipStream_t s, int mode, const XPar<double2>& q, dim3 grid, size_t lds);
template <int R>
hipError_t launch_xdct_inv_r(hipStream_t s, const XPar<double2>& q, dim3 grid, size_t lds);

// ---------------------------------------------------------------------------
// host side: parameters, LDS sizes, launches
// ---------------------------------------------------------------------------
bool xfwd_wide(const FftDesc& fx, int R, int T) {
    return (long)R * T * fx.n > (long)inplace_capacity(fx, 512, 2, true);  // 512 threads per x workgroup
}

template <typename C>
void xpar_columns(XPar<C>& q, const dfp_plan* p, int inv) {
    const Geo& g = p->geo;
    q.T = inv ? p->tk_xi : p->tk_x;
    q.nbx = (g.nzp / p->zr) / q.T;
    // A row of tiles moves about 24 * nxp * nzp bytes of u*, v*, w* (or phi and
    // b); from nxp = 512 up that exceeds an XCD's 4 MB L2 before the next row
    // reads it again as its halo, so the workgroups walk strips of 8 column
    // blocks instead (measured at 512x384x512: the FP32 form's residual pass
    // ran 4 % of a projection faster, FP64 unchanged; at 256x192x256
    // row-major order was 0.7 % faster).
    q.strip = g.nxp >= 512 ? std::min(q.nbx, 8) : q.nbx;
    q.div_t = FastDiv(q.T);
}
template void xpar_columns<double2>(XPar<double2>&, const dfp_plan*, int);
template void xpar_columns<float2>(XPar<float2>&, const dfp_plan*, int);

template <typename C>
XPar<C> make_xpar(const dfp_plan* p, int inv) {
    XPar<C> q{};
    const Geo& g = p->geo;
    q.g = g;
    q.o = p->ops;
    q.fd = p->fx;
    q.R = p->zr;
    xpar_columns(q, p, inv);
    q.npart0 = (g.nzp / p->zr) / p->tk_xi;
    q.ntw = twiddle_count(p->fx);
    q.div_nxp = FastDiv(g.nxp);
    if (g.xbc == 1) {
        // DCT-IV kernels: lines of M = nxp/2 packed points; REDFT11 twice is 2 nxp
        q.div_half = FastDiv(p->fx.n);
        q.scale = 1.0 / (2.0 * (double)g.nxp * (double)g.nzp);
    } else {
        q.div_half = FastDiv(p->fx.n / 2 + 1);
        // the duct solves z in physical space: Real()/nxp (projection.f90:326-331)
        q.scale = g.zbc == 1 ? 1.0 / (double)g.nxp : 1.0 / ((double)g.nxp * (double)g.nzp);
    }
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
template XPar<double2> make_xpar<double2>(const dfp_plan*, int);
template XPar<float2> make_xpar<float2>(const dfp_plan*, int);

// dynamic LDS of the forward (inv = 0) or inverse (inv = 1) x kernel of a plan
size_t x_lds_bytes(const dfp_plan* p, int inv, int f32) {
    const int R = p->zr, T = inv ? p->tk_xi : p->tk_x, nz2 = p->geo.nzp / R, M = p->fx.n;
    const size_t cs = f32 ? sizeof(float2) : sizeof(double2);
    if (p->geo.xbc == 1)  // DCT-IV: M packed points per line, pre- and post-twiddles
        return cs * ((size_t)R * T * M + twiddle_count(p->fx) + 2 * M + (size_t)nz2 * (R - 1));
    return cs * ((size_t)R * T * (M + 1) + twiddle_count(p->fx) + M + 1 + (size_t)nz2 * (R - 1));
}

static const char* xname(int mode, bool f32) {
    static const char* names[2][3] = {{"div_rfft_x", "rfft_x_rhs", "res_rfft_x"},
                                      {"div_rfft_x32", "rfft_x_rhs32", "res_rfft_x32"}};
    return names[f32 ? 1 : 0][mode];
}

#define DFP_R_SWITCH(R, CALL)                                                          \
    switch (R) {                                                                       \
        case 1: return CALL(1);                                                        \
        case 2: return CALL(2);                                                        \
        case 3: return CALL(3);                                                        \
        case 4: return CALL(4);                                                        \
        case 6: return CALL(6);                                                        \
        case 8: return CALL(8);                                                        \
        case 9: return CALL(9);                                                        \
        case 12: return CALL(12);                                                      \
        case 16: return CALL(16);                                                      \
        default: return hipErrorInvalidValue;                                          \
    }

template <typename C>
static hipError_t xfwd_any(int R, hipStream_t s, int mode, const XPar<C>& q, dim3 grid, size_t lds) {
#define DFP_CALL(RV) launch_xfwd_r<RV, C>(s, mode, q, grid, lds)
    DFP_R_SWITCH(R, DFP_CALL)
#undef DFP_CALL
}

template <typename C>
static hipError_t xinv_any(int R, hipStream_t s, int accum, const XPar<C>& q, dim3 grid, size_t lds) {
#define DFP_CALL(RV) launch_xinv_r<RV, C>(s, accum, q, grid, lds)
    DFP_R_SWITCH(R, DFP_CALL)
#undef DFP_CALL
}

template <typename C>
hipError_t launch_xfwd_batch(hipStream_t s, int R, int mode, const XPar<C>* cs, const int* off, const int* idx,
                             const int* nc, int grid, size_t lds) {
#define DFP_CALL(RV) launch_xfwd_batch_r<RV, C>(s, mode, cs, off, idx, nc, grid, lds)
    DFP_R_SWITCH(R, DFP_CALL)
#undef DFP_CALL
}
template hipError_t launch_xfwd_batch<double2>(hipStream_t, int, int, const XPar<double2>*, const int*, const int*,
                                               const int*, int, size_t);
template hipError_t launch_xfwd_batch<float2>(hipStream_t, int, int, const XPar<float2>*, const int*, const int*,
                                              const int*, int, size_t);

template <typename C>
hipError_t launch_xinv_batch(hipStream_t s, int R, int accum, const XPar<C>* cs, const int* off, const int* idx,
                             const int* nc, int grid, size_t lds) {
#define DFP_CALL(RV) launch_xinv_batch_r<RV, C>(s, accum, cs, off, idx, nc, grid, lds)
    DFP_R_SWITCH(R, DFP_CALL)
#undef DFP_CALL
}
template hipError_t launch_xinv_batch<double2>(hipStream_t, int, int, const XPar<double2>*, const int*, const int*,
                                               const int*, int, size_t);
template hipError_t launch_xinv_batch<float2>(hipStream_t, int, int, const XPar<float2>*, const int*, const int*,
                                              const int*, int, size_t);

static hipError_t xdct_fwd_any(int R, hipStream_t s, int mode, const XPar<double2>& q, dim3 grid, size_t lds) {
#define DFP_CALL(RV) launch_xdct_fwd_r<RV>(s, mode, q, grid, lds)
    DFP_R_SWITCH(R, DFP_CALL)
#undef DFP_CALL
}

static hipError_t xdct_inv_any(int R, hipStream_t s, const XPar<double2>& q, dim3 grid, size_t lds) {
#define DFP_CALL(RV) launch_xdct_inv_r<RV>(s, q, grid, lds)
    DFP_R_SWITCH(R, DFP_CALL)
#undef DFP_CALL
}

template <typename C>
static hipError_t launch_xfwd(dfp_plan* p, int mode, const XPar<C>& q) {
    const dim3 grid(q.nbx, q.g.nym);
    const size_t lds = x_lds_bytes(p, 0, sizeof(C) == 8);
    const int slot = prof_begin(p, xname(mode, sizeof(C) == 8));
    const hipError_t e = xfwd_any<C>(p->zr, p->stream, mode, q, grid, lds);
    prof_end(p, slot);
    return e;
}

hipError_t launch_div_rfft_x(dfp_plan* p, const double* U, const double* V, const double* W, const double* rhs,
                             int from_rhs, const double* phib) {
    XPar<double2> q = make_xpar<double2>(p);
    q.a.U = U; q.a.V = V; q.a.W = W; q.a.rhs = rhs;
    if (p->geo.xbc == 1) {
        q.a.phib = from_rhs ? nullptr : phib;
        q.a.idx2 = 1.0 / std::pow(p->dx, 2.0);
        const dim3 grid(q.nbx, q.g.nym);
        const size_t lds = x_lds_bytes(p, 0, 0);
        const int slot = prof_begin(p, from_rhs ? "dct_x_rhs" : "div_dct_x");
        const hipError_t e = xdct_fwd_any(p->zr, p->stream, from_rhs ? 1 : 0, q, grid, lds);
        prof_end(p, slot);
        return e;
    }
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
    const hipError_t e = xinv_any<C>(p->zr, p->stream, ACCUM, q, grid, lds);
    prof_end(p, slot);
    return e;
}

hipError_t launch_irfft_x(dfp_plan* p, double* rhs, double* P, double dtc, int save_p) {
    XPar<double2> q = make_xpar<double2>(p, 1);
    q.rhs = rhs; q.P = P; q.dtc = dtc; q.save_p = save_p;
    if (p->geo.xbc == 1) {
        const dim3 grid(q.nbx, q.g.nym);
        const size_t lds = x_lds_bytes(p, 1, 0);
        const int slot = prof_begin(p, "idct_x");
        const hipError_t e = xdct_inv_any(p->zr, p->stream, q, grid, lds);
        prof_end(p, slot);
        return e;
    }
    return launch_xinv<double2, 0>(p, q);
}

hipError_t launch_irfft_x32(dfp_plan* p, double* rhs, int accumulate, const double* pbar, double* part0, double* P,
                            double dtc, int save_p) {
    XPar<float2> q = make_xpar<float2>(p, 1);
    q.rhs = rhs; q.pbar = pbar; q.part0 = part0;
    q.P = P; q.dtc = save_p ? dtc : 1.0; q.save_p = save_p;
    if (accumulate) return launch_xinv<float2, 1>(p, q);
    return launch_xinv<float2, 0>(p, q);
}

}  // namespace dfp
