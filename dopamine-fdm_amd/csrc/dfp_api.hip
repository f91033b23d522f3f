// C ABI of the dfp projection library: plans, entry points, profiling, errors.
//
// Plan set-up restates the reference's operator construction
// (initialization.f90:416-434 kxx/kzz, :506-535 Dyy) so that the library
// needs only the grid; the spectral solve keeps the x modes kx = 0..nxp/2 of
// the real-to-complex transform, as the reference's GPU build does
// (poisson_gpu.f90:92-97).
#include <cmath>
#include <cstdarg>
#include <algorithm>
#include <limits>
#include <numeric>
#include <cstdio>
#include <cstring>
#include <new>

#include "fft_lds.h"
#include "symtri_eigen.h"

using namespace dfp;

namespace {

int fail(dfp_plan* p, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(dfp_plan* p, int code, const char* fmt, ...) {
    if (p) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        p->last_error = buf;
    }
    return code;
}

#define HIPCHK(p, expr)                                                                                 \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail((p), DFP_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

// radices 4, 2, 3 for n = 2^a 3^b; false if another prime divides n
bool factor(int n, FftDesc& d) {
    d = FftDesc{};
    d.n = n;
    int m = n, nst = 0;
    while (m % 4 == 0) { if (nst >= kMaxStages) return false; d.rad[nst++] = 4; m /= 4; }
    while (m % 2 == 0) { if (nst >= kMaxStages) return false; d.rad[nst++] = 2; m /= 2; }
    while (m % 3 == 0) { if (nst >= kMaxStages) return false; d.rad[nst++] = 3; m /= 3; }
    d.nst = nst;
    int ncur = n, stride = 1;
    for (int st = 0; st < nst; ++st) {
        d.nbdiv[st] = FastDiv((uint32_t)(ncur / d.rad[st] * stride));
        d.sdiv[st] = FastDiv((uint32_t)stride);
        stride *= d.rad[st];
        ncur /= d.rad[st];
    }
    return m == 1;
}

// twiddles exp(-2 pi i t / n) from the exact index, rounded from long double
double2 expi(long t, long n) {
    long tt = t % n;
    if (tt < 0) tt += n;
    const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)tt / (long double)n;
    return make_double2((double)cosl(a), (double)(-sinl(a)));
}

void stage_twiddles(FftDesc& d, std::vector<double2>& tw) {
    int ncur = d.n;
    for (int st = 0; st < d.nst; ++st) {
        const int r = d.rad[st], m = ncur / r;
        d.twoff[st] = (int)tw.size();
        for (int p = 0; p < m; ++p)
            for (int u = 1; u < r; ++u) tw.push_back(expi((long)p * u, ncur));
        ncur = m;
    }
}

// initialization.f90:420-425 (and :428-433 for z)
void wavenumbers(int n, double d, std::vector<double>& k) {
    const double pi = 4.0 * std::atan(1.0);
    k.assign(n, 0.0);
    const int half = (int)std::ceil((double)(float)n / 2.0);
    for (int i = 0; i <= half && i <= n - 1; ++i) k[i] = 2.0 * (std::cos(2.0 * pi * (double)i / (double)n) - 1.0) / std::pow(d, 2.0);
    for (int i = half + 1; i <= n - 1; ++i)
        k[i] = 2.0 * (std::cos(2.0 * pi * (double)(-n + i) / (double)n) - 1.0) / std::pow(d, 2.0);
}

// initialization.f90:443-445: DCT-IV (quarter-wave) modified wavenumbers of the
// inflow/outflow x, kxx(i) = 2 (cos(pi (2i+1) / (2n)) - 1) / d^2, i = 0..n-1
void wavenumbers_dct(int n, double d, std::vector<double>& k) {
    const double pi = 4.0 * std::atan(1.0);
    k.assign(n, 0.0);
    for (int i = 0; i < n; ++i)
        k[i] = 2.0 * (std::cos(pi * (double)(2 * i + 1) / (2.0 * (double)n)) - 1.0) / std::pow(d, 2.0);
}

// initialization.f90:506-535, rows j = 2..nyg-1 stored at j-2
void dyy(const std::vector<double>& y, const std::vector<double>& yg, std::vector<double>& dl, std::vector<double>& dd,
         std::vector<double>& du) {
    const int nyg = (int)yg.size(), nym = nyg - 2;
    dl.assign(nym > 1 ? nym - 1 : 1, 0.0);
    du.assign(nym > 1 ? nym - 1 : 1, 0.0);
    dd.assign(nym, 0.0);
    auto Y = [&](int j) { return y[j - 1]; };
    auto G = [&](int j) { return yg[j - 1]; };
    for (int j = 2; j <= nyg - 1; ++j) {
        const double a = 1.0 / (Y(j) - Y(j - 1)) / (G(j + 1) - G(j));
        const double b = 1.0 / (Y(j) - Y(j - 1)) * (-1.0 / (G(j + 1) - G(j)) - 1.0 / (G(j) - G(j - 1)));
        const double c = 1.0 / (Y(j) - Y(j - 1)) / (G(j) - G(j - 1));
        const int jj = j - 2;
        if (j == 2 && j == nyg - 1) {
            // one cell between the walls: the reference assigns the row twice,
            // b+c for the first wall, then a+b for the last, which stays
            dd[jj] = a + b;
        } else if (j == 2) {
            dd[jj] = b + c;
            du[jj] = a;
        } else if (j == nyg - 1) {
            dd[jj] = a + b;
            dl[jj - 1] = c;
        } else {
            du[jj] = a;
            dl[jj - 1] = c;
            dd[jj] = b;
        }
    }
}

// Launch shape of a plan: z-split radix R (the x kernels apply the first
// radix-R stage of the z FFT), x columns per workgroup T, and either the
// fused z kernel (whole (kx,u) sub-plane in LDS) or the split pair of z
// kernels with G rows per group.
struct Shape {
    int R = 1, T = 1, Ti = 1, G = 1, fused = 0;  // T, Ti: columns of the forward, inverse x kernels
    int gy = 1;                                  // periodic y: columns per y-transform call
    FftDesc fx{}, fz{}, fy{};
};

constexpr int kRadices[] = {1, 2, 3, 4, 6, 8, 9, 12, 16};

// x columns per workgroup of the forward (inv = 0) or inverse x kernel: the
// largest divisor of nz2 within about 2048 points, or the caller's, if the
// kernel fits; 0 if none does.  Where the x lines are long the workgroup takes
// as many points as its LDS transform holds (4096 when every stage is radix 4,
// or radix 2 paired), for longer spectrum segments: from nxp = 256 up in the
// forward kernels (which pair their radix-2 butterflies) and in the FP32
// inverse, from nxp = 512 up in the FP64 inverse; from nxp = 512 the FP32
// forward kernel takes up to 8192 points (four butterfly slots per thread).
// Measured: 512x384x512 FP64 3.90 -> 3.58 ms, FP32+refine 6.12 -> 5.47 ms
// (8192-point FP32 forward tiles: forward kernels -12 % and -23 %);
// 256x192x256 FP32+refine 650 -> 628 us, the FP64 forward kernel
// 113 -> 106 us (its inverse with 8 columns holds 66 KB of LDS, two
// workgroups per CU instead of three, and lost 8 us, so it keeps 4).
// dct: the DCT-IV kernels of the inflow/outflow x (M points per line, M pre-
// and M post-twiddles; the inverse holds at most 4096 points)
int x_columns(const Shape& sh, int nxp, int nzp, size_t lds_limit, size_t cs, int want, int inv, int dct = 0) {
    const int M = nxp / 2, nz2 = nzp / sh.R;
    const bool f32 = cs == sizeof(float2);
    const bool pair = f32 || !inv;         // the kernels that pair radix-2 butterflies
    const int slots = f32 && !inv ? 4 : 2;  // FP32 forward tiles may take four butterfly slots per thread
    const int cap = inplace_capacity(sh.fx, 512, slots, pair);
    auto fits = [&](int T) {
        const size_t pts = (size_t)sh.R * T;
        if (dct)
            return pts * M <= (size_t)cap && (!inv || pts * M <= 4096) &&
                   cs * (pts * M + twiddle_count(sh.fx) + 2 * M + (size_t)nz2 * (sh.R - 1)) + 512 <= lds_limit;
        return pts * M <= (size_t)cap &&
               cs * (pts * (M + 1) + twiddle_count(sh.fx) + M + 1 + (size_t)nz2 * (sh.R - 1)) + 512 <= lds_limit;
    };
    if (want > 0) return (nz2 % want == 0 && fits(want)) ? want : 0;
    const int pts = nxp >= 512 ? cap : nxp >= 256 && pair ? std::min(cap, 4096) : 2048;
    for (int t = std::max(1, std::min(nz2, pts / (sh.R * M))); t >= 1; --t)
        if (nz2 % t == 0 && fits(t)) return t;
    return 0;
}

// 0 on success, else a DFP_E_* code
int choose_shape(const dfp_grid* g, int f32, size_t lds_limit, const dfp_tuning* tune, Shape& out) {
    const size_t cs = f32 ? sizeof(float2) : sizeof(double2);
    const int want_R = tune ? tune->z_split : 0, want_T = tune ? tune->x_columns : 0;
    const int want_G = tune ? tune->z_rows : 0, want_f = tune ? tune->z_fused : 0;
    if (want_R < 0 || want_T < 0 || want_G < 0 || want_f < 0 || want_f > 2) return DFP_E_ARG;
    if (g->z_bc_type == 1) {
        // z walls: no z transform (R = 1), the z eigenbasis kernels instead
        // (with a periodic y, the y transform of k_yfft_solve between them)
        if ((want_R > 1) || want_G > 0 || want_f > 0) return DFP_E_UNSUPPORTED;
        Shape sh;
        if (!factor(g->nxp / 2, sh.fx)) return DFP_E_UNSUPPORTED;
        factor(1, sh.fz);
        if (g->y_bc_type == 0) {
            if (!factor(g->nym - 1, sh.fy)) return DFP_E_UNSUPPORTED;
            sh.gy = yfft_solve_columns(sh.fy, g->nym, lds_limit);
            if (sh.gy == 0) return DFP_E_UNSUPPORTED;
        }
        sh.T = x_columns(sh, g->nxp, g->nzp, lds_limit, cs, want_T, 0);
        sh.Ti = x_columns(sh, g->nxp, g->nzp, lds_limit, cs, want_T, 1);
        if (sh.T == 0 || sh.Ti == 0) return DFP_E_UNSUPPORTED;
        out = sh;
        return DFP_OK;
    }
    if (want_R > 0) {
        bool known = false;
        for (int r : kRadices) known = known || r == want_R;
        if (!known || g->nzp % want_R != 0) return DFP_E_ARG;
    }
    // the periodic y (y FFT in place of the wall solve) has the fused kernel only
    const bool py = g->y_bc_type == 0;
    FftDesc fy{};
    if (py && (want_f == 1 || want_G > 0)) return DFP_E_UNSUPPORTED;
    if (py && !factor(g->nym - 1, fy)) return DFP_E_UNSUPPORTED;
    const bool allow_fused = want_f == 2 || (want_f == 0 && want_G == 0);
    const bool allow_split = want_f != 2 && !py;
    std::vector<Shape> fused, split;
    for (int R : kRadices) {
        if (g->nzp % R != 0 || (want_R > 0 && R != want_R)) continue;
        Shape sh;
        sh.R = R;
        if (!factor(g->nxp / 2, sh.fx) || !factor(g->nzp / R, sh.fz)) continue;
        const int dct = g->x_bc_type == 1 ? 1 : 0;
        sh.T = x_columns(sh, g->nxp, g->nzp, lds_limit, cs, want_T, 0, dct);
        sh.Ti = x_columns(sh, g->nxp, g->nzp, lds_limit, cs, want_T, 1, dct);
        if (sh.T == 0 || sh.Ti == 0 || sh.fz.n > 256) continue;
        if (py) {
            const int gy = zf_py_columns(fy, sh.fz.n);
            if (gy > 0 && zf_py_lds_bytes(g->nym, sh.fz.n, twiddle_count(sh.fz), twiddle_count(fy)) <= lds_limit) {
                Shape f = sh;
                f.fused = 1;
                f.G = zf_rows_per_call(f.fz);
                f.fy = fy;
                f.gy = gy;
                fused.push_back(f);
            }
        } else if (allow_fused && zf_lds_bytes(g->nym, sh.fz.n, twiddle_count(sh.fz), f32) <= lds_limit) {
            Shape f = sh;
            f.fused = 1;
            f.G = zf_rows_per_call(f.fz);
            fused.push_back(f);
        }
        if (allow_split) {
            Shape q = sh;
            q.G = zsolve_rows_per_group(q.fz, g->nym, lds_limit);
            if (want_G > 0) q.G = want_G <= q.G ? want_G : 0;
            if (q.G >= 1) split.push_back(q);
        }
    }
    // default: the split pair with the classic split (R = 4 when it leaves at
    // least 8 points per row, else 2, else 3, else none); the fused form when
    // asked for (or when the split pair cannot run).  Measured at
    // 256x192x256 the split pair is faster: the fused kernel's sweeps run on
    // nzp/R lanes per workgroup only.
    if (want_f == 2 && !fused.empty()) { out = fused[0]; return DFP_OK; }
    if (!split.empty()) {
        if (want_R > 0) { out = split[0]; return DFP_OK; }
        const int nzp = g->nzp;
        int pick = 1;
        if (nzp % 4 == 0 && nzp / 4 >= 8) pick = 4;
        else if (nzp % 2 == 0 && nzp / 2 >= 4) pick = 2;
        else if (nzp % 3 == 0 && nzp / 3 >= 4) pick = 3;
        for (const Shape& q : split)
            if (q.R == pick) { out = q; return DFP_OK; }
        out = split[0];
        return DFP_OK;
    }
    if (!fused.empty()) { out = fused.back(); return DFP_OK; }
    return DFP_E_UNSUPPORTED;
}

bool finite_all(const double* a, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

void free_plan_memory(dfp_plan* p) {
    (void)hipFree(p->d_tables);
    (void)hipFree(p->d_tw);
    (void)hipFree(p->d_tw32);
    (void)hipFree(p->d_b);
    (void)hipFree(p->d_rowpart);
    (void)hipFree(p->d_part0);
    (void)hipFree(p->d_pbar);
    (void)hipFree(p->d_spec);
    (void)hipFree(p->d_lu);
    (void)hipFree(p->d_lucol);
    (void)hipFree(p->d_spec2);
    (void)hipFree(p->d_duct);
    (void)hipFree(p->d_z00);
    (void)hipFree(p->d_red);
    (void)hipFree(p->d_U);
    (void)hipFree(p->d_V);
    (void)hipFree(p->d_W);
    (void)hipFree(p->d_rhs);
    (void)hipFree(p->d_P);
    (void)hipFree(p->d_pb);
    (void)hipFree(p->d_phib);
    for (auto& pe : p->pending) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    for (auto ev : p->event_pool) (void)hipEventDestroy(ev);
    if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
    if (p->join_ev) (void)hipEventDestroy(p->join_ev);
    if (p->aux) (void)hipStreamDestroy(p->aux);
}

// the outlet backflow term is on (dfp_set_outflow with beta > 0)
bool backflow(const dfp_plan* p) { return p->geo.xbc == 1 && p->beta > 0.0; }

// device staging of Pb and phib for host-resident calls
int ensure_outflow_staging(dfp_plan* p) {
    if (p->d_pb) return DFP_OK;
    const size_t n = (size_t)(p->geo.nyg - 2) * (size_t)(p->geo.nzg - 2);
    HIPCHK(p, hipMalloc(&p->d_pb, sizeof(double) * n));
    HIPCHK(p, hipMalloc(&p->d_phib, sizeof(double) * n));
    HIPCHK(p, hipMemset(p->d_phib, 0, sizeof(double) * n));
    return DFP_OK;
}

int64_t face_size(const dfp_plan* p) { return (int64_t)(p->geo.nyg - 2) * (p->geo.nzg - 2); }

int ensure_staging(dfp_plan* p) {
    if (p->d_U) return DFP_OK;
    HIPCHK(p, hipMalloc(&p->d_U, sizeof(double) * p->sz_u));
    HIPCHK(p, hipMalloc(&p->d_V, sizeof(double) * p->sz_v));
    HIPCHK(p, hipMalloc(&p->d_W, sizeof(double) * p->sz_w));
    HIPCHK(p, hipMalloc(&p->d_rhs, sizeof(double) * p->sz_rhs));
    HIPCHK(p, hipMalloc(&p->d_P, sizeof(double) * p->sz_p));
    HIPCHK(p, hipMemset(p->d_rhs, 0, sizeof(double) * p->sz_rhs));
    return DFP_OK;
}

int h2d(dfp_plan* p, double* dst, const double* src, int64_t n) {
    HIPCHK(p, hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, p->stream));
    return DFP_OK;
}

int d2h(dfp_plan* p, double* dst, const double* src, int64_t n) {
    HIPCHK(p, hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, p->stream));
    return DFP_OK;
}

// the monitor's NaN/overflow guard (monitor.f90:32-38) on device fields:
// DFP_E_NONFINITE if a value is NaN or above 1e8 in magnitude
int guard_finite(dfp_plan* p, const double* U, const double* V, const double* W) {
    double* d_out = p->d_red + p->red_len;
    double bad = 0.0;
    HIPCHK(p, launch_guard_finite(p, U, V, W, d_out));
    HIPCHK(p, hipMemcpyAsync(&bad, d_out, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(p, hipStreamSynchronize(p->stream));
    if (bad != 0.0) return fail(p, DFP_E_NONFINITE, "u* holds NaN or a value above 1e8 in magnitude");
    return DFP_OK;
}

int finish(dfp_plan* p, int flags) {
    HIPCHK(p, hipGetLastError());
    if (!(flags & DFP_ASYNC) || !(flags & DFP_FIELDS_ON_DEVICE)) HIPCHK(p, hipStreamSynchronize(p->stream));
    return DFP_OK;
}

int check_flags(dfp_plan* p, int flags, int extra = 0) {
    if (flags & ~(DFP_FIELDS_ON_DEVICE | DFP_SAVE_P | DFP_ASYNC | extra))
        return fail(p, DFP_E_ARG, "unknown flag bits 0x%x", flags);
    HIPCHK(p, hipSetDevice(p->device));
    return DFP_OK;
}

// z transforms and y solves of the spectrum: the fused kernel or the split pair
hipError_t z_solve(dfp_plan* p, int f32) {
    if (p->geo.zbc == 1) return launch_duct_solve(p);
    if (p->geo.ybc == 0) return launch_zfused_py(p);
    if (p->zfused) {
        if (f32) return launch_zfused<float2>(p, make_zpar<float2>(p, 1));
        return launch_zfused<double2>(p, make_zpar<double2>(p, 0));
    }
    // elimination and back substitution of each sub-plane in one workgroup
    return f32 ? launch_zsolve32(p) : launch_zsolve(p);
}

// The (0,0)-mode solve (row sums -> pbar) and the z solves (the spectrum,
// whose (0,0) mode they leave at zero) touch disjoint data, so the former runs
// on the plan's second stream beside the latter and joins before the inverse x
// transform reads pbar.  With per-kernel profiling on, both stay in order on
// the plan's stream.
int mode00_and_z(dfp_plan* p) {
    if (p->prof) {
        HIPCHK(p, launch_mode00(p, p->stream));
        HIPCHK(p, z_solve(p, 1));
        return DFP_OK;
    }
    HIPCHK(p, hipEventRecord(p->fork_ev, p->stream));
    HIPCHK(p, hipStreamWaitEvent(p->aux, p->fork_ev, 0));
    HIPCHK(p, launch_mode00(p, p->aux));
    HIPCHK(p, hipEventRecord(p->join_ev, p->aux));
    HIPCHK(p, z_solve(p, 1));
    HIPCHK(p, hipStreamWaitEvent(p->stream, p->join_ev, 0));
    return DFP_OK;
}

// FP32 form: one FP32 spectral solve with the (0,0) mode in FP64, then
// p->sweeps refinement sweeps on the FP64 residual b - L_pin phi
int mixed_solve(dfp_plan* p, const double* U, const double* V, const double* W, double* rhs, double* P, double dtc,
                int save_p, int from_fields) {
    XArgs a{};
    a.U = U; a.V = V; a.W = W; a.rhs = rhs;
    a.b_out = p->d_b;
    a.rowpart = p->d_rowpart;
    HIPCHK(p, launch_rfft_x32(p, from_fields ? 0 : 1, a));
    if (int rc = mode00_and_z(p)) return rc;
    // P = phi/(dt*c) is written by the last inverse transform
    HIPCHK(p, launch_irfft_x32(p, rhs, 0, p->d_pbar, p->d_part0, P, dtc, save_p && p->sweeps == 0));
    for (int s = 0; s < p->sweeps; ++s) {
        XArgs r{};
        r.rhs = rhs;
        r.b = p->d_b;
        r.rowpart = p->d_rowpart;
        r.part0 = p->d_part0;
        r.idx2 = 1.0 / (p->dx * p->dx);
        r.idz2 = 1.0 / (p->dz * p->dz);
        HIPCHK(p, launch_rfft_x32(p, 2, r));
        if (int rc = mode00_and_z(p)) return rc;
        HIPCHK(p, launch_irfft_x32(p, rhs, 1, p->d_pbar, p->d_part0, P, dtc, save_p && s == p->sweeps - 1));
    }
    return DFP_OK;
}

// the spectral chain on device arrays: rhs (read unless from_fields) -> phi in rhs (+P);
// phib: outlet face values of the backflow term (inflow/outflow x, from the fields only)
int spectral_solve(dfp_plan* p, const double* U, const double* V, const double* W, double* rhs, double* P, double dtc,
                   int save_p, int from_fields, const double* phib = nullptr) {
    if (p->precision == DFP_FP32_REFINE) return mixed_solve(p, U, V, W, rhs, P, dtc, save_p, from_fields);
    HIPCHK(p, launch_div_rfft_x(p, U, V, W, rhs, from_fields ? 0 : 1, phib));
    HIPCHK(p, z_solve(p, 0));
    HIPCHK(p, launch_irfft_x(p, rhs, P, dtc, save_p));
    return DFP_OK;
}

}  // namespace

namespace dfp {

int prof_begin(dfp_plan* p, const char* name) {
    if (!p->prof) return -1;
    int slot = -1;
    for (size_t s = 0; s < p->slots.size(); ++s)
        if (p->slots[s].name == name) slot = (int)s;
    if (slot < 0) {
        p->slots.push_back(ProfSlot{name, 0, 0.0});
        slot = (int)p->slots.size() - 1;
    }
    hipEvent_t a, b;
    if (p->event_pool.size() >= 2) {
        a = p->event_pool.back(); p->event_pool.pop_back();
        b = p->event_pool.back(); p->event_pool.pop_back();
    } else {
        if (hipEventCreate(&a) != hipSuccess) return -1;
        if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return -1; }
    }
    (void)hipEventRecord(a, p->stream);
    p->pending.push_back({slot, {a, b}});
    return (int)p->pending.size() - 1;
}

void prof_end(dfp_plan* p, int idx) {
    if (idx < 0 || idx >= (int)p->pending.size()) return;
    (void)hipEventRecord(p->pending[idx].second.second, p->stream);
}

}  // namespace dfp

static int prof_collect(dfp_plan* p) {
    for (auto& pe : p->pending) {
        HIPCHK(p, hipEventSynchronize(pe.second.second));
        float ms = 0.f;
        HIPCHK(p, hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
        p->slots[pe.first].launches += 1;
        p->slots[pe.first].ms += ms;
        p->event_pool.push_back(pe.second.first);
        p->event_pool.push_back(pe.second.second);
    }
    p->pending.clear();
    return DFP_OK;
}

extern "C" {

int dfp_plan_create(const dfp_grid* g, int precision, int device, dfp_plan** out) {
    return dfp_plan_create_tuned(g, precision, device, nullptr, out);
}

int dfp_plan_tuning(const dfp_plan* p, dfp_tuning* t) {
    if (!p || !t) return DFP_E_ARG;
    t->z_split = p->zr;
    t->x_columns = p->tk_x;
    t->z_rows = p->g_zs;
    t->z_fused = p->zfused ? 2 : 1;
    return DFP_OK;
}

int dfp_plan_create_tuned(const dfp_grid* g, int precision, int device, const dfp_tuning* tune, dfp_plan** out) {
    if (!out) return DFP_E_ARG;
    *out = nullptr;
    if (!g || !g->y || !g->z || !g->yg || !g->zg) return DFP_E_ARG;
    if (g->x_bc_type != 0 && g->x_bc_type != 1) return DFP_E_UNSUPPORTED;
    if ((g->y_bc_type != 0 && g->y_bc_type != 1) || (g->z_bc_type != 0 && g->z_bc_type != 1)) return DFP_E_UNSUPPORTED;
    // z walls: the 4-wall duct with a periodic x and y walls (initialization.f90:98-100), FP64
    if (g->z_bc_type == 1 && (g->x_bc_type != 0 || precision != DFP_FP64)) return DFP_E_UNSUPPORTED;
    // the pinned kx = ky = 0 line of the z walls with a periodic y is solved in one workgroup's LDS
    if (g->z_bc_type == 1 && g->y_bc_type == 0 && g->nzp > 2048) return DFP_E_UNSUPPORTED;
    // periodic y (initialization.f90:96-97, :128-129): periodic x, a uniform
    // y grid and nym-1 = 2^a 3^b independent rows; solved in FP64
    if (g->y_bc_type == 0 && (g->x_bc_type != 0 || precision != DFP_FP64 || g->nym < 2))
        return DFP_E_UNSUPPORTED;
    if (precision != DFP_FP64 && precision != DFP_FP32_REFINE) return DFP_E_ARG;
    // the inflow/outflow x is solved in FP64 only
    if (g->x_bc_type == 1 && precision != DFP_FP64) return DFP_E_UNSUPPORTED;
    if (g->nxp < 2 || (g->nxp & 1) || g->nzp < 1) return DFP_E_UNSUPPORTED;
    if (g->nym < 1) return DFP_E_UNSUPPORTED;
    // an x block holds zr lines of nxp/2 complex points in 32 KiB of LDS; a z
    // block holds one mode per thread
    if (g->nxp > 1024 || g->nzp > 1024 || g->nym > 65535) return DFP_E_UNSUPPORTED;
    if (!(g->dx > 0.0) || !(g->dz > 0.0)) return DFP_E_ARG;
    dfp_plan* p = new (std::nothrow) dfp_plan();
    if (!p) return DFP_E_NOMEM;
    auto bail = [&](int code) {
        free_plan_memory(p);
        delete p;
        return code;
    };
    p->device = device;
    p->precision = precision;
    {
        // FFT lengths 2^a 3^b, checked before the device is touched
        FftDesc probe;
        if (!factor(g->nxp / 2, probe) || (g->z_bc_type == 0 && !factor(g->nzp, probe))) return bail(DFP_E_UNSUPPORTED);
        if (g->y_bc_type == 0) {
            if (!factor(g->nym - 1, probe)) return bail(DFP_E_UNSUPPORTED);
            // the reference takes dy = y(2) - y(1) for every row (:481-489)
            const double dy = g->y[1] - g->y[0];
            if (!(dy > 0.0)) return bail(DFP_E_ARG);
            for (int j = 1; j < g->nym + 1; ++j)
                if (std::fabs((g->y[j] - g->y[j - 1]) - dy) > 1e-9 * dy) return bail(DFP_E_UNSUPPORTED);
        }
    }
    if (hipSetDevice(device) != hipSuccess) return bail(DFP_E_HIP);
    int lds_attr = 0;
    if (hipDeviceGetAttribute(&lds_attr, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds_attr > 0)
        p->lds_limit = (size_t)lds_attr;
    Shape sh;
    if (int rc = choose_shape(g, precision == DFP_FP32_REFINE, p->lds_limit, tune, sh)) return bail(rc);
    const int zr = sh.R;
    p->zr = sh.R;
    p->fx = sh.fx;
    p->fz = sh.fz;
    p->tk_x = sh.T;
    p->tk_xi = sh.Ti;
    p->g_zs = sh.G;
    p->zfused = sh.fused;
    p->fy = sh.fy;
    p->gy = sh.gy;
    if (!sh.fused) z_two_pass(sh.fz.n, p->zn1, p->zn2);
    if (hipSetDevice(device) != hipSuccess) return bail(DFP_E_HIP);
    Geo& G = p->geo;
    const bool dct = g->x_bc_type == 1;
    G.xbc = g->x_bc_type;
    G.ybc = g->y_bc_type;
    G.zbc = g->z_bc_type;
    G.nxp = g->nxp; G.nym = g->nym; G.nzp = g->nzp;
    // the periodic x keeps the R2C half of its modes (and drops the duplicate
    // cell, nxp = nxm - 1); the inflow/outflow x keeps all nxm = nxp cells and
    // all nxp DCT-IV modes (initialization.f90:331-338)
    G.nxh = dct ? g->nxp : g->nxp / 2 + 1;
    // z walls keep all nzm = nzp cells (initialization.f90:339-345)
    G.nx = g->nxp + (dct ? 1 : 2); G.ny = g->nym + 1; G.nz = g->nzp + (G.zbc ? 1 : 2);
    G.nxg = G.nx + 1; G.nyg = g->nym + 2; G.nzg = G.nz + 1;
    p->dx = g->dx; p->dz = g->dz;
    G.inv_dx = 1.0 / g->dx;
    if (g->x) p->hx.assign(g->x, g->x + G.nx);
    p->hy.assign(g->y, g->y + G.ny);
    p->hz.assign(g->z, g->z + G.nz);
    p->hyg.assign(g->yg, g->yg + G.nyg);
    p->hzg.assign(g->zg, g->zg + G.nzg);
    if ((!p->hx.empty() && !finite_all(p->hx.data(), G.nx)) || !finite_all(p->hy.data(), G.ny) ||
        !finite_all(p->hz.data(), G.nz) || !finite_all(p->hyg.data(), G.nyg) || !finite_all(p->hzg.data(), G.nzg))
        return bail(DFP_E_NONFINITE);

    // 1-D tables
    std::vector<double> inv_dyf(G.nym), inv_dzf(G.nzp + 1), inv_dyg(G.nym > 1 ? G.nym - 1 : 1, 0.0), inv_dzg(G.nzp);
    for (int j = 2; j <= G.nyg - 1; ++j) inv_dyf[j - 2] = 1.0 / (p->hy[j - 1] - p->hy[j - 2]);
    for (int k = 2; k <= G.nzg - 1; ++k) inv_dzf[k - 2] = 1.0 / (p->hz[k - 1] - p->hz[k - 2]);
    for (int j = 2; j <= G.ny - 1; ++j) inv_dyg[j - 2] = 1.0 / (p->hyg[j] - p->hyg[j - 1]);
    for (int k = 2; k <= G.nz - 1; ++k) inv_dzg[k - 2] = 1.0 / (p->hzg[k] - p->hzg[k - 1]);
    if (dct) wavenumbers_dct(G.nxp, g->dx, p->hkxx);
    else wavenumbers(G.nxp, g->dx, p->hkxx);
    wavenumbers(G.nzp, g->dz, p->hkzz);
    dyy(p->hy, p->hyg, p->hdl, p->hdd, p->hdu);
    // periodic y: the modified wavenumbers of the nym-1 independent rows,
    // spacing y(2) - y(1) (initialization.f90:481-489)
    std::vector<double> hkyy;
    if (G.ybc == 0) wavenumbers(G.nym - 1, p->hy[1] - p->hy[0], hkyy);
    // the duct (initialization.f90:544-611): S = W^(1/2) Dzz W^(-1/2), its
    // eigenvalues lambda_z and eigenvectors Qz; Qf = W^(1/2) Qz takes a plane
    // to the eigenbasis, Qi = Qz^T W^(-1/2) back; the eigenmode of the zero
    // eigenvalue (by magnitude, :207) is pinned at kx = 0
    std::vector<double> duct_tab;
    if (G.zbc == 1) {
        const int n = G.nzp;
        auto Z = [&](int k) { return p->hz[k - 1]; };
        auto ZG = [&](int k) { return p->hzg[k - 1]; };
        std::vector<double> sd(n), so(n > 1 ? n - 1 : 1, 0.0), sw(n), Q;
        for (int k = 2; k <= G.nzg - 1; ++k) {
            const double a = 1.0 / (Z(k) - Z(k - 1)) / (ZG(k + 1) - ZG(k));
            const double b = 1.0 / (Z(k) - Z(k - 1)) * (-1.0 / (ZG(k + 1) - ZG(k)) - 1.0 / (ZG(k) - ZG(k - 1)));
            const double c = 1.0 / (Z(k) - Z(k - 1)) / (ZG(k) - ZG(k - 1));
            const int kk = k - 2;
            sd[kk] = (k == 2 && k == G.nzg - 1) ? a + b : (k == 2 ? b + c : (k == G.nzg - 1 ? a + b : b));
            if (kk + 1 < n) so[kk] = 1.0 / ((ZG(k + 1) - ZG(k)) * std::sqrt((Z(k) - Z(k - 1)) * (Z(k + 1) - Z(k))));
        }
        for (int k = 0; k < n; ++k) sw[k] = std::sqrt(Z(k + 2) - Z(k + 1));
        if (!std::isfinite(std::accumulate(sd.begin(), sd.end(), 0.0)) || !symtri_eigen(n, sd, so, Q))
            return bail(DFP_E_NONFINITE);
        double lmax = 0.0;
        for (double l : sd) lmax = std::max(lmax, std::fabs(l));
        for (int m = 0; m < n && p->pin_m < 0; ++m)
            if (std::fabs(sd[m]) < 1e-8 * lmax) p->pin_m = m;
        duct_tab.resize((size_t)2 * n * n + n);
        for (int k = 0; k < n; ++k)
            for (int m = 0; m < n; ++m) {
                const double q = Q[(size_t)m * n + k];
                duct_tab[(size_t)k * n + m] = sw[k] * q;                    // Qf(k, m)
                duct_tab[(size_t)n * n + (size_t)m * n + k] = q / sw[k];    // Qi(m, k)
            }
        std::copy(sd.begin(), sd.end(), duct_tab.begin() + (size_t)2 * n * n);
    }

    std::vector<double> tab;
    auto put = [&](const std::vector<double>& v) {
        size_t off = tab.size();
        tab.insert(tab.end(), v.begin(), v.end());
        tab.resize((tab.size() + 7) & ~size_t(7), 0.0);
        return off;
    };
    std::vector<double> kxx_half(p->hkxx.begin(), p->hkxx.begin() + G.nxh);
    const size_t o_dyf = put(inv_dyf), o_dzf = put(inv_dzf), o_dyg = put(inv_dyg), o_dzg = put(inv_dzg);
    const size_t o_kxx = put(kxx_half), o_kzz = put(p->hkzz);
    const size_t o_dl = put(p->hdl), o_dd = put(p->hdd), o_du = put(p->hdu);
    const size_t o_kyy = hkyy.empty() ? (size_t)-1 : put(hkyy);
    // z walls with a periodic y: Dzz (Dyy's construction on z, zg; initialization.f90:544-577)
    // and the elimination of its pinned kx = ky = 0 system (D(2) *= 3/2,
    // projection.f90:241): multipliers DL(k-1)/D'(k-1) and inverse pivots 1/D'(k)
    std::vector<double> zl, zd, zu, zmul, zip;
    if (G.zbc == 1 && G.ybc == 0) {
        dyy(p->hz, p->hzg, zl, zd, zu);
        const int n = G.nzp;
        zmul.assign(n, 0.0);
        zip.assign(n, 0.0);
        double dp = 3.0 / 2.0 * zd[0];
        zip[0] = 1.0 / dp;
        for (int k = 1; k < n; ++k) {
            zmul[k] = zl[k - 1] / dp;
            dp = zd[k] - zmul[k] * zu[k - 1];
            zip[k] = 1.0 / dp;
        }
    }
    const size_t o_zl = zd.empty() ? (size_t)-1 : put(zl);
    const size_t o_zd = zd.empty() ? (size_t)-1 : put(zd);
    const size_t o_zu = zd.empty() ? (size_t)-1 : put(zu);
    const size_t o_zm = zd.empty() ? (size_t)-1 : put(zmul);
    const size_t o_zi = zd.empty() ? (size_t)-1 : put(zip);
    const size_t o_x = p->hx.empty() ? (size_t)-1 : put(p->hx);
    const size_t o_y = put(p->hy), o_z = put(p->hz), o_yg = put(p->hyg);
    if (hipMalloc(&p->d_tables, sizeof(double) * tab.size()) != hipSuccess) return bail(DFP_E_NOMEM);
    if (hipMemcpy(p->d_tables, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(DFP_E_HIP);
    double* T = p->d_tables;
    G.inv_dyf = T + o_dyf; G.inv_dzf = T + o_dzf; G.inv_dyg = T + o_dyg; G.inv_dzg = T + o_dzg;
    G.x = o_x == (size_t)-1 ? nullptr : T + o_x;
    G.y = T + o_y; G.z = T + o_z; G.yg = T + o_yg;
    p->ops.kxx = T + o_kxx; p->ops.kzz = T + o_kzz;
    p->ops.dl = T + o_dl; p->ops.dd = T + o_dd; p->ops.du = T + o_du;
    p->ops.kyy = o_kyy == (size_t)-1 ? nullptr : T + o_kyy;
    if (o_zd != (size_t)-1) {
        p->dzz_l = T + o_zl;
        p->dzz_d = T + o_zd;
        p->dzz_u = T + o_zu;
        p->zw_mult = T + o_zm;
        p->zw_ip = T + o_zi;
    }

    // twiddles, in the order the x kernels stage them into LDS with one
    // copy: x stages, R2C/C2R packing (M+1) or the DCT-IV pre-twiddles
    // e^{-i pi (4n+1)/(4 nxp)} and post-twiddles 2 e^{-i pi n/nxp} (M each),
    // first z stage (nz2*(zr-1)); then the z-kernel stages
    std::vector<double2> tw;
    stage_twiddles(p->fx, tw);
    const size_t o_r2c = tw.size();
    if (dct) {
        for (int n = 0; n < G.nxp / 2; ++n) tw.push_back(expi(4L * n + 1, 8L * G.nxp));
        for (int n = 0; n < G.nxp / 2; ++n) {
            const double2 w = expi(n, 2L * G.nxp);
            tw.push_back(make_double2(2.0 * w.x, 2.0 * w.y));
        }
    } else {
        for (int k = 0; k <= G.nxp / 2; ++k) tw.push_back(expi(k, G.nxp));
    }
    const size_t o_zs = tw.size();
    for (int k = 0; k < G.nzp / zr; ++k)
        for (int u = 1; u < zr; ++u) tw.push_back(expi((long)k * u, G.nzp));
    const size_t o_tz = tw.size();
    stage_twiddles(p->fz, tw);
    tw.push_back(make_double2(1.0, 0.0));
    for (int st = 0; st < p->fz.nst; ++st) p->fz.twoff[st] -= (int)o_tz;
    const size_t o_t2 = tw.size();
    for (int e = 0; e < p->fz.n; ++e) tw.push_back(expi(e, p->fz.n));
    const size_t o_ty = tw.size();
    if (G.ybc == 0) {
        stage_twiddles(p->fy, tw);
        for (int st = 0; st < p->fy.nst; ++st) p->fy.twoff[st] -= (int)o_ty;
    }
    if (hipMalloc(&p->d_tw, sizeof(double2) * tw.size()) != hipSuccess) return bail(DFP_E_NOMEM);
    if (hipMemcpy(p->d_tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(DFP_E_HIP);
    p->tw_x = p->d_tw;
    p->tw_z = p->d_tw + o_tz;
    p->tw_r2c = p->d_tw + o_r2c;
    p->tw_zs = p->d_tw + o_zs;
    p->tw_z2 = p->d_tw + o_t2;
    p->tw_y = p->d_tw + o_ty;
    if (precision == DFP_FP32_REFINE) {
        std::vector<float2> tw32(tw.size());
        for (size_t e = 0; e < tw.size(); ++e) tw32[e] = make_float2((float)tw[e].x, (float)tw[e].y);
        if (hipMalloc(&p->d_tw32, sizeof(float2) * tw32.size()) != hipSuccess) return bail(DFP_E_NOMEM);
        if (hipMemcpy(p->d_tw32, tw32.data(), sizeof(float2) * tw32.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(DFP_E_HIP);
        p->tw_x32 = p->d_tw32;
        p->tw_z32 = p->d_tw32 + o_tz;
        p->tw_z2_32 = p->d_tw32 + o_t2;
    }

    const size_t H = (size_t)G.nxh * G.nym * G.nzp;
    if (hipMalloc(&p->d_spec, sizeof(double2) * H) != hipSuccess) return bail(DFP_E_NOMEM);
    if (G.zbc == 1) {  // z walls: eigenbasis tables, second spectrum; the duct's pivots per (kx, eigenmode)
        if (hipMalloc(&p->d_duct, sizeof(double) * duct_tab.size()) != hipSuccess ||
            hipMalloc(&p->d_spec2, sizeof(double2) * H) != hipSuccess)
            return bail(DFP_E_NOMEM);
        if (G.ybc == 1 && hipMalloc(&p->d_lu, sizeof(double) * (size_t)G.nxh * G.nym * G.nzp) != hipSuccess)
            return bail(DFP_E_NOMEM);
        if (G.ybc == 0 && hipMalloc(&p->d_z00, sizeof(double2) * (size_t)G.nzp) != hipSuccess)
            return bail(DFP_E_NOMEM);
        if (hipMemcpy(p->d_duct, duct_tab.data(), sizeof(double) * duct_tab.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(DFP_E_HIP);
        const size_t nn = (size_t)G.nzp * G.nzp;
        p->duct_qf = p->d_duct;
        p->duct_qi = p->d_duct + nn;
        p->duct_lam = p->d_duct + 2 * nn;
        if (G.ybc == 1 && (launch_duct_lu(p) != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess))
            return bail(DFP_E_HIP);
    } else if (G.ybc == 1) {  // pivots of the wall solve
        const size_t nlu = (size_t)G.nxh * G.nym * (G.nzp / 2 + 1);
        if (hipMalloc(&p->d_lu, sizeof(double) * nlu) != hipSuccess) return bail(DFP_E_NOMEM);
        if (launch_lu_table(p) != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess) return bail(DFP_E_HIP);
        // the table's column of each (u, m) mode of the z split
        const int nz2c = G.nzp / p->zr;
        std::vector<int> cols((size_t)G.nzp);
        for (int u = 0; u < p->zr; ++u)
            for (int m = 0; m < nz2c; ++m) cols[(size_t)u * nz2c + m] = lu_pos_mu(m, u, p->zr, nz2c);
        if (hipMalloc(&p->d_lucol, sizeof(int) * cols.size()) != hipSuccess) return bail(DFP_E_NOMEM);
        if (hipMemcpy(p->d_lucol, cols.data(), sizeof(int) * cols.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(DFP_E_HIP);
    }
    p->red_len = (int)((long)ceil_div((long)(G.nxg - 2) * G.nym, 256) * (G.nzg - 2));
    if (hipMalloc(&p->d_red, sizeof(double) * (p->red_len + 8)) != hipSuccess) return bail(DFP_E_NOMEM);

    p->sz_u = (int64_t)G.nx * G.nyg * G.nzg;
    p->sz_v = (int64_t)G.nxg * G.ny * G.nzg;
    p->sz_w = (int64_t)G.nxg * G.nyg * G.nz;
    p->sz_p = (int64_t)G.nxg * G.nyg * G.nzg;
    p->sz_rhs = (int64_t)(G.nxg - 1) * G.nym * (G.nzg - 1);

    const int nz2 = G.nzp / zr;
    if (precision == DFP_FP32_REFINE) {
        const size_t nparts = (size_t)ceil_div(nz2, p->tk_x);     // forward kernel blocks per row
        const size_t nparts_i = (size_t)ceil_div(nz2, p->tk_xi);  // inverse kernel blocks per row
        if (hipMalloc(&p->d_b, sizeof(double) * p->sz_rhs) != hipSuccess ||
            hipMalloc(&p->d_rowpart, sizeof(double) * nparts * G.nym) != hipSuccess ||
            hipMalloc(&p->d_part0, sizeof(double) * nparts_i) != hipSuccess ||
            hipMalloc(&p->d_pbar, sizeof(double) * G.nym) != hipSuccess)
            return bail(DFP_E_NOMEM);
        if (hipStreamCreateWithFlags(&p->aux, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming) != hipSuccess)
            return bail(DFP_E_HIP);
    } else if (G.zbc == 1 && G.ybc == 0) {
        // the pinned kx = ky = 0 line of the z walls with a periodic y runs beside the forward GEMM
        if (hipStreamCreateWithFlags(&p->aux, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming) != hipSuccess)
            return bail(DFP_E_HIP);
    }
    *out = p;
    return DFP_OK;
}

int dfp_plan_destroy(dfp_plan* plan) {
    if (!plan) return DFP_E_ARG;
    (void)hipSetDevice(plan->device);
    (void)hipStreamSynchronize(plan->stream);
    free_plan_memory(plan);
    delete plan;
    return DFP_OK;
}

int dfp_set_refine_sweeps(dfp_plan* p, int sweeps) {
    if (!p) return DFP_E_ARG;
    if (sweeps < 0 || sweeps > 8) return fail(p, DFP_E_ARG, "refinement sweeps must be 0..8, got %d", sweeps);
    p->sweeps = sweeps;
    return DFP_OK;
}

int dfp_set_stream(dfp_plan* p, void* s) {
    if (!p) return DFP_E_ARG;
    p->stream = reinterpret_cast<hipStream_t>(s);
    return DFP_OK;
}

int dfp_field_sizes(const dfp_plan* p, int64_t sizes[5]) {
    if (!p || !sizes) return DFP_E_ARG;
    sizes[0] = p->sz_u; sizes[1] = p->sz_v; sizes[2] = p->sz_w; sizes[3] = p->sz_p; sizes[4] = p->sz_rhs;
    return DFP_OK;
}

int dfp_project(dfp_plan* p, double* U, double* V, double* W, double* rhs_p, double* P, double dt, double c,
                int flags) {
    if (!p) return DFP_E_ARG;
    int rc = check_flags(p, flags, DFP_CHECK_FINITE);
    if (rc) return rc;
    const int save_p = (flags & DFP_SAVE_P) ? 1 : 0;
    if (!U || !V || !W) return fail(p, DFP_E_ARG, "U, V and W are required");
    if (save_p && (!P || !(dt * c != 0.0) || !std::isfinite(dt * c)))
        return fail(p, DFP_E_ARG, "DFP_SAVE_P needs P and a finite nonzero dt*c");
    const bool bf = backflow(p);
    if (bf && (!(dt * c != 0.0) || !std::isfinite(dt * c)))
        return fail(p, DFP_E_ARG, "the outlet backflow term needs a finite nonzero dt*c");
    const bool guard = (flags & DFP_CHECK_FINITE) != 0;
    if (flags & DFP_FIELDS_ON_DEVICE) {
        if (!rhs_p) return fail(p, DFP_E_ARG, "rhs_p is required with DFP_FIELDS_ON_DEVICE");
        if (guard && (rc = guard_finite(p, U, V, W))) return rc;
        // compute_projection_step (projection.f90:21-34): divergence, backflow
        // face values, solve, correction
        if (bf) HIPCHK(p, launch_outflow_phib(p, U, V, W, p->pb, p->phib, nullptr, dt * c));
        rc = spectral_solve(p, U, V, W, rhs_p, P, dt * c, save_p, 1, bf ? p->phib : nullptr);
        if (rc) return rc;
        HIPCHK(p, launch_correct(p, U, V, W, rhs_p, bf ? p->phib : nullptr));
        return finish(p, flags);
    }
    if ((rc = ensure_staging(p))) return rc;
    if (bf && (rc = ensure_outflow_staging(p))) return rc;
    if ((rc = h2d(p, p->d_U, U, p->sz_u)) || (rc = h2d(p, p->d_V, V, p->sz_v)) || (rc = h2d(p, p->d_W, W, p->sz_w)))
        return rc;
    if (guard && (rc = guard_finite(p, p->d_U, p->d_V, p->d_W))) return rc;
    if (rhs_p && (rc = h2d(p, p->d_rhs, rhs_p, p->sz_rhs))) return rc;
    if (bf) {
        if ((rc = h2d(p, p->d_pb, p->pb, face_size(p)))) return rc;
        HIPCHK(p, launch_outflow_phib(p, p->d_U, p->d_V, p->d_W, p->d_pb, p->d_phib, nullptr, dt * c));
    }
    rc = spectral_solve(p, p->d_U, p->d_V, p->d_W, p->d_rhs, p->d_P, dt * c, save_p, 1, bf ? p->d_phib : nullptr);
    if (rc) return rc;
    HIPCHK(p, launch_correct(p, p->d_U, p->d_V, p->d_W, p->d_rhs, bf ? p->d_phib : nullptr));
    if ((rc = d2h(p, U, p->d_U, p->sz_u)) || (rc = d2h(p, V, p->d_V, p->sz_v)) || (rc = d2h(p, W, p->d_W, p->sz_w)))
        return rc;
    if (rhs_p && (rc = d2h(p, rhs_p, p->d_rhs, p->sz_rhs))) return rc;
    if (save_p && (rc = d2h(p, P, p->d_P, p->sz_p))) return rc;
    if (bf && (rc = d2h(p, p->phib, p->d_phib, face_size(p)))) return rc;
    return finish(p, flags);
}

int dfp_set_outflow(dfp_plan* p, double beta, double outflow_uc, const double* Pb, double* phib) {
    if (!p) return DFP_E_ARG;
    if (!(beta >= 0.0) || !std::isfinite(beta) || !std::isfinite(outflow_uc))
        return fail(p, DFP_E_ARG, "outflow_backflow_beta must be finite and >= 0, outflow_Uc finite");
    if (beta > 0.0 && p->geo.xbc != 1)
        return fail(p, DFP_E_ARG, "outflow_backflow_beta > 0 needs x_bc_type = 1");
    if (beta > 0.0 && (!Pb || !phib)) return fail(p, DFP_E_ARG, "outflow_backflow_beta > 0 needs Pb and phib");
    p->beta = beta;
    p->outflow_uc = outflow_uc;
    p->pb = Pb;
    p->phib = phib;
    return DFP_OK;
}

int dfp_outflow_backflow(dfp_plan* p, const double* U, const double* V, const double* W, double* rhs_p, double dt,
                         double c, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = check_flags(p, flags);
    if (rc) return rc;
    if (!U || !V || !W || !rhs_p) return fail(p, DFP_E_ARG, "U, V, W and rhs_p are required");
    if (!backflow(p)) return fail(p, DFP_E_ARG, "no outlet backflow term is set (dfp_set_outflow, beta > 0)");
    if (!(dt * c != 0.0) || !std::isfinite(dt * c))
        return fail(p, DFP_E_ARG, "the outlet backflow term needs a finite nonzero dt*c");
    if (flags & DFP_FIELDS_ON_DEVICE) {
        HIPCHK(p, launch_outflow_phib(p, U, V, W, p->pb, p->phib, rhs_p, dt * c));
        return finish(p, flags);
    }
    if ((rc = ensure_staging(p)) || (rc = ensure_outflow_staging(p))) return rc;
    if ((rc = h2d(p, p->d_U, U, p->sz_u)) || (rc = h2d(p, p->d_V, V, p->sz_v)) || (rc = h2d(p, p->d_W, W, p->sz_w)) ||
        (rc = h2d(p, p->d_rhs, rhs_p, p->sz_rhs)) || (rc = h2d(p, p->d_pb, p->pb, face_size(p))))
        return rc;
    HIPCHK(p, launch_outflow_phib(p, p->d_U, p->d_V, p->d_W, p->d_pb, p->d_phib, p->d_rhs, dt * c));
    if ((rc = d2h(p, rhs_p, p->d_rhs, p->sz_rhs)) || (rc = d2h(p, p->phib, p->d_phib, face_size(p)))) return rc;
    return finish(p, flags);
}

// the step either side of the projection: device-resident fields, channel case
static int step_checks(dfp_plan* p, int flags) {
    if (flags & ~(DFP_FIELDS_ON_DEVICE | DFP_ASYNC | DFP_SAVE_P))
        return fail(p, DFP_E_ARG, "unknown flag bits 0x%x", flags);
    if (!(flags & DFP_FIELDS_ON_DEVICE))
        return fail(p, DFP_E_ARG, "the RK update and boundary conditions take device-resident fields only");
    if (p->geo.xbc != 0)
        return fail(p, DFP_E_UNSUPPORTED, "the boundary conditions here are those of a periodic x");
    HIPCHK(p, hipSetDevice(p->device));
    return DFP_OK;
}

static int rk_args(dfp_plan* p, int stage, const double* U, const double* V, const double* W, const double* Uo,
                   const double* Vo, const double* Wo, const double* const* Fu, const double* const* Fv,
                   const double* const* Fw, const double* coef, double dt) {
    if (stage < 1 || stage > 3) return fail(p, DFP_E_ARG, "stage must be 1, 2 or 3, got %d", stage);
    if (!U || !V || !W || !Uo || !Vo || !Wo || !Fu || !Fv || !Fw || !coef)
        return fail(p, DFP_E_ARG, "U, V, W, Uo, Vo, Wo, Fu, Fv, Fw and coef are required");
    for (int s = 0; s < stage; ++s)
        if (!Fu[s] || !Fv[s] || !Fw[s]) return fail(p, DFP_E_ARG, "right-hand side %d of the stage is null", s + 1);
    for (int s = 0; s < stage; ++s)
        if (!std::isfinite(coef[s])) return fail(p, DFP_E_ARG, "rk_coef(%d,%d) is not finite", stage, s + 1);
    if (!std::isfinite(dt)) return fail(p, DFP_E_ARG, "dt is not finite");
    return DFP_OK;
}

int dfp_rk_update(dfp_plan* p, int stage, double* U, double* V, double* W, const double* Uo, const double* Vo,
                  const double* Wo, const double* const* Fu, const double* const* Fv, const double* const* Fw,
                  const double* coef, double dt, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = step_checks(p, flags & ~DFP_SAVE_P);
    if (rc) return rc;
    if ((rc = rk_args(p, stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef, dt))) return rc;
    const double* Fo[3] = {Uo, Vo, Wo};
    const double* const* R[3] = {Fu, Fv, Fw};
    HIPCHK(p, launch_rk_update(p, stage, U, V, W, Fo, R, coef, dt));
    return finish(p, flags);
}

int dfp_apply_bc(dfp_plan* p, double* U, double* V, double* W, const double* alpha_x, const double* alpha_y,
                 const double* alpha_z, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = step_checks(p, flags & ~DFP_SAVE_P);
    if (rc) return rc;
    if (!U || !V || !W) return fail(p, DFP_E_ARG, "U, V and W are required");
    if (p->geo.ybc == 1 && (!alpha_x || !alpha_y || !alpha_z))
        return fail(p, DFP_E_ARG, "alpha_x, alpha_y and alpha_z are required with y walls");
    HIPCHK(p, launch_apply_bc(p, U, V, W, alpha_x, alpha_y, alpha_z));
    return finish(p, flags);
}

int dfp_substep(dfp_plan* p, int stage, double* U, double* V, double* W, const double* Uo, const double* Vo,
                const double* Wo, const double* const* Fu, const double* const* Fv, const double* const* Fw,
                const double* coef, const double* alpha_x, const double* alpha_y, const double* alpha_z,
                double* rhs_p, double* P, double dt, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = step_checks(p, flags);
    if (rc) return rc;
    if ((rc = rk_args(p, stage, U, V, W, Uo, Vo, Wo, Fu, Fv, Fw, coef, dt))) return rc;
    if (!rhs_p) return fail(p, DFP_E_ARG, "rhs_p is required");
    if (p->geo.ybc == 1 && (!alpha_x || !alpha_y || !alpha_z))
        return fail(p, DFP_E_ARG, "alpha_x, alpha_y and alpha_z are required with y walls");
    const int save_p = (flags & DFP_SAVE_P) ? 1 : 0;
    const double c = coef[stage - 1];
    if (save_p && (!P || !(dt * c != 0.0) || !std::isfinite(dt * c)))
        return fail(p, DFP_E_ARG, "DFP_SAVE_P needs P and a finite nonzero dt*rk_coef(stage,stage)");
    const double* Fo[3] = {Uo, Vo, Wo};
    const double* const* R[3] = {Fu, Fv, Fw};
    // one substage of time_integration.f90 without immersed boundary, VOF or
    // scalars: the update :256-258 (:394-396, :526-531), apply_boundary_conditions
    // :297 (:428, :564), the projection :300 (:431, :567) and the boundary
    // conditions after it :304 (:435, :571), boundary_conditions.f90:186-260
    HIPCHK(p, launch_rk_update(p, stage, U, V, W, Fo, R, coef, dt));
    HIPCHK(p, launch_apply_bc(p, U, V, W, alpha_x, alpha_y, alpha_z));
    if ((rc = spectral_solve(p, U, V, W, rhs_p, P, dt * c, save_p, 1))) return rc;
    HIPCHK(p, launch_correct(p, U, V, W, rhs_p, nullptr));
    HIPCHK(p, launch_apply_bc(p, U, V, W, alpha_x, alpha_y, alpha_z));
    return finish(p, flags);
}

int dfp_divergence(dfp_plan* p, const double* U, const double* V, const double* W, double* rhs_p, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = check_flags(p, flags);
    if (rc) return rc;
    if (!U || !V || !W || !rhs_p) return fail(p, DFP_E_ARG, "U, V, W and rhs_p are required");
    if (flags & DFP_FIELDS_ON_DEVICE) {
        HIPCHK(p, launch_divergence(p, U, V, W, rhs_p));
        return finish(p, flags);
    }
    if ((rc = ensure_staging(p))) return rc;
    if ((rc = h2d(p, p->d_U, U, p->sz_u)) || (rc = h2d(p, p->d_V, V, p->sz_v)) || (rc = h2d(p, p->d_W, W, p->sz_w)) ||
        (rc = h2d(p, p->d_rhs, rhs_p, p->sz_rhs)))
        return rc;
    HIPCHK(p, launch_divergence(p, p->d_U, p->d_V, p->d_W, p->d_rhs));
    if ((rc = d2h(p, rhs_p, p->d_rhs, p->sz_rhs))) return rc;
    return finish(p, flags);
}

int dfp_solve(dfp_plan* p, double* rhs_p, double* P, double dt, double c, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = check_flags(p, flags);
    if (rc) return rc;
    const int save_p = (flags & DFP_SAVE_P) ? 1 : 0;
    if (!rhs_p) return fail(p, DFP_E_ARG, "rhs_p is required");
    if (save_p && (!P || !(dt * c != 0.0) || !std::isfinite(dt * c)))
        return fail(p, DFP_E_ARG, "DFP_SAVE_P needs P and a finite nonzero dt*c");
    if (flags & DFP_FIELDS_ON_DEVICE) {
        rc = spectral_solve(p, nullptr, nullptr, nullptr, rhs_p, P, dt * c, save_p, 0);
        if (rc) return rc;
        return finish(p, flags);
    }
    if ((rc = ensure_staging(p))) return rc;
    if ((rc = h2d(p, p->d_rhs, rhs_p, p->sz_rhs))) return rc;
    rc = spectral_solve(p, nullptr, nullptr, nullptr, p->d_rhs, p->d_P, dt * c, save_p, 0);
    if (rc) return rc;
    if ((rc = d2h(p, rhs_p, p->d_rhs, p->sz_rhs))) return rc;
    if (save_p && (rc = d2h(p, P, p->d_P, p->sz_p))) return rc;
    return finish(p, flags);
}

int dfp_correct(dfp_plan* p, double* U, double* V, double* W, const double* rhs_p, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = check_flags(p, flags);
    if (rc) return rc;
    if (!U || !V || !W || !rhs_p) return fail(p, DFP_E_ARG, "U, V, W and rhs_p are required");
    const bool bf = backflow(p);
    if (flags & DFP_FIELDS_ON_DEVICE) {
        HIPCHK(p, launch_correct(p, U, V, W, rhs_p, bf ? p->phib : nullptr));
        return finish(p, flags);
    }
    if ((rc = ensure_staging(p))) return rc;
    if (bf && (rc = ensure_outflow_staging(p))) return rc;
    if ((rc = h2d(p, p->d_U, U, p->sz_u)) || (rc = h2d(p, p->d_V, V, p->sz_v)) || (rc = h2d(p, p->d_W, W, p->sz_w)) ||
        (rc = h2d(p, p->d_rhs, rhs_p, p->sz_rhs)))
        return rc;
    if (bf && (rc = h2d(p, p->d_phib, p->phib, face_size(p)))) return rc;
    HIPCHK(p, launch_correct(p, p->d_U, p->d_V, p->d_W, p->d_rhs, bf ? p->d_phib : nullptr));
    if ((rc = d2h(p, U, p->d_U, p->sz_u)) || (rc = d2h(p, V, p->d_V, p->sz_v)) || (rc = d2h(p, W, p->d_W, p->sz_w)))
        return rc;
    return finish(p, flags);
}

int dfp_max_divergence(dfp_plan* p, const double* U, const double* V, const double* W, double* out, int flags) {
    if (!p) return DFP_E_ARG;
    int rc = check_flags(p, flags & ~DFP_ASYNC);
    if (rc) return rc;
    if (!U || !V || !W || !out) return fail(p, DFP_E_ARG, "U, V, W and out are required");
    const double *dU = U, *dV = V, *dW = W;
    if (!(flags & DFP_FIELDS_ON_DEVICE)) {
        if ((rc = ensure_staging(p))) return rc;
        if ((rc = h2d(p, p->d_U, U, p->sz_u)) || (rc = h2d(p, p->d_V, V, p->sz_v)) || (rc = h2d(p, p->d_W, W, p->sz_w)))
            return rc;
        dU = p->d_U; dV = p->d_V; dW = p->d_W;
    }
    double* d_out = p->d_red + p->red_len;
    HIPCHK(p, launch_max_div(p, dU, dV, dW, d_out));
    HIPCHK(p, hipMemcpyAsync(out, d_out, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(p, hipStreamSynchronize(p->stream));
    return DFP_OK;
}

int dfp_synth_fields(dfp_plan* p, double* U, double* V, double* W, uint64_t seed, double sigma) {
    return dfp_synth_fields_ex(p, U, V, W, seed, sigma, 0);
}

int dfp_synth_fields_ex(dfp_plan* p, double* U, double* V, double* W, uint64_t seed, double sigma, int flags) {
    if (!p || !U || !V || !W) return DFP_E_ARG;
    if (flags & ~DFP_ASYNC) return fail(p, DFP_E_ARG, "unknown flag bits 0x%x", flags);
    HIPCHK(p, hipSetDevice(p->device));
    HIPCHK(p, launch_synth(p, U, V, W, seed, sigma));
    if (!(flags & DFP_ASYNC)) HIPCHK(p, hipStreamSynchronize(p->stream));
    return DFP_OK;
}

int dfp_profile_enable(dfp_plan* p, int enable) {
    if (!p) return DFP_E_ARG;
    if (!enable) {
        int rc = prof_collect(p);
        if (rc) return rc;
    }
    p->prof = enable != 0;
    return DFP_OK;
}

int dfp_profile_reset(dfp_plan* p) {
    if (!p) return DFP_E_ARG;
    int rc = prof_collect(p);
    if (rc) return rc;
    p->slots.clear();
    return DFP_OK;
}

int dfp_profile_read(dfp_plan* p, int index, const char** name, int64_t* launches, double* ms) {
    if (!p) return DFP_E_ARG;
    int rc = prof_collect(p);
    if (rc) return rc;
    if (index < 0 || index >= (int)p->slots.size()) return DFP_E_ARG;
    if (name) *name = p->slots[index].name.c_str();
    if (launches) *launches = p->slots[index].launches;
    if (ms) *ms = p->slots[index].ms;
    return DFP_OK;
}

const char* dfp_strerror(int code) {
    switch (code) {
        case DFP_OK: return "success";
        case DFP_E_ARG: return "invalid argument";
        case DFP_E_UNSUPPORTED: return "unsupported configuration";
        case DFP_E_HIP: return "HIP runtime error";
        case DFP_E_NOMEM: return "device allocation failed";
        case DFP_E_NONFINITE: return "non-finite value";
        default: return "unknown error";
    }
}

const char* dfp_last_error(const dfp_plan* p) { return p ? p->last_error.c_str() : ""; }

}  // extern "C"
