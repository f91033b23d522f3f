// Internal declarations shared by the dfp HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dfp.h"

namespace dfp {

constexpr int kMaxStages = 16;

// Division by a runtime constant d >= 1 for 0 <= n < 2^31:
//   q = (umulhi(n, mul) + n) >> shift, shift = ceil(log2 d),
//   mul = floor(2^32 (2^shift - d) / d) + 1  (round-up multiply method).
struct FastDiv {
    uint32_t d = 1, mul = 1, shift = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t dd) : d(dd) {
        shift = 0;
        while ((1ull << shift) < dd) ++shift;
        mul = (uint32_t)(((1ull << 32) * ((1ull << shift) - dd)) / dd + 1);
    }
    __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
        return (__umulhi(n, mul) + n) >> shift;
#else
        return (uint32_t)((((uint64_t)n * mul) >> 32) + n) >> shift;
#endif
    }
};

// Stockham plan of one FFT length n = 2^a 3^b on the device.
// Stage st has radix rad[st]; its twiddles exp(-2 pi i p u / ncur) for
// p = 0..m-1, u = 1..r-1 sit at tw[twoff[st] + p*(r-1) + (u-1)], where
// ncur = n / (rad[0]*...*rad[st-1]) and m = ncur / r.
struct FftDesc {
    int n;
    int nst;
    int rad[kMaxStages];
    int twoff[kMaxStages];
    FastDiv nbdiv[kMaxStages];  // butterflies per line of stage st: (n / rad[st])
    FastDiv sdiv[kMaxStages];   // stride of stage st: product of the earlier radices
};

// Geometry and 1-D metric tables of one case, as the kernels read it.
// Index conventions (reference indices are 1-based):
//   inv_dyf[j-2] = 1/(y(j)-y(j-1)),    j = 2..nyg-1   (nym entries)
//   inv_dzf[k-2] = 1/(z(k)-z(k-1)),    k = 2..nzg-1   (nzp+1 entries)
//   inv_dyg[j-2] = 1/(yg(j+1)-yg(j)),  j = 2..ny-1    (nym-1 entries)
//   inv_dzg[k-2] = 1/(zg(k+1)-zg(k)),  k = 2..nz-1    (nzp entries)
struct Geo {
    int nxp, nym, nzp, nxh;   // nxh: x planes of the spectrum (nxp/2+1 periodic, nxp inflow/outflow)
    int xbc;                  // x_bc_type: 0 periodic, 1 inflow/outflow (DCT-IV in x)
    int ybc;                  // y_bc_type: 1 walls (tridiagonal y solve), 0 periodic (y FFT, uniform y)
    int zbc;                  // z_bc_type: 0 periodic (FFT in z), 1 walls (the duct: z eigenbasis, nzp = nzm)
    int nx, ny, nz, nxg, nyg, nzg;
    double inv_dx;
    const double* inv_dyf;
    const double* inv_dzf;
    const double* inv_dyg;
    const double* inv_dzg;
    const double* x;   // x(1:nx)
    const double* y;   // y(1:ny)
    const double* z;   // z(1:nz)
    const double* yg;  // yg(1:nyg)
};

// y-operator of the spectral solve: Dyy(j+1,j), Dyy(j,j), Dyy(j,j+1) and the
// modified wavenumbers of the kept modes.
struct Ops {
    const double* dl;   // nym-1
    const double* dd;   // nym
    const double* du;   // nym-1
    const double* kxx;  // nxh: kx = 0..nxp/2 (periodic) or the nxp DCT-IV modes
    const double* kzz;  // nzp
    const double* kyy;  // nym-1 (periodic y), else null
};

// Inputs of the forward x-transform kernel beyond the spectrum.
//   mode 0: U, V, W (divergence); mode 1: rhs; mode 2: rhs = phi and b (the
//   residual b - L_pin phi of the FP32 form).
struct XArgs {
    const double* U;
    const double* V;
    const double* W;
    const double* rhs;
    const double* b;
    double* b_out;          // modes 0, 1: FP64 copy of the field (rhs_p layout), or null
    double* rowpart;        // per-(row, block) sums of the field, or null
    const double* part0;    // mode 2: per-block sums of phi's first row
    double idx2, idz2;      // mode 2: 1/dx^2, 1/dz^2
    const double* phib;     // mode 0, inflow/outflow x: outlet face values phib(2:nyg-1,2:nzg-1)
                            // (the last cell's equation takes -2 phib/dx^2), or null
};

// Per-case parameters of the x-direction kernels (k_spectral.hip).  C is
// double2 or float2 (the spectrum's precision).
template <typename C>
struct XPar {
    Geo g;
    Ops o;
    FftDesc fd;          // x FFT of length nxp/2
    const C* tw;         // x stages | R2C packing (nxp/2+1) | first z stage (nz2*(R-1))
    C* spec;
    int ntw, R, T, nbx;  // stage-twiddle count, z split radix, columns per block, blocks per row
    int strip;           // column blocks per strip of the workgroup order (nbx: row-major)
    int npart0;          // blocks per row of the inverse kernel: entries of part0
    FastDiv div_nxp, div_t, div_half;
    XArgs a;             // forward kernel inputs
    double* rhs;         // inverse: phi out (and in, when accumulating)
    double* P;
    double scale, dtc;
    int save_p;
    const double* pbar;  // inverse: FP64 (0,0)-mode part added per row, or null
    double* part0;       // inverse: per-block sums of phi's first row, or null
};

// Per-case parameters of the fused z-transform / y-solve kernels (k_zsolve.hip).
template <typename C>
struct ZPar {
    C* spec;
    const double* lu;
    const int* lucol;    // pivot column of sub-plane u, mode m: lucol[u * nzp/R + m]
    const double* dl;
    const double* du;
    FftDesc fz;          // z FFT of length nzp/R
    const C* tw;
    const C* tw2;        // w_n^e, e < n, of the two-pass transforms
    int nym, nzp, R, ntw, G, zero00, nblocks;
    int nplanes;         // (kx, u) sub-planes: nxh * R
    int n1, n2;          // two-pass factorization n = n1*n2 (0: staged transform)
    FastDiv div_n;
};

// Per-case parameters of the periodic-y z/y kernel (k_zfused.hip): the fused
// kernel's record (no y solve tables) plus the y transform over the nym-1
// independent rows and the three modified-wavenumber tables.
template <typename C>
struct ZyPar {
    ZPar<C> z;
    FftDesc fy;          // y FFT of length nym-1
    const C* twy;        // its stage twiddles
    const double* kxx;   // nxh
    const double* kyy;   // nym-1
    const double* kzz;   // nzp
    int ntwy;            // twiddle count of fy
    int gy;              // columns per y-transform call (divides nzp/R)
    FastDiv div_gy;
    double inv_ny;       // 1/(nym-1)
};

// Per-case parameters of the stencil kernels (k_stencil.hip, k_mixed.hip).
struct CPar {
    double* U;
    double* V;
    double* W;
    const double* rhs;
    const double* phib;  // inflow/outflow x with the backflow term: outlet face values, else null
    Geo g;
    FastDiv div_ni;
    int nbx;             // blocks over the (x,y) plane
};
struct MPar {
    const double* rowpart;
    const double* lu;
    Ops o;
    double* x;
    double count;
    int nparts, nym, nzs;
};
struct SPar {
    const double* rhs;
    double* P;
    Geo g;
    double dtc;
    FastDiv div_x, div_y;
    int nblocks;
};

// Launch-time record of one kernel for the profiler.
struct ProfSlot {
    std::string name;
    int64_t launches = 0;
    double ms = 0.0;
};

}  // namespace dfp

struct dfp_plan {
    int device = 0;
    int precision = DFP_FP64;
    dfp::Geo geo{};
    dfp::Ops ops{};
    double dx = 0, dz = 0;
    std::vector<double> hx, hy, hz, hyg, hzg, hkxx, hkzz, hdl, hdd, hdu;
    // device tables
    double* d_tables = nullptr;  // one allocation holding every 1-D table
    double2* d_tw = nullptr;     // FFT twiddles: x stages, z stages, R2C post twiddles
    dfp::FftDesc fx{}, fz{};     // x: length nxp/2 (half-length complex), z: length nzp/zr
    dfp::FftDesc fy{};           // periodic y: length nym-1
    const double2* tw_y = nullptr;  // periodic y: stage twiddles of fy
    int gy = 1;                  // periodic y: columns per y-transform call
    // the duct (z walls): second spectrum (the eigenmode planes), the z
    // eigenbasis tables and the pinned eigenmode (the zero eigenvalue)
    double2* d_spec2 = nullptr;
    double* d_duct = nullptr;    // Qf (nzp x nzp, row k) | Qi (nzp x nzp, row m) | lambda_z (nzp)
    const double* duct_qf = nullptr;
    const double* duct_qi = nullptr;
    const double* duct_lam = nullptr;
    int pin_m = -1;
    // z walls with a periodic y: Dzz's three diagonals, the multipliers DL/D'
    // and inverse pivots of the pinned kx = ky = 0 z system (in d_tables),
    // and that line's eigenbasis image
    const double *dzz_l = nullptr, *dzz_d = nullptr, *dzz_u = nullptr;
    const double *zw_mult = nullptr, *zw_ip = nullptr;
    double2* d_z00 = nullptr;
    int zr = 1;                  // radix of the z stage applied by the x kernels (z split)
    const double2* tw_x = nullptr;
    const double2* tw_z = nullptr;
    const double2* tw_z2 = nullptr;   // exp(-2 pi i e / n), e < n = nzp/zr (two-pass z transforms)
    const double2* tw_r2c = nullptr;  // exp(-2 pi i k / nxp), k = 0..nxp/2
    const double2* tw_zs = nullptr;   // exp(-2 pi i k' u / nzp), [k'*(zr-1) + u-1], k' < nzp/zr, 1 <= u < zr
    float2* d_tw32 = nullptr;         // the same table rounded to FP32 (FP32 form)
    const float2* tw_x32 = nullptr;
    const float2* tw_z32 = nullptr;
    const float2* tw_z2_32 = nullptr;
    // device work arrays
    double2* d_spec = nullptr;   // (nxp/2+1) x nym x nzp complex, layout [kx][u][j][m], kz = zr*m + u
    double* d_lu = nullptr;      // inverse pivots of the y solve, [kx][j][lu_pos(kz)]
    int* d_lucol = nullptr;      // lu_pos of mode (u, m) of the z split: [u * nzp/R + m]
    double* d_red = nullptr;     // reduction partials
    int red_len = 0;
    // FP32 form: the FP64 right-hand side, partial sums and the (0,0) mode
    int sweeps = 1;              // refinement sweeps after the first FP32 solve
    double* d_b = nullptr;       // rhs_p layout
    double* d_rowpart = nullptr; // nym x (x-kernel blocks per row)
    double* d_part0 = nullptr;   // x-kernel blocks per row
    double* d_pbar = nullptr;    // nym
    // x inflow/outflow: outlet backflow state (dfp_set_outflow) and the staging
    // of its face arrays for host-resident calls, (nyg-2) x (nzg-2) each
    double beta = 0.0, outflow_uc = 0.0;
    const double* pb = nullptr;  // caller's Pb(2:nyg-1,2:nzg-1)
    double* phib = nullptr;      // caller's phib(2:nyg-1,2:nzg-1)
    double *d_pb = nullptr, *d_phib = nullptr;
    // staging for host-resident fields
    double *d_U = nullptr, *d_V = nullptr, *d_W = nullptr, *d_rhs = nullptr, *d_P = nullptr;
    int64_t sz_u = 0, sz_v = 0, sz_w = 0, sz_p = 0, sz_rhs = 0;
    hipStream_t stream = nullptr;
    // second stream: the FP32 form's (0,0)-mode solve beside the z solves, or the
    // pinned kx = ky = 0 line of the z walls with a periodic y beside the forward GEMM
    hipStream_t aux = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // profiling
    bool prof = false;
    std::vector<dfp::ProfSlot> slots;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    std::string last_error;
    // launch shapes
    int tk_x = 8;   // z columns per block of the forward x kernel (zr lines each)
    int tk_xi = 8;  // z columns per block of the inverse x kernel
    int g_zs = 4;   // rows per group of the split z kernels, or rows per FFT call of the fused one
    int zfused = 0; // 1: one fused z kernel (whole sub-plane in LDS); 0: the split pair
    int zn1 = 0, zn2 = 0;  // two-pass z transforms of the split pair (0: staged transform)
    size_t lds_limit = 65536;  // LDS bytes per block on this device
};

namespace dfp {

// launchers (defined next to their kernels)
// inflow/outflow x: DCT-IV kernels; phib (mode 0 only) adds the outlet backflow term
hipError_t launch_div_rfft_x(dfp_plan* p, const double* U, const double* V, const double* W,
                             const double* rhs, int from_rhs, const double* phib = nullptr);
hipError_t launch_lu_table(dfp_plan* p);  // pins the (0,0) mode of the periodic x only
// z FFT + forward elimination, then back substitution + inverse z FFT, per
// sub-plane in one workgroup
hipError_t launch_zsolve(dfp_plan* p);
int zsolve_rows_per_group(const FftDesc& fz, int nym, size_t lds_limit);
bool z_two_pass(int n, int& n1, int& n2);
// z kernel variant of a plan: -1 fused, else n1*32 + n2 of the split pair (0: staged transform)
int z_variant(const dfp_plan* p);
hipError_t launch_irfft_x(dfp_plan* p, double* rhs, double* P, double dtc, int save_p);
// fused z transform + y solve (whole sub-plane in LDS)
template <typename C> hipError_t launch_zfused(dfp_plan* p, const ZPar<C>& zp);
template <typename C>
hipError_t launch_zfused_batch(hipStream_t s, const ZPar<C>* cs, const int* off, const int* idx, const int* nc,
                               int grid, size_t lds);
size_t zf_lds_bytes(int nym, int n, int ntw, int f32);
// the duct: z eigenbasis transforms and y solves of the spectrum (k_duct.hip)
hipError_t launch_duct_lu(dfp_plan* p);
hipError_t launch_duct_solve(dfp_plan* p);
size_t yfft_solve_lds_bytes(int nym, int G, int ntwy);
int yfft_solve_columns(const FftDesc& fy, int nym, size_t lds_limit);
// periodic y: z transform + y FFT solve of each (kx, u) sub-plane in LDS (FP64)
hipError_t launch_zfused_py(dfp_plan* p);
size_t zf_py_lds_bytes(int nym, int n, int ntwz, int ntwy);
// columns per y-transform call: the largest divisor of n whose columns one call holds
int zf_py_columns(const FftDesc& fy, int n);
size_t zf_lds_bytes(const dfp_plan* p, int f32);
int zf_rows_per_call(const FftDesc& fz);
// FP32 form
hipError_t launch_rfft_x32(dfp_plan* p, int mode, const XArgs& a);
hipError_t launch_zsolve32(dfp_plan* p);
hipError_t launch_irfft_x32(dfp_plan* p, double* rhs, int accumulate, const double* pbar, double* part0, double* P,
                            double dtc, int save_p);
hipError_t launch_mode00(dfp_plan* p, hipStream_t s);  // rowpart -> pbar of the plan's FP32 buffers
int correct_blocks_z(const Geo& g);
hipError_t launch_divergence(dfp_plan* p, const double* U, const double* V, const double* W, double* rhs);
// phib: outlet face values of the inflow/outflow x with the backflow term, or null
hipError_t launch_correct(dfp_plan* p, double* U, double* V, double* W, const double* rhs,
                          const double* phib = nullptr);
// set_outflow_backflow_pressure (projection.f90:65-96): phib from u* at the
// outlet and Pb, with dtc = dt*rk_coef(rk_step,rk_step); rhs, if given, takes
// -2 phib/dx^2 in its last x cell (the fused x kernel applies it itself)
hipError_t launch_outflow_phib(dfp_plan* p, const double* U, const double* V, const double* W, const double* Pb,
                               double* phib, double* rhs, double dtc);
hipError_t launch_max_div(dfp_plan* p, const double* U, const double* V, const double* W, double* out_dev);
// RK3 stage update of the interiors (Fo: Uo, Vo, Wo; R[field][s]: right-hand sides, s < stage;
// c = rk_coef(stage, 1:stage)) and the channel's apply_boundary_conditions (k_step.hip)
hipError_t launch_rk_update(dfp_plan* p, int stage, double* U, double* V, double* W, const double* const Fo[3],
                            const double* const* R[3], const double* c, double dt);
hipError_t launch_apply_bc(dfp_plan* p, double* U, double* V, double* W, const double* ax, const double* ay,
                           const double* az);
// out_dev = 1 if a velocity value in the monitor's ranges is NaN or above 1e8 in magnitude, else 0
hipError_t launch_guard_finite(dfp_plan* p, const double* U, const double* V, const double* W, double* out_dev);
hipError_t launch_synth(dfp_plan* p, double* U, double* V, double* W, uint64_t seed, double sigma);

// parameter records and batched (persistent) launches: cs (records), off
// (per-listed-case work offsets, off[0] = 0, off[*nc] = total items), idx
// (listed case -> record) and the count *nc are in device memory
template <typename C> XPar<C> make_xpar(const dfp_plan* p, int inv = 0);
// set the column fields (T, nbx, strip, div_t) of q for the forward (inv = 0) or inverse x kernel
template <typename C> void xpar_columns(XPar<C>& q, const dfp_plan* p, int inv);
template <typename C> ZPar<C> make_zpar(const dfp_plan* p, int zero00);
CPar make_cpar(const dfp_plan* p, double* U, double* V, double* W, const double* rhs);
MPar make_mpar(const dfp_plan* p);
SPar make_spar(const dfp_plan* p, const double* rhs, double* P, double dtc);
size_t x_lds_bytes(const dfp_plan* p, int inv, int f32);
size_t z_lds_bytes(const dfp_plan* p, int f32);
template <typename C>
hipError_t launch_xfwd_batch(hipStream_t s, int R, int mode, const XPar<C>* cs, const int* off, const int* idx,
                             const int* nc, int grid, size_t lds);
template <typename C>
hipError_t launch_xinv_batch(hipStream_t s, int R, int accum, const XPar<C>* cs, const int* off, const int* idx,
                             const int* nc, int grid, size_t lds);
template <typename C>
hipError_t launch_zsolve_batch(hipStream_t s, int variant, const ZPar<C>* cs, const int* off, const int* idx,
                               const int* nc, int grid, size_t lds);
hipError_t launch_correct_batch(hipStream_t s, const CPar* cs, const int* off, const int* idx, const int* nc,
                                int grid);
hipError_t launch_mode00_batch(hipStream_t s, const MPar* cs, const int* idx, const int* nc, int grid);
hipError_t launch_save_p_batch(hipStream_t s, const SPar* cs, const int* off, const int* idx, const int* nc,
                               int grid);

// profiler hooks used by the launchers
int prof_begin(dfp_plan* p, const char* name);
void prof_end(dfp_plan* p, int slot);

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// A forward x tile of R*T lines of the x transform fx holds more points than
// two butterfly slots per thread (512 threads, radix-2 stages paired) can
// transform: its kernel takes four slots each (FP32 form only).
bool xfwd_wide(const FftDesc& fx, int R, int T);

// Workgroups of a persistent (batched) launch: the caller's count, capped at
// what the device holds at once of this kernel with this LDS.  A workgroup
// beyond that would start only once another has finished all of its work
// items, and then run its own items with the GPU mostly idle.
int resident_limit(const void* kernel, int threads, size_t lds);  // per device, cached; 0 if unknown
template <typename K>
inline int resident_grid(K kernel, int grid, int threads, size_t lds) {
    const int cap = resident_limit(reinterpret_cast<const void*>(kernel), threads, lds);
    return cap > 0 && cap < grid ? cap : grid;
}

// The pointers in the case records all address global memory.  Saying so
// lets the compiler use global (not flat) memory instructions where a record
// is read from memory rather than from the kernel arguments (the batched
// kernels): flat accesses would also count against the LDS wait counter.
template <typename T>
using Glob = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ Glob<T>* gptr(T* p) {
    return (Glob<T>*)p;
}

// two consecutive doubles at an 8-byte aligned address
typedef double dbl2u __attribute__((ext_vector_type(2), aligned(8)));

__device__ __forceinline__ dbl2u ld2(const Glob<double>* p) { return *reinterpret_cast<const Glob<dbl2u>*>(p); }

// store the elements e with ok_e; both with one 16-byte store
__device__ __forceinline__ void st2(Glob<double>* p, double a, double b, bool ok0, bool ok1) {
    if (ok0 && ok1) {
        dbl2u v;
        v.x = a;
        v.y = b;
        *reinterpret_cast<Glob<dbl2u>*>(p) = v;
    } else {
        if (ok0) p[0] = a;
        if (ok1) p[1] = b;
    }
}

// Column of the pivot table for z mode kz of a plan split into R z sub-planes
// of nz2 = nzp/R modes (kz = R m + u).  The pivots depend on kz only through
// min(kz, nzp - kz), and mode nzp - kz lies in sub-plane R - u (mod R) at
// position nz2 - 1 - m (u > 0) or nz2 - m (u = 0).  Columns are laid out so
// that each sub-plane's modes read one contiguous run: u = 0 first (m and
// nz2 - m share a column), then per pair {u, R - u} the nz2 columns of u in
// m order (R - u reads them in reverse), and u = R/2 pairs m with
// nz2 - 1 - m.  A bijection of the nzp/2 + 1 folded modes onto
// [0, nzp/2 + 1); R = 1 gives column min(kz, nzp - kz).
__host__ __device__ __forceinline__ int lu_pos_mu(int m, int u, int R, int nz2) {
    if (u == 0) return m > 0 ? min(m, nz2 - m) : 0;
    const int c = min(u, R - u);
    const int base = nz2 / 2 + 1 + (c - 1) * nz2;
    if (2 * u == R) return base + min(m, nz2 - 1 - m);
    return base + (u < R - u ? m : nz2 - 1 - m);
}
__host__ __device__ __forceinline__ int lu_pos(int kz, int R, int nz2) {
    const int m = kz / R;
    return lu_pos_mu(m, kz - m * R, R, nz2);
}

// MI355X dispatches workgroup w of a launch to XCD w % 8, and each XCD has its
// own L2.  xcd_remap(w, n) renumbers the n workgroups so that each XCD gets a
// contiguous range of logical indices: neighbouring tiles (which read the
// same halo rows, pivot rows or cache lines) then share an L2.  A bijection
// on [0, n) for every n; only locality, never correctness, depends on the
// dispatch order.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_remap(int w, int n) {
    const int q = n / kXcds, r = n - q * kXcds;
    const int x = w % kXcds, i = w / kXcds;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

}  // namespace dfp
