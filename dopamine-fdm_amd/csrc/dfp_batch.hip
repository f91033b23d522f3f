// Batched projections: many plans (any sizes, either precision) projected
// together with one persistent launch per stage.
//
// Every stage kernel has a batched form (kb_*) whose workgroups walk the
// concatenated work lists of all cases of the stage: the case records
// (XPar/ZPar/CPar/MPar/SPar) and the per-case work offsets are written on the
// host into one pinned staging buffer, copied to the device with one
// asynchronous copy, and read by the kernels.  The kernels' own arguments are
// then only addresses inside that device buffer, the case counts live in it
// too, and the grid is fixed (workgroups per CU x CUs), so a launch sequence
// captured once in a hipGraph replays for any batch with the same stage list
// and LDS sizes.  Each plan keeps its own spectrum scratch, so the cases of a
// batch are independent; a plan may appear once per batch.
//
// Stage list of one call (FP64 cases a, FP32+refine cases b, the latter
// ordered by descending sweep count so sweep s runs the first n_s of them):
//   x forward  (a: divergence, b: divergence + FP64 copy + row sums)
//   (0,0) mode (b)
//   z forward, z backward (a, b)
//   x inverse  (a: phi + P, b: phi + pbar + first-row sums)
//   per sweep s: x forward residual, (0,0) mode, z fwd, z bwd, x inverse ACCUM (b[0:n_s])
//   save P (b with DFP_SAVE_P)
//   correction (a and b)
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <tuple>

#include "dfp_internal.h"

using namespace dfp;

namespace {

constexpr int kMaxSweepStages = 9;

// offsets (in bytes) of the tables inside the staging buffer
struct Layout {
    size_t x64, z64, x32, x32s, z32, m32, s32, c, counts, off;
    size_t total;
};

// per-stage work offsets, each an (n+1)-array of ints, and the case counts
enum OffId {
    O_X64 = 0,   // x kernels of the FP64 cases (forward and inverse)
    O_Z64,       // z kernels of the FP64 cases
    O_X32,       // x kernels of the FP32 cases
    O_Z32,       // z kernels of the FP32 cases
    O_S32,       // save P of the FP32 cases with DFP_SAVE_P
    O_C,         // correction of every case (FP64 first, then FP32)
    O_NUM
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

Layout layout_for(int cap) {
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    L.x64 = take(sizeof(XPar<double2>) * cap);
    L.z64 = take(sizeof(ZPar<double2>) * cap);
    L.x32 = take(sizeof(XPar<float2>) * cap);
    L.x32s = take(sizeof(XPar<float2>) * cap);
    L.z32 = take(sizeof(ZPar<float2>) * cap);
    L.m32 = take(sizeof(MPar) * cap);
    L.s32 = take(sizeof(SPar) * cap);
    L.c = take(sizeof(CPar) * cap);
    // counts: one int per (stage offset array, prefix length) used by a launch
    L.counts = take(sizeof(int) * 64);
    L.off = take(sizeof(int) * (size_t)O_NUM * (cap + 1));
    L.total = o;
    return L;
}

}  // namespace

struct dfp_batch {
    int device = 0;
    int cap = 0;
    int n_cu = 0;
    int occ_x = 4, occ_z = 4, occ_s = 8;  // workgroups per CU of the persistent launches
    hipStream_t stream = nullptr;
    Layout L{};
    unsigned char* h_buf = nullptr;  // pinned staging
    unsigned char* d_buf = nullptr;
    hipEvent_t copied = nullptr;
    bool copy_pending = false;
    bool use_graphs = false;
    std::map<std::vector<long>, hipGraphExec_t> graphs;
    std::string last_error;
};

namespace {

int bfail(dfp_batch* b, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int bfail(dfp_batch* b, int code, const char* fmt, ...) {
    if (b) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        b->last_error = buf;
    }
    return code;
}

#define BCHK(b, expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return bfail((b), DFP_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

// One launch of the batch: which kernel, on which tables, with which count slot.
struct Launch {
    enum Kind { XF, XI, ZF, ZB, M00, SAVE, CORR } kind;
    int f32;       // spectrum precision
    int mode;      // x forward mode, or ACCUM for the inverse
    size_t tab;    // offset of the case records
    int off_id;    // work offsets
    int count_at;  // index of the case count in the counts table
    int ncases;    // host-side count (for LDS sizing and mode00's grid)
    size_t lds;
};

void destroy_graphs(dfp_batch* b) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

// LDS of a batched launch rounded up to a power of two of at least 4 KiB
// (few distinct graphs), unless that would leave no room for the kernels'
// static LDS under the device limit
size_t lds_bucket(size_t v, size_t limit) {
    size_t a = 4096;
    while (a < v) a *= 2;
    return a + 512 <= limit ? a : v;
}

}  // namespace

extern "C" {

int dfp_batch_create(int device, int max_cases, dfp_batch** out) {
    if (!out) return DFP_E_ARG;
    *out = nullptr;
    if (max_cases < 1 || max_cases > 4096) return DFP_E_ARG;
    dfp_batch* b = new (std::nothrow) dfp_batch();
    if (!b) return DFP_E_NOMEM;
    b->device = device;
    b->cap = max_cases;
    b->L = layout_for(max_cases);
    auto bail = [&](int code) {
        (void)hipHostFree(b->h_buf);
        (void)hipFree(b->d_buf);
        if (b->copied) (void)hipEventDestroy(b->copied);
        delete b;
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(DFP_E_HIP);
    if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || b->n_cu < 1)
        return bail(DFP_E_HIP);
    if (hipHostMalloc((void**)&b->h_buf, b->L.total, hipHostMallocDefault) != hipSuccess) return bail(DFP_E_NOMEM);
    if (hipMalloc(&b->d_buf, b->L.total) != hipSuccess) return bail(DFP_E_NOMEM);
    if (hipEventCreateWithFlags(&b->copied, hipEventDisableTiming) != hipSuccess) return bail(DFP_E_HIP);
    memset(b->h_buf, 0, b->L.total);
    *out = b;
    return DFP_OK;
}

int dfp_batch_destroy(dfp_batch* b) {
    if (!b) return DFP_E_ARG;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    destroy_graphs(b);
    (void)hipHostFree(b->h_buf);
    (void)hipFree(b->d_buf);
    (void)hipEventDestroy(b->copied);
    delete b;
    return DFP_OK;
}

int dfp_batch_set_stream(dfp_batch* b, void* s) {
    if (!b) return DFP_E_ARG;
    if (reinterpret_cast<hipStream_t>(s) != b->stream) destroy_graphs(b);
    b->stream = reinterpret_cast<hipStream_t>(s);
    return DFP_OK;
}

int dfp_batch_set_occupancy(dfp_batch* b, int x_per_cu, int z_per_cu, int s_per_cu) {
    if (!b) return DFP_E_ARG;
    if (x_per_cu < 1 || x_per_cu > 16 || z_per_cu < 1 || z_per_cu > 16 || s_per_cu < 1 || s_per_cu > 32)
        return bfail(b, DFP_E_ARG, "workgroups per CU out of range");
    b->occ_x = x_per_cu;
    b->occ_z = z_per_cu;
    b->occ_s = s_per_cu;
    destroy_graphs(b);
    return DFP_OK;
}

int dfp_batch_use_graphs(dfp_batch* b, int enable) {
    if (!b) return DFP_E_ARG;
    b->use_graphs = enable != 0;
    if (!enable) destroy_graphs(b);
    return DFP_OK;
}

const char* dfp_batch_last_error(const dfp_batch* b) { return b ? b->last_error.c_str() : ""; }

int dfp_batch_project(dfp_batch* b, int n, dfp_plan* const* plans, double* const* U, double* const* V,
                      double* const* W, double* const* rhs_p, double* const* P, const double* dt, const double* c,
                      const int* flags, int batch_flags) {
    if (!b) return DFP_E_ARG;
    if (n < 0 || n > b->cap) return bfail(b, DFP_E_ARG, "batch of %d cases (capacity %d)", n, b->cap);
    if (batch_flags & ~DFP_ASYNC) return bfail(b, DFP_E_ARG, "unknown batch flag bits 0x%x", batch_flags);
    if (n == 0) return DFP_OK;
    if (!plans || !U || !V || !W || !rhs_p || !dt || !c) return bfail(b, DFP_E_ARG, "null argument array");
    BCHK(b, hipSetDevice(b->device));
    // validate and split by precision (FP32 cases by descending sweeps)
    std::vector<int> i64, i32;
    for (int i = 0; i < n; ++i) {
        dfp_plan* p = plans[i];
        if (!p) return bfail(b, DFP_E_ARG, "case %d: null plan", i);
        if (p->device != b->device) return bfail(b, DFP_E_ARG, "case %d: plan on device %d", i, p->device);
        for (int k = 0; k < i; ++k)
            if (plans[k] == p) return bfail(b, DFP_E_ARG, "cases %d and %d share a plan", k, i);
        if (!U[i] || !V[i] || !W[i] || !rhs_p[i]) return bfail(b, DFP_E_ARG, "case %d: null field", i);
        const int fl = flags ? flags[i] : 0;
        if (fl & ~(DFP_SAVE_P | DFP_FIELDS_ON_DEVICE))
            return bfail(b, DFP_E_ARG, "case %d: flags 0x%x (device fields, DFP_SAVE_P only)", i, fl);
        if ((fl & DFP_SAVE_P) && (!P || !P[i] || !(dt[i] * c[i] != 0.0)))
            return bfail(b, DFP_E_ARG, "case %d: DFP_SAVE_P needs P and a nonzero dt*c", i);
        (p->precision == DFP_FP32_REFINE ? i32 : i64).push_back(i);
    }
    std::stable_sort(i32.begin(), i32.end(), [&](int a, int q) { return plans[a]->sweeps > plans[q]->sweeps; });
    const int n64 = (int)i64.size(), n32 = (int)i32.size();
    const int max_sweeps = n32 ? std::min(plans[i32[0]]->sweeps, kMaxSweepStages) : 0;

    // wait until the previous call's copy has left the staging buffer
    if (b->copy_pending) {
        BCHK(b, hipEventSynchronize(b->copied));
        b->copy_pending = false;
    }
    unsigned char* H = b->h_buf;
    const Layout& L = b->L;
    auto* x64 = reinterpret_cast<XPar<double2>*>(H + L.x64);
    auto* z64 = reinterpret_cast<ZPar<double2>*>(H + L.z64);
    auto* x32 = reinterpret_cast<XPar<float2>*>(H + L.x32);
    auto* x32s = reinterpret_cast<XPar<float2>*>(H + L.x32s);
    auto* z32 = reinterpret_cast<ZPar<float2>*>(H + L.z32);
    auto* m32 = reinterpret_cast<MPar*>(H + L.m32);
    auto* s32 = reinterpret_cast<SPar*>(H + L.s32);
    auto* cc = reinterpret_cast<CPar*>(H + L.c);
    int* counts = reinterpret_cast<int*>(H + L.counts);
    int* off = reinterpret_cast<int*>(H + L.off);
    auto offp = [&](int id) { return off + (size_t)id * (b->cap + 1); };
    for (int id = 0; id < O_NUM; ++id) offp(id)[0] = 0;

    size_t lds_xf64 = 0, lds_xi64 = 0, lds_z64 = 0, lds_xf32 = 0, lds_xi32 = 0, lds_z32 = 0;
    for (int a = 0; a < n64; ++a) {
        const int i = i64[a];
        dfp_plan* p = plans[i];
        const int fl = flags ? flags[i] : 0;
        XPar<double2> q = make_xpar<double2>(p);
        q.a.U = U[i]; q.a.V = V[i]; q.a.W = W[i];
        q.rhs = rhs_p[i];
        q.save_p = (fl & DFP_SAVE_P) ? 1 : 0;
        q.P = q.save_p ? P[i] : nullptr;
        q.dtc = q.save_p ? dt[i] * c[i] : 1.0;
        x64[a] = q;
        z64[a] = make_zpar<double2>(p, 0);
        offp(O_X64)[a + 1] = offp(O_X64)[a] + q.nbx * p->geo.nym;
        offp(O_Z64)[a + 1] = offp(O_Z64)[a] + z64[a].nblocks;
        lds_xf64 = std::max(lds_xf64, x_lds_bytes(p, 0, 0));
        lds_xi64 = std::max(lds_xi64, x_lds_bytes(p, 1, 0));
        lds_z64 = std::max(lds_z64, z_lds_bytes(p, 0));
    }
    int ns = 0;
    std::vector<int> n_at_sweep(kMaxSweepStages + 1, 0);  // FP32 cases with sweeps > s
    for (int a = 0; a < n32; ++a) {
        const int i = i32[a];
        dfp_plan* p = plans[i];
        const int fl = flags ? flags[i] : 0;
        XPar<float2> q = make_xpar<float2>(p);
        q.a.U = U[i]; q.a.V = V[i]; q.a.W = W[i];
        q.a.b_out = p->d_b;
        q.a.rowpart = p->d_rowpart;
        q.rhs = rhs_p[i];
        q.pbar = p->d_pbar;
        q.part0 = p->d_part0;
        x32[a] = q;
        XPar<float2> r = q;  // sweeps: residual forward, accumulating inverse
        r.a = XArgs{};
        r.a.rhs = rhs_p[i];
        r.a.b = p->d_b;
        r.a.rowpart = p->d_rowpart;
        r.a.part0 = p->d_part0;
        r.a.idx2 = 1.0 / (p->dx * p->dx);
        r.a.idz2 = 1.0 / (p->dz * p->dz);
        x32s[a] = r;
        z32[a] = make_zpar<float2>(p, 1);
        m32[a] = make_mpar(p);
        offp(O_X32)[a + 1] = offp(O_X32)[a] + q.nbx * p->geo.nym;
        offp(O_Z32)[a + 1] = offp(O_Z32)[a] + z32[a].nblocks;
        for (int s = 0; s < std::min(p->sweeps, kMaxSweepStages); ++s) n_at_sweep[s] = a + 1;
        if (fl & DFP_SAVE_P) {
            s32[ns] = make_spar(p, rhs_p[i], P[i], dt[i] * c[i]);
            offp(O_S32)[ns + 1] = offp(O_S32)[ns] + s32[ns].nblocks;
            ++ns;
        }
        lds_xf32 = std::max(lds_xf32, x_lds_bytes(p, 0, 1));
        lds_xi32 = std::max(lds_xi32, x_lds_bytes(p, 1, 1));
        lds_z32 = std::max(lds_z32, z_lds_bytes(p, 1));
    }
    for (int a = 0; a < n; ++a) {
        const int i = a < n64 ? i64[a] : i32[a - n64];
        dfp_plan* p = plans[i];
        cc[a] = make_cpar(p, U[i], V[i], W[i], rhs_p[i]);
        offp(O_C)[a + 1] = offp(O_C)[a] + cc[a].nbx * correct_blocks_z(p->geo);
    }

    // launch list; counts[k] is the case count a launch reads
    std::vector<Launch> ls;
    int nct = 0;
    const size_t lds_limit = plans[0]->lds_limit;
    auto add = [&](Launch::Kind k, int f32, int mode, size_t tab, int off_id, int ncase, size_t lds) {
        if (nct < 64) counts[nct] = ncase;
        ls.push_back(Launch{k, f32, mode, tab, off_id, nct, ncase, lds_bucket(lds, lds_limit)});
        ++nct;
    };
    if (n64) {
        add(Launch::XF, 0, 0, L.x64, O_X64, n64, lds_xf64);
        add(Launch::ZF, 0, 0, L.z64, O_Z64, n64, lds_z64);
        add(Launch::ZB, 0, 0, L.z64, O_Z64, n64, lds_z64);
        add(Launch::XI, 0, 0, L.x64, O_X64, n64, lds_xi64);
    }
    if (n32) {
        add(Launch::XF, 1, 0, L.x32, O_X32, n32, lds_xf32);
        add(Launch::M00, 1, 0, L.m32, O_X32, n32, 0);
        add(Launch::ZF, 1, 0, L.z32, O_Z32, n32, lds_z32);
        add(Launch::ZB, 1, 0, L.z32, O_Z32, n32, lds_z32);
        add(Launch::XI, 1, 0, L.x32, O_X32, n32, lds_xi32);
        for (int s = 0; s < max_sweeps; ++s) {
            const int m = n_at_sweep[s];
            add(Launch::XF, 1, 2, L.x32s, O_X32, m, lds_xf32);
            add(Launch::M00, 1, 0, L.m32, O_X32, m, 0);
            add(Launch::ZF, 1, 0, L.z32, O_Z32, m, lds_z32);
            add(Launch::ZB, 1, 0, L.z32, O_Z32, m, lds_z32);
            add(Launch::XI, 1, 1, L.x32s, O_X32, m, lds_xi32);
        }
        if (ns) add(Launch::SAVE, 1, 0, L.s32, O_S32, ns, 0);
    }
    add(Launch::CORR, 0, 0, L.c, O_C, n, 0);
    if (nct > 64) return bfail(b, DFP_E_UNSUPPORTED, "too many stages (%d)", nct);

    // tables to the device, then the launches (or their graph)
    BCHK(b, hipMemcpyAsync(b->d_buf, b->h_buf, L.total, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipEventRecord(b->copied, b->stream));
    b->copy_pending = true;

    unsigned char* D = b->d_buf;
    const int* d_counts = reinterpret_cast<const int*>(D + L.counts);
    const int* d_off = reinterpret_cast<const int*>(D + L.off);
    auto d_offp = [&](int id) { return d_off + (size_t)id * (b->cap + 1); };
    const int gx = b->n_cu * b->occ_x, gz = b->n_cu * b->occ_z, gs = b->n_cu * b->occ_s;
    auto run_all = [&](hipStream_t st) -> hipError_t {
        for (const Launch& l : ls) {
            const int* o = d_offp(l.off_id);
            const int* nc = d_counts + l.count_at;
            hipError_t e = hipSuccess;
            switch (l.kind) {
                case Launch::XF:
                    e = l.f32 ? launch_xfwd_batch(st, l.mode, reinterpret_cast<const XPar<float2>*>(D + l.tab), o, nc,
                                                  gx, l.lds)
                              : launch_xfwd_batch(st, l.mode, reinterpret_cast<const XPar<double2>*>(D + l.tab), o,
                                                  nc, gx, l.lds);
                    break;
                case Launch::XI:
                    e = l.f32 ? launch_xinv_batch(st, l.mode, reinterpret_cast<const XPar<float2>*>(D + l.tab), o, nc,
                                                  gx, l.lds)
                              : launch_xinv_batch(st, l.mode, reinterpret_cast<const XPar<double2>*>(D + l.tab), o,
                                                  nc, gx, l.lds);
                    break;
                case Launch::ZF:
                case Launch::ZB: {
                    const int dir = l.kind == Launch::ZF ? -1 : 1;
                    e = l.f32 ? launch_zsolve_batch(st, dir, reinterpret_cast<const ZPar<float2>*>(D + l.tab), o, nc,
                                                    gz, l.lds)
                              : launch_zsolve_batch(st, dir, reinterpret_cast<const ZPar<double2>*>(D + l.tab), o, nc,
                                                    gz, l.lds);
                    break;
                }
                case Launch::M00:
                    e = launch_mode00_batch(st, reinterpret_cast<const MPar*>(D + l.tab), nc, std::min(b->cap, 256));
                    break;
                case Launch::SAVE:
                    e = launch_save_p_batch(st, reinterpret_cast<const SPar*>(D + l.tab), o, nc, gs);
                    break;
                case Launch::CORR:
                    e = launch_correct_batch(st, reinterpret_cast<const CPar*>(D + l.tab), o, nc, gs);
                    break;
            }
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };

    if (!b->use_graphs) {
        BCHK(b, run_all(b->stream));
    } else {
        // key: the launch list as the kernels see it (kind, precision, mode, table, counts slot, LDS)
        std::vector<long> key;
        for (const Launch& l : ls) {
            key.push_back(l.kind);
            key.push_back(l.f32 * 16 + l.mode);
            key.push_back((long)l.tab);
            key.push_back(l.off_id * 64 + l.count_at);
            key.push_back((long)l.lds);
        }
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipStream_t cap_stream;
            BCHK(b, hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
            hipGraph_t graph;
            hipError_t e = hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t el = run_all(cap_stream);
                e = hipStreamEndCapture(cap_stream, &graph);
                if (el != hipSuccess) e = el;
            }
            hipGraphExec_t exec = nullptr;
            if (e == hipSuccess) {
                e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
            }
            (void)hipStreamDestroy(cap_stream);
            if (e != hipSuccess) return bfail(b, DFP_E_HIP, "graph capture: %s", hipGetErrorString(e));
            it = b->graphs.emplace(key, exec).first;
        }
        BCHK(b, hipGraphLaunch(it->second, b->stream));
    }
    BCHK(b, hipGetLastError());
    if (!(batch_flags & DFP_ASYNC)) BCHK(b, hipStreamSynchronize(b->stream));
    return DFP_OK;
}

}  // extern "C"
