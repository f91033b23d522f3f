// Batched projections: many plans (any sizes, either precision) projected
// together with one persistent launch per stage.
//
// Every stage kernel has a batched form (kb_*) whose workgroups walk the
// concatenated work lists of all cases of the stage: the case records
// (XPar/ZPar/CPar/MPar/SPar) and the per-case work offsets are written on the
// host into one pinned staging buffer, copied to the device with one
// asynchronous copy, and read by the kernels.  The kernels' own arguments are
// then only addresses inside that device buffer, the case counts live in it
// too, and the grid is fixed (workgroups per CU x CUs, capped at what the
// device holds at once of the kernel with its LDS), so a launch sequence
// captured once in a hipGraph replays for any batch with the same stage list
// and LDS sizes.  Each plan keeps its own spectrum scratch, so the cases of a
// batch are independent; a plan may appear once per batch.
//
// Stage list of one call (FP64 cases a, FP32+refine cases b, the latter
// ordered by descending sweep count so sweep s runs the first n_s of them);
// each x stage is one launch per z-split radix present (the x kernels are
// compiled per radix), each z stage one launch for the cases with the fused
// z kernel and one pair for those with the split pair:
//   x forward  (a: divergence, b: divergence + FP64 copy + row sums)
//   (0,0) mode (b)
//   z transforms and y solves (a, b)
//   x inverse  (a: phi + P, b: phi + pbar + first-row sums)
//   per sweep s: x forward residual, (0,0) mode, z, x inverse ACCUM (b[0:n_s])
//   save P (b with DFP_SAVE_P)
//   correction (a and b)
// A launch reads a case list: the work offsets of the listed cases, their
// record indices and their count, all in the device buffer.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <tuple>

#include <mutex>

#include "dfp_internal.h"

using namespace dfp;

namespace dfp {

int resident_limit(const void* kernel, int threads, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, int, size_t>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    const auto key = std::make_tuple(dev, kernel, threads, lds);
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int per_cu = 0, ncu = 0;
    int cap = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && per_cu > 0 && ncu > 0)
        cap = per_cu * ncu;
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = cap;
    return cap;
}

}  // namespace dfp

namespace {

constexpr int kMaxSweepStages = 9;
constexpr int kRadixList[] = {1, 2, 3, 4, 6, 8, 9, 12, 16};  // z-split radices the x kernels are compiled for

// offsets (in bytes) of the tables inside the staging buffer
struct Layout {
    size_t x64, x64i, z64, x32, x32i, x32s, x32si, z32, m32, s32, c, counts, lists;  // *i: inverse x records
    size_t total;
};

constexpr int kMaxLists = 512;  // case lists per call (each launch reads one)

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
    L.x64i = take(sizeof(XPar<double2>) * cap);
    L.z64 = take(sizeof(ZPar<double2>) * cap);
    L.x32 = take(sizeof(XPar<float2>) * cap);
    L.x32i = take(sizeof(XPar<float2>) * cap);
    L.x32s = take(sizeof(XPar<float2>) * cap);
    L.x32si = take(sizeof(XPar<float2>) * cap);
    L.z32 = take(sizeof(ZPar<float2>) * cap);
    L.m32 = take(sizeof(MPar) * cap);
    L.s32 = take(sizeof(SPar) * cap);
    L.c = take(sizeof(CPar) * cap);
    // list counts, then the lists: offsets (count + 1) and record indices (count)
    L.counts = take(sizeof(int) * kMaxLists);
    L.lists = take(sizeof(int) * (size_t)kMaxLists * (2 * cap + 1));
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
    // launches of one phase (independent case lists) run side by side on the
    // caller's stream and these, forked and joined with events
    std::vector<hipStream_t> aux;
    int n_par = 3;  // auxiliary streams in use (<= aux.size())
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    std::string last_error;
};

constexpr int kAuxStreams = 3;  // with the caller's stream: four, the process's hardware queues

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

// One launch of the batch: which kernel, on which records, with which case list.
struct Launch {
    enum Kind { XF, XI, ZFB, ZFUSED, M00, SAVE, CORR } kind;
    int f32;      // spectrum precision
    int mode;     // x forward mode, or ACCUM for the inverse
    int R;        // z-split radix (x kernels)
    size_t tab;   // offset of the case records
    int list;     // case list
    size_t lds;
    int phase;    // launches of one phase are independent and may run concurrently
};

void release_streams(dfp_batch* b) {
    for (hipStream_t st : b->aux) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : b->ev_join) (void)hipEventDestroy(ev);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    b->aux.clear();
    b->ev_join.clear();
    b->ev_fork = nullptr;
}

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
        release_streams(b);
        delete b;
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(DFP_E_HIP);
    if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || b->n_cu < 1)
        return bail(DFP_E_HIP);
    if (hipHostMalloc((void**)&b->h_buf, b->L.total, hipHostMallocDefault) != hipSuccess) return bail(DFP_E_NOMEM);
    if (hipMalloc(&b->d_buf, b->L.total) != hipSuccess) return bail(DFP_E_NOMEM);
    if (hipEventCreateWithFlags(&b->copied, hipEventDisableTiming) != hipSuccess) return bail(DFP_E_HIP);
    if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess) return bail(DFP_E_HIP);
    for (int q = 0; q < kAuxStreams; ++q) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return bail(DFP_E_HIP);
        b->aux.push_back(st);
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return bail(DFP_E_HIP);
        b->ev_join.push_back(ev);
    }
    memset(b->h_buf, 0, b->L.total);
    *out = b;
    return DFP_OK;
}

int dfp_batch_destroy(dfp_batch* b) {
    if (!b) return DFP_E_ARG;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    destroy_graphs(b);
    for (hipStream_t st : b->aux) (void)hipStreamSynchronize(st);
    (void)hipHostFree(b->h_buf);
    (void)hipFree(b->d_buf);
    (void)hipEventDestroy(b->copied);
    release_streams(b);
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

int dfp_batch_set_parallel(dfp_batch* b, int aux_streams) {
    if (!b) return DFP_E_ARG;
    if (aux_streams < 0 || aux_streams > (int)b->aux.size())
        return bfail(b, DFP_E_ARG, "auxiliary streams out of range (0..%d)", (int)b->aux.size());
    if (aux_streams != b->n_par) destroy_graphs(b);
    b->n_par = aux_streams;
    return DFP_OK;
}

const char* dfp_batch_last_error(const dfp_batch* b) { return b ? b->last_error.c_str() : ""; }

int dfp_batch_project(dfp_batch* b, int n, dfp_plan* const* plans, double* const* U, double* const* V,
                      double* const* W, double* const* rhs_p, double* const* P, const double* dt, const double* c,
                      const int* flags, int batch_flags) {
    if (!b) return DFP_E_ARG;
    if (n < 0 || n > b->cap) return bfail(b, DFP_E_ARG, "batch of %d cases (capacity %d)", n, b->cap);
    if (batch_flags & ~(DFP_ASYNC | DFP_SOLVE_ONLY))
        return bfail(b, DFP_E_ARG, "unknown batch flag bits 0x%x", batch_flags);
    // solve only: rhs_p -> phi (+ P) as dfp_solve, no divergence, no correction
    const bool solve = (batch_flags & DFP_SOLVE_ONLY) != 0;
    if (n == 0) return DFP_OK;
    if (!plans || !rhs_p || !dt || !c || (!solve && (!U || !V || !W)))
        return bfail(b, DFP_E_ARG, "null argument array");
    BCHK(b, hipSetDevice(b->device));
    // validate and split by precision (FP32 cases by descending sweeps)
    std::vector<int> i64, i32;
    for (int i = 0; i < n; ++i) {
        dfp_plan* p = plans[i];
        if (!p) return bfail(b, DFP_E_ARG, "case %d: null plan", i);
        if (p->device != b->device) return bfail(b, DFP_E_ARG, "case %d: plan on device %d", i, p->device);
        if (p->geo.xbc != 0 || p->geo.ybc != 1 || p->geo.zbc != 0)
            return bfail(b, DFP_E_UNSUPPORTED, "case %d: the batch executor takes channel plans (periodic x, y walls) only",
                         i);
        for (int k = 0; k < i; ++k)
            if (plans[k] == p) return bfail(b, DFP_E_ARG, "cases %d and %d share a plan", k, i);
        if (!rhs_p[i] || (!solve && (!U[i] || !V[i] || !W[i]))) return bfail(b, DFP_E_ARG, "case %d: null field", i);
        const int fl = flags ? flags[i] : 0;
        if (fl & ~(DFP_SAVE_P | DFP_FIELDS_ON_DEVICE))
            return bfail(b, DFP_E_ARG, "case %d: flags 0x%x (device fields, DFP_SAVE_P only)", i, fl);
        if ((fl & DFP_SAVE_P) && (!P || !P[i] || !(dt[i] * c[i] != 0.0) || !std::isfinite(dt[i] * c[i])))
            return bfail(b, DFP_E_ARG, "case %d: DFP_SAVE_P needs P and a finite nonzero dt*c", i);
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
    auto* x64i = reinterpret_cast<XPar<double2>*>(H + L.x64i);
    auto* x32i = reinterpret_cast<XPar<float2>*>(H + L.x32i);
    auto* x32si = reinterpret_cast<XPar<float2>*>(H + L.x32si);
    auto* z64 = reinterpret_cast<ZPar<double2>*>(H + L.z64);
    auto* x32 = reinterpret_cast<XPar<float2>*>(H + L.x32);
    auto* x32s = reinterpret_cast<XPar<float2>*>(H + L.x32s);
    auto* z32 = reinterpret_cast<ZPar<float2>*>(H + L.z32);
    auto* m32 = reinterpret_cast<MPar*>(H + L.m32);
    auto* s32 = reinterpret_cast<SPar*>(H + L.s32);
    auto* cc = reinterpret_cast<CPar*>(H + L.c);
    int* counts = reinterpret_cast<int*>(H + L.counts);
    int* pool = reinterpret_cast<int*>(H + L.lists);
    const size_t list_stride = 2 * (size_t)b->cap + 1;

    std::vector<size_t> w64x(n64), w64xi(n64), w64z(n64), w32x(n32), w32xi(n32), w32z(n32), wc(n), ws;
    std::vector<int> R64(n64), F64(n64), R32(n32), F32(n32);
    for (int a = 0; a < n64; ++a) {
        const int i = i64[a];
        dfp_plan* p = plans[i];
        const int fl = flags ? flags[i] : 0;
        XPar<double2> q = make_xpar<double2>(p);
        if (solve) q.a.rhs = rhs_p[i];
        else { q.a.U = U[i]; q.a.V = V[i]; q.a.W = W[i]; }
        q.rhs = rhs_p[i];
        q.save_p = (fl & DFP_SAVE_P) ? 1 : 0;
        q.P = q.save_p ? P[i] : nullptr;
        q.dtc = q.save_p ? dt[i] * c[i] : 1.0;
        x64[a] = q;
        xpar_columns(q, p, 1);
        x64i[a] = q;
        z64[a] = make_zpar<double2>(p, 0);
        w64x[a] = (size_t)x64[a].nbx * p->geo.nym;
        w64xi[a] = (size_t)x64i[a].nbx * p->geo.nym;
        w64z[a] = z64[a].nblocks;
        R64[a] = p->zr;
        F64[a] = p->zfused;
    }
    int ns = 0;
    std::vector<int> n_at_sweep(kMaxSweepStages + 1, 0);  // FP32 cases with sweeps > s
    std::vector<int> save_rec;
    for (int a = 0; a < n32; ++a) {
        const int i = i32[a];
        dfp_plan* p = plans[i];
        const int fl = flags ? flags[i] : 0;
        XPar<float2> q = make_xpar<float2>(p);
        if (solve) q.a.rhs = rhs_p[i];
        else { q.a.U = U[i]; q.a.V = V[i]; q.a.W = W[i]; }
        q.a.b_out = p->d_b;
        q.a.rowpart = p->d_rowpart;
        q.rhs = rhs_p[i];
        q.pbar = p->d_pbar;
        q.part0 = p->d_part0;
        x32[a] = q;
        x32i[a] = q;
        xpar_columns(x32i[a], p, 1);
        XPar<float2> r = q;  // sweeps: residual forward, accumulating inverse
        r.a = XArgs{};
        r.a.rhs = rhs_p[i];
        r.a.b = p->d_b;
        r.a.rowpart = p->d_rowpart;
        r.a.part0 = p->d_part0;
        r.a.idx2 = 1.0 / (p->dx * p->dx);
        r.a.idz2 = 1.0 / (p->dz * p->dz);
        x32s[a] = r;
        x32si[a] = r;
        xpar_columns(x32si[a], p, 1);
        z32[a] = make_zpar<float2>(p, 1);
        m32[a] = make_mpar(p);
        w32x[a] = (size_t)q.nbx * p->geo.nym;
        w32xi[a] = (size_t)x32i[a].nbx * p->geo.nym;
        w32z[a] = z32[a].nblocks;
        R32[a] = p->zr;
        F32[a] = p->zfused;
        for (int s = 0; s < std::min(p->sweeps, kMaxSweepStages); ++s) n_at_sweep[s] = a + 1;
        if (fl & DFP_SAVE_P) {
            s32[ns] = make_spar(p, rhs_p[i], P[i], dt[i] * c[i]);
            ws.push_back(s32[ns].nblocks);
            save_rec.push_back(ns);
            ++ns;
        }
    }
    for (int a = 0; a < n && !solve; ++a) {
        const int i = a < n64 ? i64[a] : i32[a - n64];
        dfp_plan* p = plans[i];
        cc[a] = make_cpar(p, U[i], V[i], W[i], rhs_p[i]);
        wc[a] = (size_t)cc[a].nbx * correct_blocks_z(p->geo);
    }

    // case lists: records recs with work items w[rec].  A list slot holds the
    // offsets at [0, cap] and the record indices at [cap + 1, 2 cap]: fixed
    // positions, so a launch (and a graph that captured it) reads the same
    // addresses whatever the list's length
    int nlists = 0;
    bool overflow = false;
    auto make_list = [&](const std::vector<int>& recs, const std::vector<size_t>& w) {
        if (nlists >= kMaxLists) { overflow = true; return 0; }
        int* o = pool + (size_t)nlists * list_stride;
        int* ix = o + b->cap + 1;
        o[0] = 0;
        for (size_t k = 0; k < recs.size(); ++k) {
            o[k + 1] = o[k] + (int)w[recs[k]];
            ix[k] = recs[k];
        }
        counts[nlists] = (int)recs.size();
        return nlists++;
    };
    const size_t lds_limit = plans[0]->lds_limit;
    std::vector<Launch> ls;
    int phase = 0;
    auto add = [&](Launch::Kind k, int f32, int mode, int R, size_t tab, int list, size_t lds) {
        ls.push_back(Launch{k, f32, mode, R, tab, list, lds_bucket(lds, lds_limit), phase});
    };
    // the x launches of a set of records, one per radix present; the z launches
    // for its fused and split members
    auto x_launches = [&](Launch::Kind k, int f32, int mode, size_t tab, const std::vector<int>& recs,
                          const std::vector<int>& Rs, const std::vector<size_t>& w, int inv) {
        for (int R : kRadixList) {
            std::vector<int> sel;
            size_t lds = 0;
            bool wide = false;  // a forward tile too large for two butterfly slots: the four-slot kernel
            for (int a : recs)
                if (Rs[a] == R) {
                    const dfp_plan* p = plans[f32 ? i32[a] : i64[a]];
                    sel.push_back(a);
                    lds = std::max(lds, x_lds_bytes(p, inv, f32));
                    wide = wide || (!inv && xfwd_wide(p->fx, R, p->tk_x));
                }
            if (!sel.empty()) add(k, f32, wide ? mode | 8 : mode, R, tab, make_list(sel, w), lds);
        }
        ++phase;
    };
    auto z_launches = [&](int f32, size_t tab, const std::vector<int>& recs, const std::vector<int>& Fs,
                          const std::vector<size_t>& w) {
        // split-pair cases grouped by kernel variant (Launch::R holds it)
        std::vector<int> fz;
        std::vector<std::pair<int, std::vector<int>>> sz;
        std::vector<size_t> lsz;
        size_t lf = 0;
        for (int a : recs) {
            dfp_plan* p = plans[f32 ? i32[a] : i64[a]];
            if (Fs[a]) { fz.push_back(a); lf = std::max(lf, zf_lds_bytes(p, f32)); continue; }
            const int v = z_variant(p);
            size_t k = 0;
            while (k < sz.size() && sz[k].first != v) ++k;
            if (k == sz.size()) { sz.push_back({v, {}}); lsz.push_back(0); }
            sz[k].second.push_back(a);
            lsz[k] = std::max(lsz[k], z_lds_bytes(p, f32));
        }
        // one phase: the fused kernel and the split pairs (each workgroup
        // eliminates a sub-plane and substitutes back through it)
        if (!fz.empty()) add(Launch::ZFUSED, f32, 0, 0, tab, make_list(fz, w), lf);
        for (size_t k = 0; k < sz.size(); ++k) add(Launch::ZFB, f32, 0, sz[k].first, tab, make_list(sz[k].second, w), lsz[k]);
        ++phase;
    };
    auto first = [](int k) { std::vector<int> v(k); for (int a = 0; a < k; ++a) v[a] = a; return v; };
    if (n64) {
        const std::vector<int> all = first(n64);
        x_launches(Launch::XF, 0, solve ? 1 : 0, L.x64, all, R64, w64x, 0);
        z_launches(0, L.z64, all, F64, w64z);
        x_launches(Launch::XI, 0, 0, L.x64i, all, R64, w64xi, 1);
    }
    if (n32) {
        const std::vector<int> all = first(n32);
        const std::vector<size_t> one(n32, 1);
        x_launches(Launch::XF, 1, solve ? 1 : 0, L.x32, all, R32, w32x, 0);
        add(Launch::M00, 1, 0, 0, L.m32, make_list(all, one), 0);
        ++phase;
        z_launches(1, L.z32, all, F32, w32z);
        x_launches(Launch::XI, 1, 0, L.x32i, all, R32, w32xi, 1);
        for (int s = 0; s < max_sweeps; ++s) {
            const std::vector<int> pre = first(n_at_sweep[s]);
            x_launches(Launch::XF, 1, 2, L.x32s, pre, R32, w32x, 0);
            add(Launch::M00, 1, 0, 0, L.m32, make_list(pre, one), 0);
            ++phase;
            z_launches(1, L.z32, pre, F32, w32z);
            x_launches(Launch::XI, 1, 1, L.x32si, pre, R32, w32xi, 1);
        }
        if (ns) add(Launch::SAVE, 1, 0, 0, L.s32, make_list(save_rec, ws), 0);
        ++phase;
    }
    if (!solve) add(Launch::CORR, 0, 0, 0, L.c, make_list(first(n), wc), 0);
    if (overflow) return bfail(b, DFP_E_UNSUPPORTED, "too many case lists in one call");

    // tables to the device (up to the last list used), then the launches (or their graph)
    const size_t used = L.lists + sizeof(int) * (size_t)nlists * list_stride;
    BCHK(b, hipMemcpyAsync(b->d_buf, b->h_buf, used, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipEventRecord(b->copied, b->stream));
    b->copy_pending = true;

    unsigned char* D = b->d_buf;
    const int* d_counts = reinterpret_cast<const int*>(D + L.counts);
    const int* d_pool = reinterpret_cast<const int*>(D + L.lists);
    const int gx = b->n_cu * b->occ_x, gz = b->n_cu * b->occ_z, gs = b->n_cu * b->occ_s;
    // one launch on stream st
    auto run_one = [&](const Launch& l, hipStream_t st) -> hipError_t {
        {
            const int* o = d_pool + (size_t)l.list * list_stride;
            const int* ix = o + b->cap + 1;
            const int* nc = d_counts + l.list;
            hipError_t e = hipSuccess;
            switch (l.kind) {
                case Launch::XF:
                    e = l.f32 ? launch_xfwd_batch(st, l.R, l.mode, reinterpret_cast<const XPar<float2>*>(D + l.tab),
                                                  o, ix, nc, gx, l.lds)
                              : launch_xfwd_batch(st, l.R, l.mode, reinterpret_cast<const XPar<double2>*>(D + l.tab),
                                                  o, ix, nc, gx, l.lds);
                    break;
                case Launch::XI:
                    e = l.f32 ? launch_xinv_batch(st, l.R, l.mode, reinterpret_cast<const XPar<float2>*>(D + l.tab),
                                                  o, ix, nc, gx, l.lds)
                              : launch_xinv_batch(st, l.R, l.mode, reinterpret_cast<const XPar<double2>*>(D + l.tab),
                                                  o, ix, nc, gx, l.lds);
                    break;
                case Launch::ZFB:
                    e = l.f32 ? launch_zsolve_batch(st, l.R, reinterpret_cast<const ZPar<float2>*>(D + l.tab), o,
                                                    ix, nc, gz, l.lds)
                              : launch_zsolve_batch(st, l.R, reinterpret_cast<const ZPar<double2>*>(D + l.tab), o,
                                                    ix, nc, gz, l.lds);
                    break;
                case Launch::ZFUSED:
                    e = l.f32 ? launch_zfused_batch(st, reinterpret_cast<const ZPar<float2>*>(D + l.tab), o, ix, nc,
                                                    gz, l.lds)
                              : launch_zfused_batch(st, reinterpret_cast<const ZPar<double2>*>(D + l.tab), o, ix, nc,
                                                    gz, l.lds);
                    break;
                case Launch::M00:
                    e = launch_mode00_batch(st, reinterpret_cast<const MPar*>(D + l.tab), ix, nc, std::min(b->cap, 256));
                    break;
                case Launch::SAVE:
                    e = launch_save_p_batch(st, reinterpret_cast<const SPar*>(D + l.tab), o, ix, nc, gs);
                    break;
                case Launch::CORR:
                    e = launch_correct_batch(st, reinterpret_cast<const CPar*>(D + l.tab), o, ix, nc, gs);
                    break;
            }
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    // every launch in order; a phase with several launches forks onto the
    // auxiliary streams and joins back into st before the next phase
    auto run_all = [&](hipStream_t st) -> hipError_t {
        size_t i = 0;
        while (i < ls.size()) {
            size_t k = i;
            while (k < ls.size() && ls[k].phase == ls[i].phase) ++k;
            if (k - i == 1) {
                if (hipError_t e = run_one(ls[i], st)) return e;
            } else {
                const int ns = std::min((int)(k - i), b->n_par + 1);
                if (hipError_t e = hipEventRecord(b->ev_fork, st)) return e;
                for (int q = 1; q < ns; ++q)
                    if (hipError_t e = hipStreamWaitEvent(b->aux[q - 1], b->ev_fork, 0)) return e;
                for (size_t m = i; m < k; ++m) {
                    const int q = (int)((m - i) % ns);
                    if (hipError_t e = run_one(ls[m], q == 0 ? st : b->aux[q - 1])) return e;
                }
                for (int q = 1; q < ns; ++q) {
                    if (hipError_t e = hipEventRecord(b->ev_join[q - 1], b->aux[q - 1])) return e;
                    if (hipError_t e = hipStreamWaitEvent(st, b->ev_join[q - 1], 0)) return e;
                }
            }
            i = k;
        }
        return hipSuccess;
    };

    if (!b->use_graphs) {
        BCHK(b, run_all(b->stream));
    } else {
        // key: the launch list as the kernels see it (kind, precision, mode, radix, records, list, LDS)
        std::vector<long> key;
        for (const Launch& l : ls) {
            key.push_back(l.kind);
            key.push_back(l.f32 * 16 + l.mode);
            key.push_back(l.R);
            key.push_back((long)l.tab);
            key.push_back(l.list);
            key.push_back((long)l.lds);
            key.push_back(l.phase);
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
