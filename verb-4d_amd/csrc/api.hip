// C-ABI of the VERB-4D step library: contexts, per-call drop-in entries
// (host pointers, synchronous) and batched full-grid entries (device
// pointers, asynchronous on the context's stream).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "gmres.h"

namespace verb {
// the learned warm start (unet.hip)
int unet_warm_start(verb_unet* u, const verb_grid& g, const double* psd, int node0, int count, uint16_t* feat,
                    float* pin, double* scale, float* delta, double* x0, hipStream_t st);
int unet_max_batch(const verb_unet* u);
verb_ctx* unet_ctx(const verb_unet* u);
}  // namespace verb


struct verb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int* status = nullptr;  // device status words
    std::mutex mu;
    std::string last_error;
    // grow-only device scratch for the per-call entries and the plain
    // (uncached) local-diffusion path
    std::vector<std::pair<void*, size_t>> bufs;
    size_t scratch_bytes = 0;
    // bumped whenever a scratch slot moves or the device-side extent tables
    // change: launches captured into a graph hold those addresses/contents
    unsigned long long scratch_gen = 0;
    // kernel-phase timing
    bool timing = false;
    struct Mark { int phase; hipEvent_t a, b; };
    std::vector<Mark> marks;
    double ms[VERB_NUM_PHASES] = {0};
    long counts[VERB_NUM_PHASES] = {0};
    // launch shapes (verb_set_tuning / verb_autotune)
    int conv_tq = 0;
    int conv_kernel = 0;  // 0: line kernel where instantiated, 1: tile kernel
    int lu_variant = -1;
    int lu_persistent = 0;
    // captured plain step (verb_step_graph) and what it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    verb_grid graph_grid{};
    double* graph_psd = nullptr;
    int graph_tune[4] = {0, -1, 0, 0};
    std::vector<int> graph_k0;  // pitch-angle extents the graph was captured for (empty: uniform)
    unsigned long long graph_gen = 0;  // scratch_gen at capture
    // the uneven extents last uploaded (first column, unknowns, dispatch order)
    std::vector<int> uneven_k0;
    // rocBLAS handle of the plane data assimilation (created on first use)
    void* blas = nullptr;
};

namespace {

int set_err(verb_ctx* c, int code, const char* what) {
    if (c) c->last_error = what;
    return code;
}

#define HIPCHK(ctx, call)                                                          \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) return set_err(ctx, VERB_ERR_HIP, hipGetErrorString(e_)); \
    } while (0)

// Scratch slot `slot` of at least `bytes` bytes (contents not preserved).
void* scratch(verb_ctx* c, int slot, size_t bytes) {
    if ((int)c->bufs.size() <= slot) c->bufs.resize(slot + 1, {nullptr, 0});
    auto& b = c->bufs[slot];
    if (b.second < bytes) {
        if (b.first) hipFree(b.first);
        c->scratch_bytes -= b.second;
        b.first = nullptr;
        b.second = 0;
        c->scratch_gen++;
        if (hipMalloc(&b.first, bytes) != hipSuccess) return nullptr;
        b.second = bytes;
        c->scratch_bytes += bytes;
    }
    return b.first;
}

int reset_status(verb_ctx* c) {
    int init[verb::kStatusWords] = {0, 0x7fffffff, -1, 0};
    HIPCHK(c, hipMemcpyAsync(c->status, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    return VERB_OK;
}

// Synchronise and translate the device status words into a return code.
int finish(verb_ctx* c) {
    int st[verb::kStatusWords];
    HIPCHK(c, hipMemcpyAsync(st, c->status, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st[0] == VERB_ERR_COURANT) return set_err(c, VERB_ERR_COURANT, "max(CourNum) > 1, calculation can't be performed");
    if (st[0] == VERB_ERR_BC) return set_err(c, VERB_ERR_BC, "unknown boundary type");
    if (st[0] == VERB_ERR_SINGULAR) return set_err(c, VERB_ERR_SINGULAR, "trans_solve: zero pivot");
    if (st[0] == VERB_ERR_DOMAIN)
        return set_err(c, VERB_ERR_DOMAIN, "non-finite input or result the reference rejects (interpolation input, "
                                           "a NaN Kalman analysis)");
    if (st[0] != 0) return set_err(c, st[0], "device error");
    if (st[1] != 0x7fffffff) return set_err(c, st[1], "Error computing LU decomposition (zero pivot)");
    return VERB_OK;
}

bool bc_ok(int t) { return t >= 0 && t <= 2; }

// Upload `count` doubles to scratch slot `slot`.
const double* up(verb_ctx* c, int slot, const double* h, size_t count, int& rc) {
    if (rc) return nullptr;
    if (!h) { rc = set_err(c, VERB_ERR_ARG, "null input array"); return nullptr; }
    double* d = (double*)scratch(c, slot, count * sizeof(double));
    if (!d) { rc = set_err(c, VERB_ERR_NOMEM, "device allocation failed"); return nullptr; }
    if (hipMemcpyAsync(d, h, count * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, VERB_ERR_HIP, "host to device copy failed");
    return d;
}

struct LocalBuffers {
    verb::LocalWork w;
    int rc = 0;
};

constexpr int kSlotQueue = 48;  // persistent-LU node counter

LocalBuffers local_buffers(verb_ctx* c, const verb_grid* g, double* lu_cache, int64_t* ipiv_cache, int base_slot) {
    LocalBuffers lb;
    const size_t n = (size_t)g->nV * g->nK;
    const size_t nodes = (size_t)g->nP * g->nR;
    const size_t ldab = 3 * (size_t)(g->nV + 1) + 1;
    if (lu_cache) {
        lb.w.ab = lu_cache;
        lb.w.ipiv = ipiv_cache;
    } else {
        lb.w.ab = (double*)scratch(c, base_slot + 0, ldab * n * nodes * sizeof(double));
        lb.w.ipiv = (int64_t*)scratch(c, base_slot + 1, n * nodes * sizeof(int64_t));
    }
    lb.w.b0 = (double*)scratch(c, base_slot + 2, n * nodes * sizeof(double));
    lb.w.c0 = (double*)scratch(c, base_slot + 3, n * nodes * sizeof(double));
    lb.w.rhs = (double*)scratch(c, base_slot + 4, n * nodes * sizeof(double));
    lb.w.diag = (double*)scratch(c, base_slot + 5, 9 * n * nodes * sizeof(double));
    lb.w.y = (double*)scratch(c, base_slot + 6, n * nodes * sizeof(double));
    lb.w.uext = (int*)scratch(c, base_slot + 7, n * nodes * sizeof(int));
    // the persistent LU's counters: node slots, back-substitution slots, one
    // "factors written" flag per node
    lb.w.queue = c->lu_persistent ? (int*)scratch(c, kSlotQueue, sizeof(int) * (2 + (size_t)nodes)) : nullptr;
    lb.w.lu_variant = c->lu_variant;
    if (!lb.w.ab || !lb.w.ipiv || !lb.w.b0 || !lb.w.c0 || !lb.w.rhs || !lb.w.diag || !lb.w.y || !lb.w.uext ||
        (c->lu_persistent && !lb.w.queue))
        lb.rc = set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    return lb;
}

constexpr int kSlotLocal = 40;  // scratch slots 40..47 for the local workspace

hipEvent_t new_event() {
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

// Records a start event for `phase` when timing is on; returns its index.
int mark_begin(verb_ctx* c, int phase) {
    if (!c->timing) return -1;
    verb_ctx::Mark m{phase, new_event(), new_event()};
    hipEventRecord(m.a, c->stream);
    c->marks.push_back(m);
    return (int)c->marks.size() - 1;
}

void mark_end(verb_ctx* c, int idx) {
    if (idx >= 0) hipEventRecord(c->marks[idx].b, c->stream);
}

}  // namespace

namespace verb {
int ctx_error(verb_ctx* c, int code, const char* what) { return set_err(c, code, what); }
void* ctx_scratch(verb_ctx* c, int slot, size_t bytes) { return scratch(c, slot, bytes); }
}  // namespace verb

extern "C" {

int verb_ctx_create(verb_ctx** out) {
    if (!out) return VERB_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return VERB_ERR_NODEVICE;
    verb_ctx* c = new verb_ctx();
    hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VERB_ERR_HIP;
    }
    if (hipMalloc(&c->status, sizeof(int) * verb::kStatusWords) != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return VERB_ERR_NOMEM;
    }
    if (reset_status(c) != VERB_OK || hipStreamSynchronize(c->stream) != hipSuccess) {
        hipFree(c->status);
        hipStreamDestroy(c->stream);
        delete c;
        return VERB_ERR_HIP;
    }
    *out = c;
    return VERB_OK;
}

int verb_ctx_destroy(verb_ctx* c) {
    if (!c) return VERB_ERR_ARG;
    hipStreamSynchronize(c->stream);
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    verb::da_conv_release(&c->blas);
    for (auto& b : c->bufs)
        if (b.first) hipFree(b.first);
    hipFree(c->status);
    hipStreamDestroy(c->stream);
    delete c;
    return VERB_OK;
}

void* verb_ctx_stream(verb_ctx* c) { return c ? (void*)c->stream : nullptr; }

int verb_ctx_synchronize(verb_ctx* c) {
    if (!c) return VERB_ERR_ARG;
    return hipStreamSynchronize(c->stream) == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

const char* verb_ctx_last_error(verb_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

size_t verb_ctx_scratch_bytes(verb_ctx* c) { return c ? c->scratch_bytes : 0; }

int verb_ctx_set_timing(verb_ctx* c, int enable) {
    if (!c) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->timing = enable != 0;
    return VERB_OK;
}

int verb_ctx_get_timings(verb_ctx* c, double* ms, long* counts) {
    if (!c) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& m : c->marks) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, m.a, m.b) == hipSuccess) {
            c->ms[m.phase] += t;
            c->counts[m.phase] += 1;
        }
        hipEventDestroy(m.a);
        hipEventDestroy(m.b);
    }
    c->marks.clear();
    for (int i = 0; i < VERB_NUM_PHASES; i++) {
        if (ms) ms[i] = c->ms[i];
        if (counts) counts[i] = c->counts[i];
        c->ms[i] = 0;
        c->counts[i] = 0;
    }
    return VERB_OK;
}

// ---------------------------------------------------------------------------
// Batched entries
// ---------------------------------------------------------------------------
int verb_check_status(verb_ctx* c) {
    if (!c) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = finish(c);
    int r2 = reset_status(c);
    hipStreamSynchronize(c->stream);
    return rc ? rc : r2;
}

static int grid_ok(const verb_grid* g) {
    if (!g || g->nP < 1 || g->nR < 1 || g->nV < 1 || g->nK < 1) return 0;
    if (!(g->dt > 0)) return 0;
    return 1;
}

int verb_convection_batched_ex(verb_ctx* c, const verb_grid* g, double* psd, const verb_convection_options* opt) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const int m = mark_begin(c, VERB_PHASE_CONVECTION);
    int rc = verb::launch_convection(*g, psd, c->status, c->stream, c->conv_tq, opt ? opt->psd_lost : nullptr,
                                     opt ? opt->fast_convection : 0, c->conv_kernel);
    mark_end(c, m);
    return rc ? set_err(c, rc, "convection launch rejected") : VERB_OK;
}

int verb_convection_batched(verb_ctx* c, const verb_grid* g, double* psd) {
    return verb_convection_batched_ex(c, g, psd, nullptr);
}

int verb_radial_diffusion_batched(verb_ctx* c, const verb_grid* g, double* psd) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    if (!bc_ok(g->Ll_type) || !bc_ok(g->Lu_type)) return VERB_ERR_BC;
    std::lock_guard<std::mutex> lk(c->mu);
    const int m = mark_begin(c, VERB_PHASE_RADIAL);
    int rc = verb::launch_radial(*g, psd, c->status, c->stream);
    mark_end(c, m);
    return rc ? set_err(c, rc, "radial launch rejected") : VERB_OK;
}

int verb_local_assembly(verb_ctx* c, const verb_grid* g, const double* psd, double* diag, double* b0, double* c0) {
    if (!c || !grid_ok(g) || !psd || !diag || !b0 || !c0 || !(g->sub_dt > 0)) return VERB_ERR_ARG;
    if (g->nV < 3 || g->nK < 3) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t n = (size_t)g->nV * g->nK * g->nP * g->nR;
    double* rhs = (double*)scratch(c, kSlotLocal + 4, n * sizeof(double));
    if (!rhs) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    int rc = verb::launch_local_assembly(*g, psd, diag, b0, c0, rhs, c->status, c->stream);
    return rc ? set_err(c, rc, "assembly launch rejected") : VERB_OK;
}

size_t verb_lu_cache_doubles(const verb_grid* g) {
    if (!g) return 0;
    const size_t n = (size_t)g->nV * g->nK;
    return (3 * (size_t)(g->nV + 1) + 1) * n * (size_t)g->nP * g->nR;
}

size_t verb_lu_cache_ints(const verb_grid* g) {
    if (!g) return 0;
    return (size_t)g->nV * g->nK * (size_t)g->nP * g->nR;
}

int verb_local_diffusion_batched(verb_ctx* c, const verb_grid* g, double* psd, double* lu_cache,
                                 int64_t* ipiv_cache, int recompute_lu) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    if ((lu_cache == nullptr) != (ipiv_cache == nullptr)) return VERB_ERR_ARG;
    if (!(g->sub_dt > 0)) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    LocalBuffers lb = local_buffers(c, g, lu_cache, ipiv_cache, kSlotLocal);
    if (lb.rc) return lb.rc;
    const bool factor = (lu_cache == nullptr) || recompute_lu;
    hipEvent_t ev[6];
    hipEvent_t* marks = nullptr;
    if (c->timing) {
        for (auto& e : ev) e = new_event();
        marks = ev;
    }
    int rc = verb::launch_local(*g, psd, lb.w, factor, lu_cache != nullptr, c->status, c->stream, marks);
    if (marks) {
        c->marks.push_back({VERB_PHASE_LOCAL_ASSEMBLE, ev[0], ev[1]});
        c->marks.push_back({VERB_PHASE_LOCAL_FACTOR, ev[2], ev[3]});
        c->marks.push_back({VERB_PHASE_LOCAL_SOLVE, ev[4], ev[5]});
    }
    return rc ? set_err(c, rc, "local diffusion launch rejected") : VERB_OK;
}

int verb_local_diffusion_uneven(verb_ctx* c, const verb_grid* g, double* psd, const int* k0, double* lu_cache,
                                int64_t* ipiv_cache, int recompute_lu) {
    if (!c || !grid_ok(g) || !psd || !k0) return VERB_ERR_ARG;
    if ((lu_cache == nullptr) != (ipiv_cache == nullptr)) return VERB_ERR_ARG;
    if (!(g->sub_dt > 0)) return VERB_ERR_ARG;
    const int nodes = g->nP * g->nR;
    std::vector<int> hk(k0, k0 + nodes), hn(nodes), ord(nodes);
    for (int i = 0; i < nodes; i++) {
        if (hk[i] < 0 || g->nK - hk[i] < 3) return set_err(c, VERB_ERR_ARG, "active pitch-angle extent below 3");
        hn[i] = g->nV * (g->nK - hk[i]);
        ord[i] = i;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return hn[a] > hn[b]; });
    std::lock_guard<std::mutex> lk(c->mu);
    int* dk = (int*)scratch(c, 50, nodes * sizeof(int));
    int* dn = (int*)scratch(c, 51, nodes * sizeof(int));
    int* dord = (int*)scratch(c, 52, nodes * sizeof(int));
    if (!dk || !dn || !dord) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    if (c->uneven_k0 != hk) {
        // new extents: upload them once; a call with the same extents launches
        // kernels only (so the step can be captured into a graph)
        HIPCHK(c, hipMemcpyAsync(dk, hk.data(), nodes * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dn, hn.data(), nodes * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dord, ord.data(), nodes * sizeof(int), hipMemcpyHostToDevice, c->stream));
        // the host vectors die at return: wait for the copies
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->uneven_k0 = hk;
        c->scratch_gen++;
    }
    LocalBuffers lb = local_buffers(c, g, lu_cache, ipiv_cache, kSlotLocal);
    if (lb.rc) return lb.rc;
    const bool factor = (lu_cache == nullptr) || recompute_lu;
    verb::NodeSizes sz{dk, dn, dord};
    hipEvent_t ev[6];
    hipEvent_t* marks = nullptr;
    if (c->timing) {
        for (auto& e : ev) e = new_event();
        marks = ev;
    }
    int rc = verb::launch_local(*g, psd, lb.w, factor, lu_cache != nullptr, c->status, c->stream, marks, &sz);
    if (marks) {
        c->marks.push_back({VERB_PHASE_LOCAL_ASSEMBLE, ev[0], ev[1]});
        c->marks.push_back({VERB_PHASE_LOCAL_FACTOR, ev[2], ev[3]});
        c->marks.push_back({VERB_PHASE_LOCAL_SOLVE, ev[4], ev[5]});
    }
    return rc ? set_err(c, rc, "uneven local diffusion launch rejected") : VERB_OK;
}

int verb_step_uneven(verb_ctx* c, const verb_grid* g, double* psd, const int* k0, int run_convection, int run_radial,
                     int run_local) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    int rc = VERB_OK;
    if (run_convection && (3 < g->nP || 3 < g->nR)) rc = verb_convection_batched(c, g, psd);
    if (!rc && run_radial && g->nR >= 3) rc = verb_radial_diffusion_batched(c, g, psd);
    if (!rc && run_local && g->nV >= 3 && g->nK >= 3)
        rc = verb_local_diffusion_uneven(c, g, psd, k0, nullptr, nullptr, 1);
    return rc;
}

constexpr int kSlotGmres = 60;  // scratch slots 60..66 for the GMRES workspace
constexpr int kSlotTune = 70;   // the autotuner's copy of the PSD

}  // extern "C"

namespace {

int gmres_local(verb_ctx* c, const verb_grid* g, double* psd, const double* x0, const verb_gmres_params* prm,
                verb_gmres_stats* stats) {
    if (!c || !grid_ok(g) || !psd || !prm) return VERB_ERR_ARG;
    if (!(g->sub_dt > 0) || (int)(g->dt / g->sub_dt) > 1) return set_err(c, VERB_ERR_ARG, "GMRES path: one sub-step only");
    if (prm->restart < 1 || prm->restart > verb::kGmMaxM || prm->max_cycles < 0 || prm->max_iterations < 0 ||
        !(prm->tol > 0))
        return set_err(c, VERB_ERR_ARG, "GMRES parameters out of range");
    if (g->nV + 1 > 112) return set_err(c, VERB_ERR_ARG, "GMRES path: fallback LU needs nV <= 111");
    const size_t n = (size_t)g->nV * g->nK;
    const int nodes = g->nP * g->nR;
    std::lock_guard<std::mutex> lk(c->mu);
    LocalBuffers lb = local_buffers(c, g, nullptr, nullptr, kSlotLocal);
    if (lb.rc) return lb.rc;
    verb::GmresWork gw;
    gw.x = (double*)scratch(c, kSlotGmres + 0, n * nodes * sizeof(double));
    gw.d32 = (float*)scratch(c, kSlotGmres + 1, 9 * n * nodes * sizeof(float));
    gw.pf = (float*)scratch(c, kSlotGmres + 2, 3 * n * nodes * sizeof(float));
    gw.V = (float*)scratch(c, kSlotGmres + 3, (size_t)(verb::kGmMaxM + 1) * n * nodes * sizeof(float));
    gw.z = (float*)scratch(c, kSlotGmres + 4, 2 * n * nodes * sizeof(float));
    gw.conv = (int*)scratch(c, kSlotGmres + 5, nodes * sizeof(int));
    gw.iters = (int*)scratch(c, kSlotGmres + 6, nodes * sizeof(int));
    gw.resid = (double*)scratch(c, kSlotGmres + 7, nodes * sizeof(double));
    if (!gw.x || !gw.d32 || !gw.pf || !gw.V || !gw.z || !gw.conv || !gw.iters || !gw.resid)
        return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    hipEvent_t ev[6];
    hipEvent_t* marks = nullptr;
    if (c->timing) {
        for (auto& e : ev) e = new_event();
        marks = ev;
    }
    int rc = verb::launch_local_gmres(*g, psd, lb.w, gw, prm->tol, prm->restart, prm->max_cycles,
                                      prm->max_iterations, prm->inner_tol, x0, c->status, c->stream, marks);
    if (marks) {
        c->marks.push_back({VERB_PHASE_LOCAL_ASSEMBLE, ev[0], ev[1]});
        c->marks.push_back({VERB_PHASE_LOCAL_KRYLOV, ev[2], ev[3]});
        c->marks.push_back({VERB_PHASE_LOCAL_FACTOR, ev[4], ev[5]});
    }
    if (rc) return set_err(c, rc, "GMRES local diffusion launch rejected");
    if (stats) {
        std::vector<int> conv(nodes), it(nodes);
        std::vector<double> rs(nodes);
        HIPCHK(c, hipMemcpyAsync(conv.data(), gw.conv, nodes * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(it.data(), gw.iters, nodes * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(rs.data(), gw.resid, nodes * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        verb_gmres_stats st{0, 0, 0, 0, 0, 0.0, 0.0};
        double slog = 0.0;
        for (int i = 0; i < nodes; i++) {
            const int iR = i % g->nR;
            if (!g->include_boundary && (iR == 0 || iR == g->nR - 1)) continue;
            st.nodes++;
            if (conv[i]) st.converged++;
            else st.fallback++;
            st.iterations += it[i];
            if (it[i] > st.max_iterations) st.max_iterations = it[i];
            if (rs[i] > st.max_residual) st.max_residual = rs[i];
            slog += std::log10(rs[i] > 1e-300 ? rs[i] : 1e-300);
        }
        st.mean_log10_residual = st.nodes ? slog / st.nodes : 0.0;
        *stats = st;
    }
    return VERB_OK;
}

}  // namespace

extern "C" {

int verb_local_diffusion_gmres(verb_ctx* c, const verb_grid* g, double* psd, const verb_gmres_params* prm,
                               verb_gmres_stats* stats) {
    return gmres_local(c, g, psd, nullptr, prm, stats);
}

int verb_local_diffusion_gmres_x0(verb_ctx* c, const verb_grid* g, double* psd, const double* x0,
                                  const verb_gmres_params* prm, verb_gmres_stats* stats) {
    if (!x0) return VERB_ERR_ARG;
    return gmres_local(c, g, psd, x0, prm, stats);
}

constexpr int kSlotLearned = 100;  // scratch slots 100..104 for the warm-start workspace

int verb_local_diffusion_learned(verb_ctx* c, const verb_grid* g, double* psd, verb_unet* net,
                                 const verb_gmres_params* prm, verb_gmres_stats* stats) {
    if (!c || !grid_ok(g) || !psd || !net || !prm) return VERB_ERR_ARG;
    if (verb::unet_ctx(net) != c) return set_err(c, VERB_ERR_ARG, "the network belongs to another context");
    if (!g->DKK || !g->DVV || !g->DVK || !g->Vg) return VERB_ERR_ARG;
    const int nodes = g->nP * g->nR;
    const int Hp = (g->nV + 7) / 8 * 8, Wp = (g->nK + 7) / 8 * 8;
    const int chunk = std::min(nodes, verb::unet_max_batch(net));
    const size_t n = (size_t)g->nV * g->nK;
    uint16_t* feat;
    float *pin, *delta;
    double *scale, *x0;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        feat = (uint16_t*)scratch(c, kSlotLearned + 0, (size_t)chunk * Hp * Wp * 16 * sizeof(uint16_t));
        pin = (float*)scratch(c, kSlotLearned + 1, (size_t)chunk * Hp * Wp * sizeof(float));
        delta = (float*)scratch(c, kSlotLearned + 2, (size_t)chunk * Hp * Wp * sizeof(float));
        scale = (double*)scratch(c, kSlotLearned + 3, (size_t)chunk * sizeof(double));
        x0 = (double*)scratch(c, kSlotLearned + 4, n * nodes * sizeof(double));
        if (!feat || !pin || !delta || !scale || !x0) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
        const int m = mark_begin(c, VERB_PHASE_LOCAL_WARMSTART);
        for (int node0 = 0; node0 < nodes; node0 += chunk) {
            const int cnt = std::min(chunk, nodes - node0);
            int rc = verb::unet_warm_start(net, *g, psd, node0, cnt, feat, pin, scale, delta, x0, c->stream);
            if (rc) return set_err(c, rc, "warm-start network launch rejected");
        }
        mark_end(c, m);
    }
    return gmres_local(c, g, psd, x0, prm, stats);
}

int verb_step_learned(verb_ctx* c, const verb_grid* g, double* psd, verb_unet* net, const verb_gmres_params* prm,
                      verb_gmres_stats* stats) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    int rc = VERB_OK;
    if (3 < g->nP || 3 < g->nR) rc = verb_convection_batched(c, g, psd);
    if (!rc && g->nR >= 3) rc = verb_radial_diffusion_batched(c, g, psd);
    if (!rc && g->nV >= 3 && g->nK >= 3) rc = verb_local_diffusion_learned(c, g, psd, net, prm, stats);
    return rc;
}

constexpr int kSlotMixed = 80;  // scratch slots 80..90 for the mixed-precision workspace

int verb_local_diffusion_mixed(verb_ctx* c, const verb_grid* g, double* psd, const verb_mixed_params* prm,
                               verb_mixed_stats* stats) {
    if (!c || !grid_ok(g) || !psd || !prm) return VERB_ERR_ARG;
    if (!(g->sub_dt > 0) || (int)(g->dt / g->sub_dt) > 1)
        return set_err(c, VERB_ERR_ARG, "mixed-precision path: one sub-step only");
    if (prm->max_sweeps < 1 || prm->max_sweeps > 64 || !(prm->tol > 0))
        return set_err(c, VERB_ERR_ARG, "mixed-precision parameters out of range");
    if (g->nV + 1 > 152) return set_err(c, VERB_ERR_ARG, "mixed-precision path: nV <= 151");
    const size_t n = (size_t)g->nV * g->nK;
    const int nodes = g->nP * g->nR;
    const size_t ldab = 3 * (size_t)(g->nV + 1) + 1;
    std::lock_guard<std::mutex> lk(c->mu);
    // the FP64 operator, B0, C0 and b live in the shared local workspace; the
    // FP64 band is not needed here, so its slot is not requested
    verb::LocalWork w{};
    w.b0 = (double*)scratch(c, kSlotLocal + 2, n * nodes * sizeof(double));
    w.c0 = (double*)scratch(c, kSlotLocal + 3, n * nodes * sizeof(double));
    w.rhs = (double*)scratch(c, kSlotLocal + 4, n * nodes * sizeof(double));
    w.diag = (double*)scratch(c, kSlotLocal + 5, 9 * n * nodes * sizeof(double));
    verb::MixedWork mw;
    mw.d32 = (float*)scratch(c, kSlotMixed + 0, 9 * n * nodes * sizeof(float));
    mw.ab32 = (float*)scratch(c, kSlotMixed + 1, ldab * n * nodes * sizeof(float));
    mw.ipiv = (int64_t*)scratch(c, kSlotMixed + 2, n * nodes * sizeof(int64_t));
    mw.x = (double*)scratch(c, kSlotMixed + 3, n * nodes * sizeof(double));
    mw.r32 = (float*)scratch(c, kSlotMixed + 4, n * nodes * sizeof(float));
    mw.y32 = (float*)scratch(c, kSlotMixed + 5, n * nodes * sizeof(float));
    mw.c32 = (float*)scratch(c, kSlotMixed + 6, n * nodes * sizeof(float));
    mw.bmax = (double*)scratch(c, kSlotMixed + 7, nodes * sizeof(double));
    mw.rres = (double*)scratch(c, kSlotMixed + 8, nodes * sizeof(double));
    mw.done = (int*)scratch(c, kSlotMixed + 9, nodes * sizeof(int));
    mw.sweeps = (int*)scratch(c, kSlotMixed + 10, nodes * sizeof(int));
    if (!w.b0 || !w.c0 || !w.rhs || !w.diag || !mw.d32 || !mw.ab32 || !mw.ipiv || !mw.x || !mw.r32 || !mw.y32 ||
        !mw.c32 || !mw.bmax || !mw.rres || !mw.done || !mw.sweeps)
        return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    hipEvent_t ev[6];
    hipEvent_t* marks = nullptr;
    if (c->timing) {
        for (auto& e : ev) e = new_event();
        marks = ev;
    }
    int rc = verb::launch_local_mixed(*g, psd, w, mw, prm->tol, prm->max_sweeps, c->status, c->stream, marks);
    if (marks) {
        c->marks.push_back({VERB_PHASE_LOCAL_ASSEMBLE, ev[0], ev[1]});
        c->marks.push_back({VERB_PHASE_LOCAL_FACTOR, ev[2], ev[3]});
        c->marks.push_back({VERB_PHASE_LOCAL_SOLVE, ev[4], ev[5]});
    }
    if (rc) return set_err(c, rc, "mixed-precision local diffusion launch rejected");
    if (stats) {
        std::vector<int> done(nodes), sw(nodes);
        std::vector<double> rr(nodes);
        HIPCHK(c, hipMemcpyAsync(done.data(), mw.done, nodes * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(sw.data(), mw.sweeps, nodes * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(rr.data(), mw.rres, nodes * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        verb_mixed_stats st{0, 0, 0, 0.0, 0.0};
        long total = 0;
        for (int i = 0; i < nodes; i++) {
            const int iR = i % g->nR;
            if (!g->include_boundary && (iR == 0 || iR == g->nR - 1)) continue;
            st.nodes++;
            if (done[i]) st.converged++;
            total += sw[i];
            if (sw[i] > st.max_sweeps_used) st.max_sweeps_used = sw[i];
            if (rr[i] > st.max_residual) st.max_residual = rr[i];
        }
        st.mean_sweeps = st.nodes ? (double)total / st.nodes : 0.0;
        *stats = st;
    }
    return VERB_OK;
}

int verb_step_mixed(verb_ctx* c, const verb_grid* g, double* psd, const verb_mixed_params* prm,
                    verb_mixed_stats* stats) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    int rc = VERB_OK;
    if (3 < g->nP || 3 < g->nR) rc = verb_convection_batched(c, g, psd);
    if (!rc && g->nR >= 3) rc = verb_radial_diffusion_batched(c, g, psd);
    if (!rc && g->nV >= 3 && g->nK >= 3) rc = verb_local_diffusion_mixed(c, g, psd, prm, stats);
    return rc;
}

int verb_step_gmres(verb_ctx* c, const verb_grid* g, double* psd, const verb_gmres_params* prm,
                    verb_gmres_stats* stats) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    int rc = VERB_OK;
    if (3 < g->nP || 3 < g->nR) rc = verb_convection_batched(c, g, psd);
    if (!rc && g->nR >= 3) rc = verb_radial_diffusion_batched(c, g, psd);
    if (!rc && g->nV >= 3 && g->nK >= 3) rc = verb_local_diffusion_gmres(c, g, psd, prm, stats);
    return rc;
}

// ---------------------------------------------------------------------------
// Launch shapes, autotuning and the captured step
// ---------------------------------------------------------------------------
int verb_set_tuning(verb_ctx* c, const verb_tuning* t) {
    if (!c || !t) return VERB_ERR_ARG;
    int variant = -1;
    if (t->lu_rb > 0 || t->lu_cb > 0) {
        for (int i = 0; i < verb::lu_variant_count(); i++) {
            int rb = 0, cb = 0;
            verb::lu_variant_shape(i, 0, &rb, &cb);
            if (rb == t->lu_rb && cb == t->lu_cb) variant = i;
        }
        if (variant < 0) return set_err(c, VERB_ERR_ARG, "no register-window LU of that shape");
    }
    if (t->convection_tq < 0 || t->convection_tq > 64) return set_err(c, VERB_ERR_ARG, "convection_tq out of range");
    if (t->convection_kernel < 0 || t->convection_kernel > 1)
        return set_err(c, VERB_ERR_ARG, "convection_kernel must be 0 or 1");
    std::lock_guard<std::mutex> lk(c->mu);
    c->conv_tq = t->convection_tq;
    c->conv_kernel = t->convection_kernel;
    c->lu_variant = variant;
    c->lu_persistent = t->lu_persistent ? 1 : 0;
    return VERB_OK;
}

int verb_get_tuning(verb_ctx* c, verb_tuning* t) {
    if (!c || !t) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    t->convection_tq = c->conv_tq;
    t->convection_kernel = c->conv_kernel;
    t->lu_rb = t->lu_cb = 0;
    if (c->lu_variant >= 0) verb::lu_variant_shape(c->lu_variant, 0, &t->lu_rb, &t->lu_cb);
    t->lu_persistent = c->lu_persistent;
    t->convection_ms = t->local_ms = 0.0;
    return VERB_OK;
}

}  // extern "C"

namespace {

// Milliseconds of one call of fn on the ctx stream, psd restored from `save`
// first (the restore is not timed). Negative when fn rejects the launch.
template <class F>
double time_once(verb_ctx* c, double* psd, const double* save, size_t bytes, F fn) {
    if (hipMemcpyAsync(psd, save, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) return -1.0;
    hipEvent_t a = new_event(), b = new_event();
    hipEventRecord(a, c->stream);
    const int rc = fn();
    hipEventRecord(b, c->stream);
    hipEventSynchronize(b);
    float ms = -1.0f;
    if (rc == VERB_OK) hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return rc == VERB_OK ? (double)ms : -1.0;
}

}  // namespace

extern "C" {

int verb_autotune(verb_ctx* c, const verb_grid* g, double* psd, verb_tuning* out) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    const size_t bytes = (size_t)g->nP * g->nR * g->nV * g->nK * sizeof(double);
    double* save = nullptr;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        save = (double*)scratch(c, kSlotTune, bytes);
        if (!save) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
        HIPCHK(c, hipMemcpyAsync(save, psd, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    const bool saved_timing = c->timing;
    c->timing = false;
    verb_tuning best{};
    // convection: planes per workgroup
    best.convection_ms = -1.0;
    if (3 < g->nP || 3 < g->nR) {
        // the line kernel (where it is built for this grid; otherwise these
        // trials run the tile kernel too) and the tile kernel
        for (int kern = 0; kern <= 1; kern++) {
            for (int tq : {1, 2, 4, 8, 16}) {
                if (kern == 1 && tq > 8) continue;
                c->conv_kernel = kern;
                c->conv_tq = tq;
                time_once(c, psd, save, bytes, [&] { return verb_convection_batched(c, g, psd); });  // warm-up
                const double ms = time_once(c, psd, save, bytes, [&] { return verb_convection_batched(c, g, psd); });
                if (ms >= 0.0 && (best.convection_ms < 0.0 || ms < best.convection_ms)) {
                    best.convection_ms = ms;
                    best.convection_tq = tq;
                    best.convection_kernel = kern;
                }
            }
        }
    }
    c->conv_tq = best.convection_tq;
    c->conv_kernel = best.convection_kernel;
    // local diffusion: register-window shape and persistent dispatch
    best.local_ms = -1.0;
    int best_variant = -1, best_pers = 0;
    const int kl = g->nV + 1;
    if (g->nV >= 3 && g->nK >= 3) {
        for (int v = 0; v < verb::lu_variant_count(); v++) {
            if (!verb::lu_variant_shape(v, kl, nullptr, nullptr)) continue;
            for (int pers = 0; pers <= 1; pers++) {
                c->lu_variant = v;
                c->lu_persistent = pers;
                auto run = [&] { return verb_local_diffusion_batched(c, g, psd, nullptr, nullptr, 1); };
                time_once(c, psd, save, bytes, run);  // warm-up (workspace allocation)
                const double ms = time_once(c, psd, save, bytes, run);
                if (ms >= 0.0 && (best.local_ms < 0.0 || ms < best.local_ms)) {
                    best.local_ms = ms;
                    best_variant = v;
                    best_pers = pers;
                }
            }
        }
    }
    c->lu_variant = best_variant;
    c->lu_persistent = best_pers;
    if (best_variant >= 0) verb::lu_variant_shape(best_variant, kl, &best.lu_rb, &best.lu_cb);
    best.lu_persistent = best_pers;
    c->timing = saved_timing;
    // the trials may have raised device status (e.g. a zero pivot): report it
    int rc = verb_check_status(c);
    HIPCHK(c, hipMemcpyAsync(psd, save, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (out) *out = best;
    return rc;
}

int verb_step_graph(verb_ctx* c, const verb_grid* g, double* psd, const int* k0) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    const int tune[4] = {c->conv_tq, c->lu_variant, c->lu_persistent, c->conv_kernel};
    std::vector<int> key;
    if (k0) key.assign(k0, k0 + g->nP * g->nR);
    // the captured kernels hold scratch addresses and read the extent tables:
    // any reallocation or re-upload since the capture forces a new one
    const bool match = c->graph_exec && std::memcmp(&c->graph_grid, g, sizeof(verb_grid)) == 0 &&
                       c->graph_psd == psd && std::memcmp(c->graph_tune, tune, sizeof(tune)) == 0 &&
                       c->graph_k0 == key && c->graph_gen == c->scratch_gen;
    auto run = [&] {
        return k0 ? verb_step_uneven(c, g, psd, k0, 1, 1, 1) : verb_step(c, g, psd, 1, 1, 1, nullptr, nullptr, 1);
    };
    if (match) {
        HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
        return VERB_OK;
    }
    // A new shape: this call runs the step eagerly (which also allocates the
    // workspace outside any capture), then records the same launch sequence
    // into a graph for the following calls.
    int rc = run();
    if (rc) return rc;
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    const bool saved_timing = c->timing;
    c->timing = false;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = run();
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &gr);
    c->timing = saved_timing;
    if (rc || e != hipSuccess || !gr) {
        if (gr) hipGraphDestroy(gr);
        return rc ? rc : set_err(c, VERB_ERR_HIP, "step capture failed");
    }
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) {
        hipGraphDestroy(gr);
        return set_err(c, VERB_ERR_HIP, "step graph instantiation failed");
    }
    c->graph = gr;
    c->graph_exec = ex;
    c->graph_grid = *g;
    c->graph_psd = psd;
    std::memcpy(c->graph_tune, tune, sizeof(tune));
    c->graph_k0 = key;
    c->graph_gen = c->scratch_gen;
    return VERB_OK;
}

int verb_step(verb_ctx* c, const verb_grid* g, double* psd, int run_convection, int run_radial, int run_local,
              double* lu_cache, int64_t* ipiv_cache, int recompute_lu) {
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    int rc = VERB_OK;
    if (run_convection && (3 < g->nP || 3 < g->nR)) rc = verb_convection_batched(c, g, psd);
    if (!rc && run_radial && g->nR >= 3) rc = verb_radial_diffusion_batched(c, g, psd);
    if (!rc && run_local && g->nV >= 3 && g->nK >= 3)
        rc = verb_local_diffusion_batched(c, g, psd, lu_cache, ipiv_cache, recompute_lu);
    return rc;
}

constexpr int kSlotDriver = 95;  // negative-point counter
constexpr int kSlotStability = 96;  // per-block minima and the result

int verb_field_update(verb_ctx* c, double* dst, const double* original, size_t n,
                      const verb_field_update_params* prm) {
    if (!c || !dst || !original || !prm) return VERB_ERR_ARG;
    bool limited = false;
    for (int a = 0; a < 4; a++) limited = limited || prm->grid[a];
    if (dst == original && (prm->scale != 1.0 || limited))
        return set_err(c, VERB_ERR_ARG, "field update: dst must differ from the original array");
    if (n == 0) return VERB_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = verb::launch_field_update(dst, original, n, *prm, c->stream);
    return rc ? set_err(c, rc, "field update launch failed") : VERB_OK;
}

int verb_check_stability(verb_ctx* c, const verb_grid* g, double* min_value) {
    if (!c || !grid_ok(g) || !g->DVV || !g->DKK || !g->DVK) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    double* part = (double*)scratch(c, kSlotStability, (verb::kStabilityParts + 1) * sizeof(double));
    if (!part) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    int rc = verb::launch_stability_min(*g, part, part + verb::kStabilityParts, c->stream);
    if (rc) return set_err(c, rc, "stability check launch failed");
    double m = 0.0;
    HIPCHK(c, hipMemcpyAsync(&m, part + verb::kStabilityParts, sizeof(m), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (min_value) *min_value = m;
    if (m < 0) return set_err(c, VERB_ERR_UNSTABLE, "(DVV*DKK - DVK*DVK) < 0: the computation would be unstable");
    return VERB_OK;
}

int verb_remap_batched(verb_ctx* c, double* psd, const double* lstar_old, const double* lstar_new, int nP, int nR,
                       int nV, int nK) {
    if (!c || !psd || !lstar_old || !lstar_new || nP < 0 || nR < 0 || nV < 0 || nK < 0) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = verb::launch_remap(psd, lstar_old, lstar_new, nP, nR, nV, nK, c->status, c->stream);
    return rc ? set_err(c, rc, "remap launch rejected") : VERB_OK;
}

int verb_da_radial(verb_ctx* c, const verb_grid* g, double* psd, const double* obs, double* covariance,
                   double* model_matrix, const verb_da_params* p) {
    if (!c || !grid_ok(g) || !psd || !obs || !covariance || !p || !g->Lg || !g->DLL || !g->G_radial)
        return VERB_ERR_ARG;
    if (!(p->dt > 0) || p->p_index < 0 || p->p_index >= g->nP) return VERB_ERR_ARG;
    if (!bc_ok(g->Ll_type) || !bc_ok(g->Lu_type) || g->Ll_type == VERB_BC_PERIODIC || g->Lu_type == VERB_BC_PERIODIC)
        return VERB_ERR_BC;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = verb::launch_da_radial(*g, psd, obs, covariance, model_matrix, p->p_index, p->dt, p->model_error,
                                    p->observation_error, c->status, c->stream);
    return rc ? set_err(c, rc, "data assimilation launch rejected (3 <= nR <= 32)") : VERB_OK;
}

int verb_da_convection(verb_ctx* c, const verb_grid* g, double* psd, const double* obs, const double* loss,
                       const double* saturation, double* covariance, double* model, const verb_da_conv_params* p) {
    if (!c || !grid_ok(g) || !psd || !obs || !loss || !covariance || !model || !p || !g->Pg || !g->Rg || !g->VP ||
        !g->VR)
        return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = verb::run_da_convection(c, &c->blas, c->stream, c->status, *g, psd, obs, loss, saturation, covariance,
                                     model, *p);
    return rc == VERB_ERR_ARG ? set_err(c, rc, "data assimilation arguments rejected") : rc;
}

int verb_cubic1d(verb_ctx* c, int n, const double* x_old, double* values, const double* x_new) {
    if (!c || n < 0 || (n > 0 && (!x_old || !values || !x_new))) return VERB_ERR_ARG;
    if (n == 0) return VERB_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = 0;
    const double* dxo = up(c, 1, x_old, n, rc);
    const double* dxn = up(c, 2, x_new, n, rc);
    double* dv = (double*)scratch(c, 0, n * sizeof(double));
    if (rc) return rc;
    if (!dv) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    HIPCHK(c, hipMemcpyAsync(dv, values, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = reset_status(c))) return rc;
    rc = verb::launch_remap(dv, dxo, dxn, 1, n, 1, 1, c->status, c->stream);
    if (rc) return set_err(c, rc, "remap launch rejected");
    rc = finish(c);
    reset_status(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(values, dv, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VERB_OK;
}

int verb_step_ex(verb_ctx* c, const verb_grid* g, double* psd, int run_convection, int run_radial, int run_local,
                 double* lu_cache, int64_t* ipiv_cache, int recompute_lu, const verb_step_options* opt) {
    if (!opt) return verb_step(c, g, psd, run_convection, run_radial, run_local, lu_cache, ipiv_cache, recompute_lu);
    if (!c || !grid_ok(g) || !psd) return VERB_ERR_ARG;
    const size_t n = (size_t)g->nP * g->nR * g->nV * g->nK;
    constexpr double kFloor = 1e-21;
    int rc = VERB_OK;
    if (opt->positive_psd) rc = verb::launch_clamp_floor(psd, n, kFloor, c->stream);
    verb_grid gl = *g;
    if (!rc && run_convection && (3 < g->nP || 3 < g->nR)) {
        const verb_convection_options co{opt->psd_lost, opt->fast_convection};
        rc = verb_convection_batched_ex(c, g, psd, &co);
        if (!rc && opt->vl_from_convection && g->Vl_type == VERB_BC_CONSTANT_VALUE) {
            rc = verb::launch_vl_from_psd(*g, psd, opt->vl_from_convection, c->stream);
            gl.Vl = opt->vl_from_convection;
        }
        if (!rc && opt->vu_from_convection && g->Vu_type == VERB_BC_CONSTANT_VALUE) {
            rc = verb::launch_v_slice(*g, psd, g->nV - 1, opt->vu_from_convection, c->stream);
            gl.Vu = opt->vu_from_convection;
        }
    }
    if (!rc && opt->positive_psd) rc = verb::launch_clamp_floor(psd, n, kFloor, c->stream);
    if (!rc && run_radial && g->nR >= 3) rc = verb_radial_diffusion_batched(c, &gl, psd);
    if (!rc && opt->positive_psd) rc = verb::launch_clamp_floor(psd, n, kFloor, c->stream);
    if (!rc && run_local && g->nV >= 3 && g->nK >= 3)
        rc = verb_local_diffusion_batched(c, &gl, psd, lu_cache, ipiv_cache, recompute_lu);
    if (rc) return rc;
    if (opt->positive_psd) {
        // every value is >= 1e-21 after the final clamp: no negative points
        // (src/VERB4D_Solver.cpp:1218-1221)
        if (opt->negative_points) *opt->negative_points = 0;
        return verb::launch_clamp_floor(psd, n, kFloor, c->stream);
    }
    if (opt->negative_points) {
        auto* cnt = (unsigned long long*)scratch(c, kSlotDriver, sizeof(unsigned long long));
        if (!cnt) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
        rc = verb::launch_count_negative(psd, n, cnt, c->stream);
        if (rc) return set_err(c, rc, "negative count launch failed");
        unsigned long long h = 0;
        HIPCHK(c, hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *opt->negative_points = (long long)h;
    }
    return VERB_OK;
}

// ---------------------------------------------------------------------------
// Per-call drop-in entries
// ---------------------------------------------------------------------------
int verb_diffusion_2d(verb_ctx* c, double* psd, const double* x, const double* y, int nx, int ny,
                      const double* xl, const double* xu, const double* yl, const double* yu, int xl_type,
                      int xu_type, int yl_type, int yu_type, const double* Dxx, const double* Dyy,
                      const double* Dxy, const double* Dyx, const double* G, const double* sources,
                      const double* losses, double dt, double sub_dt, double* lu_cache, int64_t* ipiv_cache,
                      int recompute_lu) {
    if (!c || !psd || nx < 3 || ny < 3 || !(sub_dt > 0)) return VERB_ERR_ARG;
    if (!bc_ok(xl_type) || !bc_ok(xu_type) || !bc_ok(yl_type) || !bc_ok(yu_type)) return VERB_ERR_BC;
    if (Dxy != Dyx) {
        // The device assembly reads one mixed coefficient for both mixed
        // terms, as the reference time loop passes DVK twice.
        const size_t n = (size_t)nx * ny;
        for (size_t i = 0; i < n; i++)
            if (Dxy[i] != Dyx[i]) return set_err(c, VERB_ERR_ARG, "Dxy != Dyx is not supported");
    }
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t n = (size_t)nx * ny;
    int rc = 0;
    verb_grid g;
    std::memset(&g, 0, sizeof(g));
    g.nP = 1; g.nR = 1; g.nV = nx; g.nK = ny;
    double* dpsd = (double*)scratch(c, 0, n * sizeof(double));
    if (!dpsd) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    HIPCHK(c, hipMemcpyAsync(dpsd, psd, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    g.Vg = up(c, 1, x, n, rc);
    g.Kg = up(c, 2, y, n, rc);
    g.Vl = up(c, 3, xl, ny, rc);
    g.Vu = up(c, 4, xu, ny, rc);
    g.Kl = up(c, 5, yl, nx, rc);
    g.Ku = up(c, 6, yu, nx, rc);
    g.DVV = up(c, 7, Dxx, n, rc);
    g.DKK = up(c, 8, Dyy, n, rc);
    g.DVK = up(c, 9, Dxy, n, rc);
    g.G_local = up(c, 10, G, n, rc);
    g.sources = up(c, 11, sources, n, rc);
    g.losses_local = up(c, 12, losses, n, rc);
    if (rc) return rc;
    g.Vl_type = xl_type; g.Vu_type = xu_type; g.Kl_type = yl_type; g.Ku_type = yu_type;
    g.dt = dt; g.sub_dt = sub_dt; g.local_sources = 1.0; g.include_boundary = 1;
    const int kl = nx + 1, ldab = 3 * kl + 1;
    double* dab = nullptr;
    int64_t* dpiv = nullptr;
    if (lu_cache) {
        dab = (double*)scratch(c, 13, (size_t)ldab * n * sizeof(double));
        dpiv = (int64_t*)scratch(c, 14, n * sizeof(int64_t));
        if (!dab || !dpiv) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
        if (!recompute_lu) {
            HIPCHK(c, hipMemcpyAsync(dab, lu_cache, (size_t)ldab * n * sizeof(double), hipMemcpyHostToDevice,
                                     c->stream));
            HIPCHK(c, hipMemcpyAsync(dpiv, ipiv_cache, n * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
        }
    }
    LocalBuffers lb = local_buffers(c, &g, dab, dpiv, 20);  // slots 20..27
    if (lb.rc) return lb.rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = reset_status(c))) return rc;
    const bool factor = (lu_cache == nullptr) || recompute_lu;
    rc = verb::launch_local(g, dpsd, lb.w, factor, lu_cache != nullptr, c->status, c->stream);
    if (rc) return set_err(c, rc, "local diffusion launch rejected");
    rc = finish(c);
    reset_status(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(psd, dpsd, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (lu_cache && recompute_lu) {
        HIPCHK(c, hipMemcpyAsync(lu_cache, lb.w.ab, (size_t)ldab * n * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(ipiv_cache, lb.w.ipiv, n * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VERB_OK;
}

int verb_diffusion_1d(verb_ctx* c, double* psd, const double* x, int n, double lbc, double ubc, int l_type,
                      int u_type, const double* Dxx, const double* G, const double* sources, const double* losses,
                      double dt) {
    if (!c || !psd || n < 3 || n > 64) return VERB_ERR_ARG;
    if (!bc_ok(l_type) || !bc_ok(u_type)) return VERB_ERR_BC;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = 0;
    verb_grid g;
    std::memset(&g, 0, sizeof(g));
    g.nP = 1; g.nR = n; g.nV = 1; g.nK = 1;
    double* dpsd = (double*)scratch(c, 0, n * sizeof(double));
    if (!dpsd) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    HIPCHK(c, hipMemcpyAsync(dpsd, psd, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    g.Lg = up(c, 1, x, n, rc);
    g.DLL = up(c, 2, Dxx, n, rc);
    g.G_radial = up(c, 3, G, n, rc);
    g.sources = up(c, 4, sources, n, rc);
    g.losses_radial = up(c, 5, losses, n, rc);
    g.Ll = up(c, 6, &lbc, 1, rc);
    g.Lu = up(c, 7, &ubc, 1, rc);
    if (rc) return rc;
    g.Ll_type = l_type; g.Lu_type = u_type; g.dt = dt; g.radial_sources = 1.0;
    if ((rc = reset_status(c))) return rc;
    rc = verb::launch_radial(g, dpsd, c->status, c->stream);
    if (rc) return set_err(c, rc, "radial launch rejected");
    rc = finish(c);
    reset_status(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(psd, dpsd, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VERB_OK;
}

static int convection_call(verb_ctx* c, double* psd, const double* P, const double* R, int nP, int nR,
                           const double* Pl, const double* Pu, const double* Rl, const double* Ru, int Pl_type,
                           int Pu_type, int Rl_type, int Ru_type, const double* VP, const double* VR,
                           const double* sources, const double* losses, const double* G, double dt,
                           int single_outer, double* lost = nullptr, int fast = 0) {
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t np = (size_t)nP * nR;
    int rc = 0;
    verb_grid g;
    std::memset(&g, 0, sizeof(g));
    g.nP = nP; g.nR = nR; g.nV = 1; g.nK = 1;
    double* dpsd = (double*)scratch(c, 0, np * sizeof(double));
    if (!dpsd) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    HIPCHK(c, hipMemcpyAsync(dpsd, psd, np * sizeof(double), hipMemcpyHostToDevice, c->stream));
    g.Pg = up(c, 1, P, np, rc);
    g.Rg = up(c, 2, R, np, rc);
    g.VP = up(c, 3, VP, np, rc);
    g.VR = up(c, 4, VR, np, rc);
    g.sources = up(c, 5, sources, np, rc);
    g.losses_conv = up(c, 6, losses, np, rc);
    g.G_conv = up(c, 7, G, np, rc);
    if (Pl_type != VERB_BC_PERIODIC) g.Pl = up(c, 8, Pl, nR, rc);
    if (Pu_type != VERB_BC_PERIODIC) g.Pu = up(c, 9, Pu, nR, rc);
    if (Rl_type != VERB_BC_PERIODIC) g.Rl = up(c, 10, Rl, nP, rc);
    if (Ru_type != VERB_BC_PERIODIC) g.Ru = up(c, 11, Ru, nP, rc);
    if (rc) return rc;
    g.Pl_type = Pl_type; g.Pu_type = Pu_type; g.Rl_type = Rl_type; g.Ru_type = Ru_type;
    g.dt = dt;
    double* dlost = nullptr;
    if (lost) {
        dlost = (double*)scratch(c, 12, np * sizeof(double));
        if (!dlost) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
        HIPCHK(c, hipMemcpyAsync(dlost, lost, np * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    if ((rc = reset_status(c))) return rc;
    rc = verb::launch_convection_mode(g, dpsd, c->status, c->stream, single_outer, 0, dlost, fast);
    if (rc) return set_err(c, rc, "convection launch rejected");
    rc = finish(c);
    reset_status(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(psd, dpsd, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (lost) HIPCHK(c, hipMemcpyAsync(lost, dlost, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VERB_OK;
}

int verb_convection_2d(verb_ctx* c, double* psd, const double* P, const double* R, int nP, int nR,
                       const double* Pl, const double* Pu, const double* Rl, const double* Ru, int Pl_type,
                       int Pu_type, int Rl_type, int Ru_type, const double* VP, const double* VR,
                       const double* sources, const double* losses, const double* G_conv, double dt) {
    if (!c || !psd || nP < 1 || nR < 1) return VERB_ERR_ARG;
    if (!bc_ok(Pl_type) || !bc_ok(Pu_type) || !bc_ok(Rl_type) || !bc_ok(Ru_type)) return VERB_ERR_BC;
    return convection_call(c, psd, P, R, nP, nR, Pl, Pu, Rl, Ru, Pl_type, Pu_type, Rl_type, Ru_type, VP, VR, sources,
                           losses, G_conv, dt, 0);
}

int verb_convection_2d_ex(verb_ctx* c, double* psd, const double* P, const double* R, int nP, int nR,
                          const double* Pl, const double* Pu, const double* Rl, const double* Ru, int Pl_type,
                          int Pu_type, int Rl_type, int Ru_type, const double* VP, const double* VR,
                          const double* sources, const double* losses, const double* G_conv, double dt,
                          const verb_convection_options* opt) {
    if (!c || !psd || nP < 1 || nR < 1) return VERB_ERR_ARG;
    if (!bc_ok(Pl_type) || !bc_ok(Pu_type) || !bc_ok(Rl_type) || !bc_ok(Ru_type)) return VERB_ERR_BC;
    return convection_call(c, psd, P, R, nP, nR, Pl, Pu, Rl, Ru, Pl_type, Pu_type, Rl_type, Ru_type, VP, VR, sources,
                           losses, G_conv, dt, 0, opt ? opt->psd_lost : nullptr, opt ? opt->fast_convection : 0);
}

int verb_convection_1d_quickest6(verb_ctx* c, double* psd, const double* x, int n, double lbc, double ubc,
                                 int l_type, int u_type, const double* Ux, double dt) {
    if (!c || !psd || n < 6) return VERB_ERR_ARG;
    if (!bc_ok(l_type) || !bc_ok(u_type)) return VERB_ERR_BC;
    // A single line is the R sweep of a one-row plane with G = 1, no losses,
    // no sources and one outer step: x*1, x/1, x*exp(0) and x + 0*dt leave
    // every value bit-identical, so only the QUICKEST6 update remains.
    std::vector<double> ones(n, 1.0), zeros(n, 0.0);
    return convection_call(c, psd, x, x, 1, n, nullptr, nullptr, &lbc, &ubc, VERB_BC_PERIODIC, VERB_BC_PERIODIC,
                           l_type, u_type, zeros.data(), Ux, zeros.data(), zeros.data(), ones.data(), dt, 1);
}

int verb_dgbtrf(verb_ctx* c, int64_t m, int64_t n, int64_t kl, int64_t ku, double* ab, int64_t ldab,
                int64_t* ipiv) {
    if (!c || !ab || !ipiv || m != n || n < 1 || kl < 0 || ku < 0 || ldab < 2 * kl + ku + 1 || n > (1 << 24))
        return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t bytes = (size_t)ldab * n * sizeof(double);
    double* dab = (double*)scratch(c, 30, bytes);
    int64_t* dpiv = (int64_t*)scratch(c, 31, (size_t)n * sizeof(int64_t));
    if (!dab || !dpiv) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    HIPCHK(c, hipMemcpyAsync(dab, ab, bytes, hipMemcpyHostToDevice, c->stream));
    int rc = reset_status(c);
    if (rc) return rc;
    rc = verb::launch_dgbtrf((int)n, (int)kl, (int)ku, dab, (int)ldab, dpiv, c->status, c->stream);
    if (rc) return set_err(c, rc, "dgbtrf launch rejected");
    int st[verb::kStatusWords];
    HIPCHK(c, hipMemcpyAsync(st, c->status, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ab, dab, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ipiv, dpiv, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    reset_status(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return st[1] == 0x7fffffff ? VERB_OK : st[1];
}

int verb_dgbtrs(verb_ctx* c, int64_t n, int64_t kl, int64_t ku, const double* ab, int64_t ldab,
                const int64_t* ipiv, double* b) {
    if (!c || !ab || !ipiv || !b || n < 1 || kl < 0 || ku < 0 || ldab < 2 * kl + ku + 1) return VERB_ERR_ARG;
    for (int64_t j = 0; j < n; j++)  // interchanges stay inside the band, as dgbtrf produces them
        if (ipiv[j] < j + 1 || ipiv[j] > std::min<int64_t>(n, j + 1 + kl)) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t bytes = (size_t)ldab * n * sizeof(double);
    double* dab = (double*)scratch(c, 30, bytes);
    int64_t* dpiv = (int64_t*)scratch(c, 31, (size_t)n * sizeof(int64_t));
    double* db = (double*)scratch(c, 32, (size_t)n * sizeof(double));
    if (!dab || !dpiv || !db) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    HIPCHK(c, hipMemcpyAsync(dab, ab, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dpiv, ipiv, (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(db, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = verb::launch_dgbtrs((int)n, (int)kl, (int)ku, dab, (int)ldab, dpiv, db, c->stream);
    if (rc) return set_err(c, rc, "dgbtrs launch rejected");
    HIPCHK(c, hipMemcpyAsync(b, db, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VERB_OK;
}

int verb_tridag_batched(verb_ctx* c, const double* a, const double* b, const double* cc, const double* r, double* u,
                        int n, int lines) {
    if (!c || !a || !b || !cc || !r || !u || n < 1 || n > 64 || lines < 0) return VERB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t cnt = (size_t)n * lines;
    int rc = 0;
    const double* da = up(c, 1, a, cnt, rc);
    const double* db = up(c, 2, b, cnt, rc);
    const double* dc = up(c, 3, cc, cnt, rc);
    const double* dr = up(c, 4, r, cnt, rc);
    if (rc) return rc;
    double* du = (double*)scratch(c, 5, cnt * sizeof(double));
    if (!du) return set_err(c, VERB_ERR_NOMEM, "device allocation failed");
    rc = verb::launch_tridag(da, db, dc, dr, du, n, lines, c->stream);
    if (rc) return set_err(c, rc, "tridag launch rejected");
    HIPCHK(c, hipMemcpyAsync(u, du, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return VERB_OK;
}

}  // extern "C"
