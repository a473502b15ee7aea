// Diffusion_2D for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
// verb_diffusion_2d_noalloc takes every device buffer it needs (the node's
// inputs staged from the host, the assembled operator, the band factors,
// B0, C0, right-hand sides, pivots, the status words) from one file-scope
// device array of fixed size instead of the context's pooled scratch, and
// runs the same kernels as verb_diffusion_2d.
//
// The array exists once per GPU for the life of the process. Its device
// address is taken with hipGetSymbolAddress on the first call on each
// device (with that device current) and kept per device; host code never
// passes the array's name to a kernel. Calls on one GPU share the array:
// a process-wide mutex per device serialises the host threads of this
// process, and calls that overlap on one GPU from elsewhere race.
#include <mutex>

#include "common.h"

namespace verb {
namespace {

constexpr size_t kNoAllocBytes = size_t(32) << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_dev_mu[kMaxDevices];
void* g_dev_ptr[kMaxDevices] = {nullptr};

// Bump allocation from the array: 256-byte aligned pieces.
struct Carve {
    unsigned char* base;
    size_t used = 0;
    template <typename T>
    T* take(size_t count) {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + count * sizeof(T);
        return (used <= kNoAllocBytes) ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

// Bytes a node of nx*ny unknowns takes (the same pieces as the carve below,
// each rounded up to 256 bytes).
size_t node_bytes(int nx, int ny, bool cached) {
    const size_t n = (size_t)nx * ny;
    const size_t ldab = 3 * (size_t)(nx + 1) + 1;
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t b = r(sizeof(int) * kStatusWords);
    b += 10 * r(n * sizeof(double));                        // psd, x, y, Dxx, Dyy, Dxy, G, S, L; rhs
    b += 2 * r(ny * sizeof(double)) + 2 * r(nx * sizeof(double));  // boundary vectors
    b += r(ldab * n * sizeof(double)) + r(n * sizeof(int64_t));    // band factors, pivots
    b += 4 * r(n * sizeof(double));                         // b0, c0, y, spare
    b += r(9 * n * sizeof(double)) + r(n * sizeof(int));    // diagonals, U extents
    (void)cached;  // a caller's cache uses the band and pivot pieces above
    return b;
}

}  // namespace
}  // namespace verb

extern "C" int verb_diffusion_2d_noalloc(verb_ctx* c, double* psd, const double* x, const double* y, int nx, int ny,
                                         const double* xl, const double* xu, const double* yl, const double* yu,
                                         int xl_type, int xu_type, int yl_type, int yu_type, const double* Dxx,
                                         const double* Dyy, const double* Dxy, const double* Dyx, const double* G,
                                         const double* sources, const double* losses, double dt, double sub_dt,
                                         double* lu_cache, int64_t* ipiv_cache, int recompute_lu) {
    using namespace verb;
    if (!c || !psd || nx < 3 || ny < 3 || !(sub_dt > 0)) return VERB_ERR_ARG;
    if ((lu_cache == nullptr) != (ipiv_cache == nullptr)) return VERB_ERR_ARG;
    for (int t : {xl_type, xu_type, yl_type, yu_type})
        if (t < 0 || t > 2) return VERB_ERR_BC;
    for (const double* p : {x, y, xl, xu, yl, yu, Dxx, Dyy, Dxy, Dyx, G, sources, losses})
        if (!p) return ctx_error(c, VERB_ERR_ARG, "null input array");
    const size_t n = (size_t)nx * ny;
    if (Dxy != Dyx) {
        for (size_t i = 0; i < n; i++)
            if (Dxy[i] != Dyx[i]) return ctx_error(c, VERB_ERR_ARG, "Dxy != Dyx is not supported");
    }
    if (node_bytes(nx, ny, lu_cache != nullptr) > kNoAllocBytes)
        return ctx_error(c, VERB_ERR_ARG,
                         "verb_diffusion_2d_noalloc: the node needs more than the 32 MiB static scratch");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return ctx_error(c, VERB_ERR_HIP, "no current HIP device");
    std::lock_guard<std::mutex> lk(g_dev_mu[dev]);
    if (!g_dev_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return ctx_error(c, VERB_ERR_HIP, "hipGetSymbolAddress of the static scratch failed");
        g_dev_ptr[dev] = p;
    }
    hipStream_t st = (hipStream_t)verb_ctx_stream(c);
    Carve cv{(unsigned char*)g_dev_ptr[dev]};
    int* status = cv.take<int>(kStatusWords);
    double* dpsd = cv.take<double>(n);
    auto up = [&](const double* h, size_t count) -> const double* {
        double* d = cv.take<double>(count);
        if (d && hipMemcpyAsync(d, h, count * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) return nullptr;
        return d;
    };
    verb_grid g{};
    g.nP = 1; g.nR = 1; g.nV = nx; g.nK = ny;
    g.Vg = up(x, n); g.Kg = up(y, n);
    g.Vl = up(xl, ny); g.Vu = up(xu, ny); g.Kl = up(yl, nx); g.Ku = up(yu, nx);
    g.DVV = up(Dxx, n); g.DKK = up(Dyy, n); g.DVK = up(Dxy, n);
    g.G_local = up(G, n); g.sources = up(sources, n); g.losses_local = up(losses, n);
    g.Vl_type = xl_type; g.Vu_type = xu_type; g.Kl_type = yl_type; g.Ku_type = yu_type;
    g.dt = dt; g.sub_dt = sub_dt; g.local_sources = 1.0; g.include_boundary = 1;
    const size_t ldab = 3 * (size_t)(nx + 1) + 1;
    LocalWork w{};
    w.ab = cv.take<double>(ldab * n);
    w.ipiv = cv.take<int64_t>(n);
    w.rhs = cv.take<double>(n);
    w.b0 = cv.take<double>(n);
    w.c0 = cv.take<double>(n);
    w.y = cv.take<double>(n);
    (void)cv.take<double>(n);
    w.diag = cv.take<double>(9 * n);
    w.uext = cv.take<int>(n);
    w.queue = nullptr;
    w.lu_variant = -1;
    if (!status || !dpsd || !g.Vg || !g.Kg || !g.Vl || !g.Vu || !g.Kl || !g.Ku || !g.DVV || !g.DKK || !g.DVK ||
        !g.G_local || !g.sources || !g.losses_local || !w.ab || !w.ipiv || !w.rhs || !w.b0 || !w.c0 || !w.y ||
        !w.diag || !w.uext)
        return ctx_error(c, VERB_ERR_HIP, "staging into the static scratch failed");
    if (hipMemcpyAsync(dpsd, psd, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return ctx_error(c, VERB_ERR_HIP, "host to device copy failed");
    if (lu_cache && !recompute_lu) {
        if (hipMemcpyAsync(w.ab, lu_cache, ldab * n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(w.ipiv, ipiv_cache, n * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess)
            return ctx_error(c, VERB_ERR_HIP, "host to device copy failed");
    }
    // status words start clear: no error, no zero pivot (the array keeps the
    // previous call's contents)
    const int init[kStatusWords] = {0, 0x7fffffff, -1, 0};
    if (hipMemcpyAsync(status, init, sizeof(init), hipMemcpyHostToDevice, st) != hipSuccess)
        return ctx_error(c, VERB_ERR_HIP, "host to device copy failed");
    const bool factor = (lu_cache == nullptr) || recompute_lu;
    int rc = launch_local(g, dpsd, w, factor, lu_cache != nullptr, status, st);
    if (rc) return ctx_error(c, rc, "local diffusion launch rejected");
    int hs[kStatusWords];
    if (hipMemcpyAsync(hs, status, sizeof(hs), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return ctx_error(c, VERB_ERR_HIP, "device to host copy failed");
    if (hs[0] == VERB_ERR_BC) return ctx_error(c, VERB_ERR_BC, "unknown boundary type");
    if (hs[0] != 0) return ctx_error(c, hs[0], "device error");
    if (hs[1] != 0x7fffffff) return ctx_error(c, hs[1], "Error computing LU decomposition (zero pivot)");
    if (hipMemcpyAsync(psd, dpsd, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
        return ctx_error(c, VERB_ERR_HIP, "device to host copy failed");
    if (lu_cache && recompute_lu) {
        if (hipMemcpyAsync(lu_cache, w.ab, ldab * n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(ipiv_cache, w.ipiv, n * sizeof(int64_t), hipMemcpyDeviceToHost, st) != hipSuccess)
            return ctx_error(c, VERB_ERR_HIP, "device to host copy failed");
    }
    if (hipStreamSynchronize(st) != hipSuccess) return ctx_error(c, VERB_ERR_HIP, "stream synchronisation failed");
    return VERB_OK;
}
