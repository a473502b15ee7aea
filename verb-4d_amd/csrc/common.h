// Shared device/host declarations of the MI355X VERB-4D step library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "verb4d_hip.h"

namespace verb {

// Device status word layout (int32 array owned by each ctx):
//   [0] first error code raised by a kernel (0 = none), e.g. VERB_ERR_COURANT
//   [1] smallest 1-based zero-pivot column seen by a band factorisation
//       (INT32_MAX = none)
//   [2] node index of that pivot
constexpr int kStatusWords = 4;

__device__ inline void raise_error(int* status, int code) { atomicCAS(&status[0], 0, code); }

__device__ inline void raise_pivot(int* status, int col1, int node) {
    int old = atomicMin(&status[1], col1);
    if (col1 < old) atomicExch(&status[2], node);
}

// Records `what` as the context's last error and returns code (api.hip).
int ctx_error(verb_ctx* c, int code, const char* what);

// Launchers (defined in the .hip files). All asynchronous on `stream`.
int launch_radial(const verb_grid& g, double* psd, int* status, hipStream_t stream);
// tq: (V,K) planes per workgroup (0: automatic); lost (nullable,
// [P][R][V][K]): SAVE_PSD_LOST_CONV accumulation; fast: FAST_CONVECTION outer
// sub-step rule; kernel: 0 = the line kernel where it is instantiated for
// (nP, nR), else the tile kernel; 1 = the tile kernel.
int launch_convection(const verb_grid& g, double* psd, int* status, hipStream_t stream, int tq = 0,
                      double* lost = nullptr, int fast = 0, int kernel = 0);
// single_outer: one outer step (a bare QUICKEST6 line; tile kernel)
int launch_convection_mode(const verb_grid& g, double* psd, int* status, hipStream_t stream, int single_outer,
                           int tq, double* lost = nullptr, int fast = 0, int kernel = 0);

struct LocalWork {
    double* ab;       // band factors, per node (2kl+ku+1)*n doubles (LAPACK layout)
    int64_t* ipiv;    // per node n pivots, 1-based
    double* b0;       // per node n: diagonal of B (energy-fast order)
    double* c0;       // per node n: diagonal of C
    double* rhs;      // per node n: right-hand side / solution (energy-fast order)
    double* diag;     // per node 9n: the assembled operator (register-window LU path)
    double* y;        // per node n: forward-eliminated right-hand side
    int* uext;        // per node n: last nonzero column of each U row
    int* queue = nullptr;  // device counter for the persistent LU launch (null: one workgroup per node)
    int lu_variant = -1;   // register-window instantiation (-1: smallest that covers kl)
};

template <typename T>
struct LuArgsT;
using LuArgs = LuArgsT<double>;
// FP32 factorisation (register window up to kl = 152)
int launch_band_lu32(const LuArgsT<float>& a, int nodes, hipStream_t stream);
int launch_band_lu(const LuArgs& a, int nodes, hipStream_t stream, int variant = -1);
// register-window instantiations: count, and (rb, cb) of entry i; returns
// whether entry i covers kl
int lu_variant_count();
int lu_variant_shape(int i, int kl, int* rb, int* cb);
int launch_backsolve(const double* ab, const double* y, const int* uext, double* x, int n, int kl, int ku, int ldab,
                     int nodes, int nR, int include_boundary, hipStream_t stream, const int* nn = nullptr,
                     const int* order = nullptr, const int* done = nullptr, int ublk_cols = 0,
                     size_t ublk_stride = 0);
// y = L^-1 P (B0*rhs + C0) (B0, C0 nullable: y = L^-1 P rhs) with LAPACK
// band factors (dgbtf2-order multipliers); kl <= 112.
int launch_forward(const double* ab, const int64_t* ipiv, int n, int kl, int ku, int ldab, const double* b0,
                   const double* c0, const double* rhs, double* y, int nodes, int nR, int include_boundary,
                   hipStream_t stream);
// FP32 solves with FP32 band factors (no B0/C0 premultiplication)
int launch_forward32(const float* ab, const int64_t* ipiv, int n, int kl, int ku, int ldab, const float* rhs,
                     float* y, int nodes, int nR, int include_boundary, hipStream_t stream,
                     const int* done = nullptr);
int launch_backsolve32(const float* ab, const float* y, float* x, int n, int kl, int ku, int ldab, int nodes, int nR,
                       int include_boundary, hipStream_t stream, const int* done = nullptr);
// 0 = automatic (register-window LU where an instantiation covers kl),
// 1 = force the one-column-per-step reference-order kernel.
extern int local_lu_variant;
// marks (nullable, 6 events): begin/end of the assembly, of the band factor
// and of the solve (+ scatter), recorded on `stream`.
// Uneven pitch-angle extents (device arrays, one entry per (P,R) node):
// k0 = first active K column, n = nV*(nK-k0) unknowns, order = nodes sorted
// by decreasing n (dispatch order of the factor and solve launches).
struct NodeSizes {
    const int* k0;
    const int* n;
    const int* order;
};
// cached: the factors live in a caller's cache (LAPACK layout incl. L).
int launch_local(const verb_grid& g, double* psd, const LocalWork& w, bool factor, bool cached, int* status,
                 hipStream_t stream, hipEvent_t* marks = nullptr, const NodeSizes* sizes = nullptr);

// Local diffusion by matrix-free GMRES (FP32 inner cycles, FP64 residual
// test), with the band LU for the nodes that miss the tolerance. Single
// sub-step only. conv/iters (device, one int per node) report the outcome.
struct GmresWork {
    double* x;     // [nodes][nfull]
    float* d32;    // [nodes][9][nfull]
    float* pf;     // [nodes][3][nfull]
    float* V;      // [nodes][25][nfull]
    float* z;      // [nodes][2][nfull]
    int* conv;     // [nodes]
    int* iters;    // [nodes]
    double* resid = nullptr;  // [nodes] relative FP64 residual at the last check
};
// x0 (nullable, [P][R][V][K] like psd): initial iterate; null starts from psd.
// max_iters > 0 caps the inner iterations of a node over all its cycles.
int launch_local_gmres(const verb_grid& g, double* psd, const LocalWork& w, const GmresWork& gw, double tol,
                       int restart, int max_cycles, int max_iters, double inner_tol, const double* x0, int* status,
                       hipStream_t stream, hipEvent_t* marks = nullptr);

// Local diffusion with an FP32 band LU and FP64 iterative refinement:
// x0 = 0; repeat { r = b - A x (FP64); stop the node when max|r| <= tol*max|b|;
// solve L U d = P r in FP32; x += d }. Single sub-step only.
struct MixedWork {
    float* d32;      // [nodes][9][n] FP32 operator
    float* ab32;     // [nodes][ldab][n] FP32 factors (L in dgbtf2 order)
    int64_t* ipiv;   // [nodes][n]
    double* x;       // [nodes][n] FP64 iterate (energy-fast)
    float* r32;      // [nodes][n] residual
    float* y32;      // [nodes][n] forward-eliminated residual
    float* c32;      // [nodes][n] correction
    double* bmax;    // [nodes] max|b|
    double* rres;    // [nodes] last max|r| / max|b|
    int* done;       // [nodes] 1 once converged (or skipped)
    int* sweeps;     // [nodes] solves used
};
int launch_local_mixed(const verb_grid& g, double* psd, const LocalWork& w, const MixedWork& mw, double tol,
                       int max_sweeps, int* status, hipStream_t stream, hipEvent_t* marks = nullptr);

// The assembled local operator of every node (the plain path's assembly):
// diag [nodes][9][n], b0/c0/rhs [nodes][n], energy-fast; rhs = B0*psd + C0.
int launch_local_assembly(const verb_grid& g, const double* psd, double* diag, double* b0, double* c0, double* rhs,
                          int* status, hipStream_t stream);

// Single-system LAPACK band routines on device buffers (one system).
int launch_dgbtrf(int n, int kl, int ku, double* ab, int ldab, int64_t* ipiv, int* status, hipStream_t stream);
int launch_dgbtrs(int n, int kl, int ku, const double* ab, int ldab, const int64_t* ipiv, double* b,
                  hipStream_t stream);
int launch_tridag(const double* a, const double* b, const double* c, const double* r, double* u, int n,
                  int lines, hipStream_t stream);

// Time-loop pieces around the operators (driver.hip).
int launch_clamp_floor(double* psd, size_t n, double floor, hipStream_t stream);
int launch_count_negative(const double* psd, size_t n, unsigned long long* count, hipStream_t stream);
int launch_vl_from_psd(const verb_grid& g, const double* psd, double* vl, hipStream_t stream);
int launch_v_slice(const verb_grid& g, const double* psd, int iv, double* v, hipStream_t stream);
int launch_field_update(double* dst, const double* src, size_t n, const verb_field_update_params& p,
                        hipStream_t stream);
constexpr int kStabilityParts = 1024;
int launch_stability_min(const verb_grid& g, double* part, double* out, hipStream_t stream);

// da.hip: the Kalman update along L on every (V,K) line of P slice iP (3 <= nR <= 32)
int launch_da_radial(const verb_grid& g, double* psd, const double* obs, double* Pa, double* F_out, int iP, double dt,
                     double model_error, double observation_error, int* status, hipStream_t stream);

// da_conv.hip: the (MLT, R) plane Kalman filter on rocBLAS / rocSOLVER
int run_da_convection(verb_ctx* c, void** blas_slot, hipStream_t st, int* status, const verb_grid& g, double* psd,
                      const double* obs, const double* loss, const double* satd, double* cov, double* model,
                      const verb_da_conv_params& p);
// destroys the handle that run_da_convection created in *blas_slot (if any)
// and clears the slot
void da_conv_release(void** blas_slot);
// a context's grow-only scratch slot (api.hip)
void* ctx_scratch(verb_ctx* c, int slot, size_t bytes);

// remap.hip: Cubic1D on every (P,V,K) line (nR <= 105)
int launch_remap(double* psd, const double* lold, const double* lnew, int nP, int nR, int nV, int nK, int* status,
                 hipStream_t stream);

}  // namespace verb
