// Data assimilation in the (MLT, R) plane: the Kalman filter that
// DataAssimilationManagerConvection runs on every (V,K) plane
// (src/DataAssimilationConvection.cpp:42-238, runKalmanFilter
// src/DataAssimilation.cpp:18-73). The state is the plane's N = nP*nR
// points, so every plane carries dense N x N matrices:
//
//   M      the model matrix: the one-step upwind transport matrix of
//          getModelMatrixConvection2DNoStabilityCheck
//          (src/DataAssimilationHelper.cpp:25-102) for the Courant-split
//          sub-step, raised to the number of sub-steps (:104-148), kept
//          between calls and rebuilt when the velocities change
//   M1d    M^ceil(3/24/dt); Q = M1d diag(tmp) M1d^T + diag(D)/10
//   Pf     M Pa M^T + Q;  S = H Pf H^T + R;  K = Pf H^T S^-1
//   x     += K (obs - H x) (in log space with use_log);  Pa = Pf - K H Pf
//
// The dense products are FP64 GEMMs on rocBLAS (strided over the planes of
// a chunk); S is factored and solved by rocSOLVER (LU with partial
// pivoting, S padded to the chunk's largest observation count with the
// identity, which leaves the solution of the observed block unchanged).
// Matrix powers follow the reference's operator^ order
// (src/MatrixOperations.cpp:466-620): squares by repeated squaring, the
// squares of the set bits >= 1 multiplied in increasing order each from
// the left, then the base from the left for odd powers. The remaining
// steps (the transport matrix, column scaling by tmp, Q's diagonal,
// observation selection, the analysis) are kernels of this file.
#include <algorithm>
#include <cmath>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.h"

namespace verb {
namespace {

constexpr int kDcThreads = 256;

struct PlaneArgs {
    const double* Pg;
    const double* Rg;
    const double* VP;
    const double* VR;
    const double* loss;   // [P][R][V][K]
    const double* satd;   // saturation density [P][R][V][K] or null
    const double* psd;
    int nP, nR;
    long NQ;              // planes nV*nK
    long q0;              // first plane of the chunk
};

__device__ inline size_t at(const PlaneArgs& a, int p, int r, long q) {
    return ((size_t)p * a.nR + r) * a.NQ + q;
}
// (the grid steps dP and dR are those of the (V,K) = (0,0) plane: the
// reference's manager keeps P.yzSlice(0, 0) and R.yzSlice(0, 0),
// src/VERB4D_Solver.cpp:325-326)

// sub-step count per plane: splitTimeStepCourantCondition (:104-114)
__global__ __launch_bounds__(kDcThreads) void dc_substeps(PlaneArgs a, double dt, int* nt) {
    const long q = a.q0 + blockIdx.x;
    const double dP = a.Pg[at(a, 1, 0, 0)] - a.Pg[at(a, 0, 0, 0)];
    const double dR = a.Rg[at(a, 0, 1, 0)] - a.Rg[at(a, 0, 0, 0)];
    double mx = 0.0;
    for (int t = threadIdx.x; t < a.nP * a.nR; t += kDcThreads) {
        const int p = t / a.nR, r = t % a.nR;
        const double cp = a.VP[at(a, p, r, q)] * dt / dP;
        const double cr = a.VR[at(a, p, r, q)] * dt / dR;
        mx = fmax(mx, fabs(fabs(cp) + fabs(cr)));
    }
    __shared__ double red[kDcThreads];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kDcThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double c = red[0];
        nt[blockIdx.x] = c <= 1.0 ? 1 : (int)ceil(c / 1.0);
    }
}

// the one-step transport matrix of each plane of the chunk into M (N x N,
// row-major, zero off the stencil), getModelMatrixConvection2DNoStabilityCheck
__global__ __launch_bounds__(kDcThreads) void dc_step_matrix(PlaneArgs a, const int* nt, double dt, double corr_time,
                                                             int use_log, double* M, int* status) {
    const int c = blockIdx.x;
    const long q = a.q0 + c;
    const int nP = a.nP, nR = a.nR, N = nP * nR;
    double* Mq = M + (size_t)c * N * N;
    for (size_t t = threadIdx.x; t < (size_t)N * N; t += kDcThreads) Mq[t] = 0.0;
    __syncthreads();
    const double dts = dt / (double)nt[c];
    const double cf = exp(-dts / corr_time);
    const double dP = a.Pg[at(a, 1, 0, 0)] - a.Pg[at(a, 0, 0, 0)];
    const double dR = a.Rg[at(a, 0, 1, 0)] - a.Rg[at(a, 0, 0, 0)];
    for (int row = threadIdx.x; row < N; row += kDcThreads) {
        const int iP = row / nR, iR = row % nR;
        double* mr = Mq + (size_t)row * N;
        if (iR == 0 || iR == nR - 1) {
            mr[row] = cf;
            continue;
        }
        int iPl, iPr;
        if (iP == 0 || iP == nP - 1) { iPl = nP - 2; iPr = 1; }
        else { iPl = iP - 1; iPr = iP + 1; }
        const size_t g = at(a, iP, iR, q);
        const double vp = a.VP[g], vr = a.VR[g];
        // loss rate = Loss / PSD, zero where the PSD reaches saturation
        const bool sat = a.satd ? (a.psd[g] >= a.satd[g]) : false;
        double lc = a.loss[g] / a.psd[g];
        lc = sat ? 0.0 : lc;
        double mcc;
        if (use_log) mcc = 1. - fabs(vp) * dts / dP - fabs(vr) * dts / dR;
        else mcc = 1. - fabs(vp) * dts / dP - fabs(vr) * dts / dR + dts * lc;
        if (mcc < 0) raise_error(status, VERB_ERR_DOMAIN);  // the reference throws
        double mlc, mrc, mcl, mcr;
        if (vp > 0.) { mlc = vp * dts / dP; mrc = 0.; }
        else { mlc = 0.; mrc = -vp * dts / dP; }
        if (vr > 0.) { mcl = vr * dts / dR; mcr = 0.; }
        else { mcl = 0.; mcr = -vr * dts / dR; }
        mr[row] = mcc;
        mr[iR + iPl * nR] = mlc;
        mr[iR + iPr * nR] = mrc;
        mr[iR - 1 + iP * nR] = mcl;
        mr[iR + 1 + iP * nR] = mcr;
    }
}

// W = M1d diag(tmp): tmp = modelError, (1 - cf^2) modelError at each
// line's R ends
__global__ __launch_bounds__(kDcThreads) void dc_scale_cols(const double* M1d, double* W, int nP, int nR,
                                                            double model_error, double cf, size_t count) {
    const int N = nP * nR;
    for (size_t t = (size_t)blockIdx.x * kDcThreads + threadIdx.x; t < count; t += (size_t)gridDim.x * kDcThreads) {
        const int j = (int)(t % N);
        const int r = j % nR;
        const double d = (r == 0 || r == nR - 1) ? (1. - (cf * cf)) * model_error : model_error;
        W[t] = M1d[t] * d;
    }
}

// Q += diag(D)/10 (every element, as the reference adds the dense diagonal
// matrix), then the boundary overrides unless the boundary Q may grow
__global__ __launch_bounds__(kDcThreads) void dc_q_diag(double* Q, int nP, int nR, double model_error, double cf,
                                                        int growing, size_t count) {
    const int N = nP * nR;
    for (size_t t = (size_t)blockIdx.x * kDcThreads + threadIdx.x; t < count; t += (size_t)gridDim.x * kDcThreads) {
        const size_t e = t % ((size_t)N * N);
        const int i = (int)(e / N), j = (int)(e % N);
        double v = Q[t] + (i == j ? model_error / 10 : 0.0 / 10);
        if (!growing && i == j) {
            const int r = i % nR;
            if (r == 0 || r == nR - 1) v = (1. - (cf * cf)) * model_error;
        }
        Q[t] = v;
    }
}

// the analysis covariance starts at 5 modelError I where it is all zero
// (a covariance that was ever set has a nonzero diagonal: that is checked
// first, the whole matrix only when the diagonal is zero)
__global__ __launch_bounds__(kDcThreads) void dc_pa_init(double* Pa, int N, double model_error) {
    const int c = blockIdx.x;
    double* P = Pa + (size_t)c * N * N;
    __shared__ int nz;
    if (threadIdx.x == 0) nz = 0;
    __syncthreads();
    int mine = 0;
    for (int i = threadIdx.x; i < N; i += kDcThreads) mine |= (P[(size_t)i * N + i] != 0.0);
    if (mine) atomicOr(&nz, 1);
    __syncthreads();
    if (nz) return;
    for (size_t t = threadIdx.x; t < (size_t)N * N; t += kDcThreads) mine |= (P[t] != 0.0);
    if (mine) atomicOr(&nz, 1);
    __syncthreads();
    if (!nz)
        for (int i = threadIdx.x; i < N; i += kDcThreads) P[(size_t)i * N + i] = 5 * model_error;
}

// observations of each plane: the finite ones in (P,R) order (a block-wide
// compaction by wave ballots), their count, the forecast, the observed
// values and the innovation data - H x (log space with use_log)
__global__ __launch_bounds__(kDcThreads) void dc_observe(PlaneArgs a, const double* obs, int use_log, int* idx,
                                                         int* m, double* x, double* data, double* innov) {
    const int c = blockIdx.x;
    const long q = a.q0 + c;
    const int N = a.nP * a.nR;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int wcount[kDcThreads / 64];
    int base = 0;
    for (int t0 = 0; t0 < N; t0 += kDcThreads) {
        const int t = t0 + threadIdx.x;
        double o = 0.0, v = 0.0;
        bool f = false;
        if (t < N) {
            const size_t g = at(a, t / a.nR, t % a.nR, q);
            o = obs[g];
            v = a.psd[g];
            f = isfinite(o);
            x[(size_t)c * N + t] = use_log ? log(v) : v;
        }
        const unsigned long long bal = __ballot(f);
        if (lane == 0) wcount[wave] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; w++) off += wcount[w];
        int total = 0;
        for (int w = 0; w < kDcThreads / 64; w++) total += wcount[w];
        if (f) {
            const int k = off + __popcll(bal & ((1ull << lane) - 1));
            const double d = use_log ? log(o) : o;
            idx[(size_t)c * N + k] = t;
            data[(size_t)c * N + k] = d;
            innov[(size_t)c * N + k] = d - (use_log ? log(v) : v);
        }
        base += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) m[c] = base;
}

// S (mm x mm, padded with the identity), Pf H^T (N x mm, zero padded) and
// H Pf (mm x N, zero padded); R = diag(observationError) with use_log, else
// diag(obs * observationError). A plane without observations gets the
// reference's single all-zero pseudo-observation of value 0. Grid: planes x
// slices of the three outputs.
__global__ __launch_bounds__(kDcThreads) void dc_select(const double* Pf, const int* idx, const int* m,
                                                        const double* data, int N, int mm, int use_log,
                                                        double observation_error, double* S, double* PH,
                                                        double* HP) {
    const int c = blockIdx.x;
    const int mc = m[c];
    const double* P = Pf + (size_t)c * N * N;
    const int* id = idx + (size_t)c * N;
    const double* dv = data + (size_t)c * N;
    double* Sc = S + (size_t)c * mm * mm;
    double* PHc = PH + (size_t)c * N * mm;
    double* HPc = HP + (size_t)c * mm * N;
    const int s0 = blockIdx.y * kDcThreads + threadIdx.x, ss = gridDim.y * kDcThreads;
    for (int t = s0; t < mm * mm; t += ss) {
        const int i = t / mm, j = t % mm;
        double v;
        if (i < mc && j < mc) {
            v = P[(size_t)id[i] * N + id[j]];
            if (i == j) v += use_log ? observation_error : dv[i] * observation_error;
            else v += 0.0;
        } else if (mc == 0 && i == 0 && j == 0) {
            v = 0.0 + (use_log ? observation_error : 0.0 * observation_error);
        } else {
            v = (i == j) ? 1.0 : 0.0;
        }
        Sc[(size_t)i * mm + j] = v;
    }
    for (int t = s0; t < N * mm; t += ss) {
        const int i = t / mm, r = t % mm;
        PHc[t] = r < mc ? P[(size_t)i * N + id[r]] : 0.0;
    }
    for (int t = s0; t < mm * N; t += ss) {
        const int r = t / N, j = t % N;
        HPc[t] = r < mc ? P[(size_t)id[r] * N + j] : 0.0;
    }
}

// the identity in each plane's mm x mm block (the right-hand sides that
// give S^-1)
__global__ __launch_bounds__(kDcThreads) void dc_identity(double* I, int mm) {
    double* Ic = I + (size_t)blockIdx.x * mm * mm;
    for (int t = blockIdx.y * kDcThreads + threadIdx.x; t < mm * mm; t += gridDim.y * kDcThreads)
        Ic[t] = (t / mm == t % mm) ? 1.0 : 0.0;
}

// x += K (data - H x), a wavefront per row of K (row-major N x mm, plane
// stride kstride); NaN -> error (the reference throws); back from log
// space; the plane written back into the PSD
__global__ __launch_bounds__(kDcThreads) void dc_analysis(PlaneArgs a, double* psd, const double* x, const double* K,
                                                          size_t kstride, const int* m, const double* innov, int mm,
                                                          int use_log, int* status) {
    const int c = blockIdx.x;
    const long q = a.q0 + c;
    const int N = a.nP * a.nR;
    const int mc = m[c];
    const int lane = threadIdx.x & 63;
    const double* Kc = K + (size_t)c * kstride;
    const double* in = innov + (size_t)c * N;
    for (int i = blockIdx.y * (kDcThreads / 64) + (threadIdx.x >> 6); i < N; i += gridDim.y * (kDcThreads / 64)) {
        double corr = 0.0;
        for (int t = lane; t < mc; t += 64) corr = fma(Kc[(size_t)i * mm + t], in[t], corr);
        for (int o = 32; o > 0; o >>= 1) corr += __shfl_xor(corr, o);
        if (lane == 0) {
            if (mc == 0 && use_log) corr = 0.0 * (0.0 - log(0.0));  // log(0) - 0 times a zero gain: NaN
            double v = x[(size_t)c * N + i] + corr;
            if (isnan(v)) raise_error(status, VERB_ERR_DOMAIN);
            if (use_log) v = pow(exp(1.0), v);
            psd[at(a, i / a.nR, i % a.nR, q)] = v;
        }
    }
}

int grid_blocks(size_t count) { return (int)std::min<size_t>((count + kDcThreads - 1) / kDcThreads, 65535 * 4); }

struct Blas {
    rocblas_handle h = nullptr;
};

// row-major D (m x n) = alpha A (m x k) op(B) + beta C, op(B) = B (k x n)
// or B^T (B stored n x k), strided over `batch` matrices; D may be C
rocblas_status gemm_rm(rocblas_handle h, bool bT, int m, int n, int k, double alpha, const double* A, long sA,
                       const double* B, long sB, double beta, const double* C, long sC, double* D, long sD,
                       int batch) {
    // column-major: D^T (n x m) = op(B)^T A^T + beta C^T
    const rocblas_datatype f = rocblas_datatype_f64_r;
    return rocblas_gemm_strided_batched_ex(h, bT ? rocblas_operation_transpose : rocblas_operation_none,
                                           rocblas_operation_none, n, m, k, &alpha, B, f, bT ? k : n, sB, A, f, k, sA,
                                           &beta, C, f, n, sC, D, f, n, sD, batch, f, rocblas_gemm_algo_standard, 0,
                                           0);
}
rocblas_status gemm_rm(rocblas_handle h, bool bT, int m, int n, int k, double alpha, const double* A, long sA,
                       const double* B, long sB, double beta, double* C, long sC, int batch) {
    return gemm_rm(h, bT, m, n, k, alpha, A, sA, B, sB, beta, C, sC, C, sC, batch);
}

// res = A^p for `batch` planes (stride N*N); w0, w1: squares (ping-pong),
// w2: scratch for the product; res may not alias A, w0, w1, w2
rocblas_status power_rm(rocblas_handle h, hipStream_t st, const double* A, int p, int N, int batch, double* res,
                        double* w0, double* w1, double* w2) {
    const long s = (long)N * N;
    const size_t bytes = sizeof(double) * (size_t)s * batch;
    if (p == 1) return hipMemcpyAsync(res, A, bytes, hipMemcpyDeviceToDevice, st) == hipSuccess
                           ? rocblas_status_success : rocblas_status_internal_error;
    const double* cur = A;
    double* sq[2] = {w0, w1};
    int which = 0;
    bool have = false;
    rocblas_status rc = rocblas_status_success;
    for (int b = 1; (p >> b) > 0; b++) {
        double* nxt = sq[which];
        which ^= 1;
        rc = gemm_rm(h, false, N, N, N, 1.0, cur, s, cur, s, 0.0, nxt, s, batch);
        if (rc != rocblas_status_success) return rc;
        cur = nxt;
        if ((p >> b) & 1) {
            if (!have) {
                if (hipMemcpyAsync(res, cur, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
                    return rocblas_status_internal_error;
                have = true;
            } else {
                // res = cur res (the next square from the left)
                rc = gemm_rm(h, false, N, N, N, 1.0, cur, s, res, s, 0.0, w2, s, batch);
                if (rc != rocblas_status_success) return rc;
                if (hipMemcpyAsync(res, w2, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
                    return rocblas_status_internal_error;
            }
        }
    }
    if (p & 1) {
        rc = gemm_rm(h, false, N, N, N, 1.0, A, s, res, s, 0.0, w2, s, batch);
        if (rc != rocblas_status_success) return rc;
        if (hipMemcpyAsync(res, w2, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
            return rocblas_status_internal_error;
    }
    return rocblas_status_success;
}

}  // namespace

void da_conv_release(void** blas_slot) {
    Blas* b = (Blas*)*blas_slot;
    *blas_slot = nullptr;
    if (!b) return;
    if (b->h) rocblas_destroy_handle(b->h);
    delete b;
}

int run_da_convection(verb_ctx* c, void** blas_slot, hipStream_t st, int* status, const verb_grid& g, double* psd,
                      const double* obs, const double* loss, const double* satd, double* cov, double* model,
                      const verb_da_conv_params& p) {
    const int nP = g.nP, nR = g.nR, N = nP * nR;
    const long NQ = (long)g.nV * g.nK;
    if (nP < 3 || nR < 3 || !(p.dt > 0) || !(p.correlation_time != 0)) return VERB_ERR_ARG;
    const long qb = p.plane_begin, qn = p.plane_count > 0 ? p.plane_count : NQ - qb;
    if (qb < 0 || qb + qn > NQ || qn <= 0) return VERB_ERR_ARG;
    Blas* b = (Blas*)*blas_slot;
    if (!b) {
        b = new Blas;
        if (rocblas_create_handle(&b->h) != rocblas_status_success) {
            delete b;
            return ctx_error(c, VERB_ERR_HIP, "rocblas_create_handle failed");
        }
        *blas_slot = b;
    }
    rocblas_set_stream(b->h, st);
    rocblas_set_pointer_mode(b->h, rocblas_pointer_mode_host);
    const size_t NN = (size_t)N * N;
    // planes per chunk: four N x N work matrices per plane within ~4 GiB
    const long chunk = std::max<long>(1, std::min<long>(qn, (long)((size_t)4 << 30) / (long)(4 * NN * sizeof(double))));
    const size_t wbytes = NN * chunk * sizeof(double);
    double* W[4];
    for (int i = 0; i < 4; i++) {
        W[i] = (double*)ctx_scratch(c, 110 + i, wbytes);
        if (!W[i]) return ctx_error(c, VERB_ERR_NOMEM, "device allocation failed");
    }
    int* d_nt = (int*)ctx_scratch(c, 114, sizeof(int) * chunk);
    int* d_m = (int*)ctx_scratch(c, 115, sizeof(int) * chunk);
    int* d_idx = (int*)ctx_scratch(c, 116, sizeof(int) * N * chunk);
    double* d_x = (double*)ctx_scratch(c, 117, sizeof(double) * N * chunk);
    double* d_data = (double*)ctx_scratch(c, 118, sizeof(double) * N * chunk);
    double* d_innov = (double*)ctx_scratch(c, 125, sizeof(double) * N * chunk);
    if (!d_nt || !d_m || !d_idx || !d_x || !d_data || !d_innov)
        return ctx_error(c, VERB_ERR_NOMEM, "device allocation failed");
    const double cf = exp(-p.dt / p.correlation_time);
    const int p1 = (int)std::ceil(3. / 24 / p.dt);
    std::vector<int> nt(chunk), m(chunk);
    for (long q0 = qb; q0 < qb + qn; q0 += chunk) {
        const int C = (int)std::min<long>(chunk, qb + qn - q0);
        PlaneArgs a{g.Pg, g.Rg, g.VP, g.VR, loss, satd, psd, nP, nR, NQ, q0};
        double* Mq = model + (size_t)q0 * NN;
        double* Pa = cov + (size_t)q0 * NN;
        // the model matrices (sub-step counts read back to group equal ones)
        if (p.recompute_model) {
            hipLaunchKernelGGL(dc_substeps, dim3(C), dim3(kDcThreads), 0, st, a, p.dt, d_nt);
            if (hipMemcpyAsync(nt.data(), d_nt, sizeof(int) * C, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return ctx_error(c, VERB_ERR_HIP, "sub-step count read-back failed");
            hipLaunchKernelGGL(dc_step_matrix, dim3(C), dim3(kDcThreads), 0, st, a, d_nt, p.dt, p.correlation_time,
                               p.use_log, W[0], status);
            for (int i0 = 0; i0 < C;) {
                int i1 = i0 + 1;
                while (i1 < C && nt[i1] == nt[i0]) i1++;
                if (power_rm(b->h, st, W[0] + (size_t)i0 * NN, nt[i0], N, i1 - i0, Mq + (size_t)i0 * NN,
                             W[1] + (size_t)i0 * NN, W[2] + (size_t)i0 * NN, W[3] + (size_t)i0 * NN) !=
                    rocblas_status_success)
                    return ctx_error(c, VERB_ERR_HIP, "rocBLAS dgemm failed");
                i0 = i1;
            }
        }
        // M1d = M^p1 into W[3] (or M itself), Q = M1d diag(tmp) M1d^T +
        // diag(D)/10 into W[1] (through the caller's Q array when given:
        // built with the model matrix, read otherwise)
        double* Qc = p.process_noise ? p.process_noise + (size_t)q0 * NN : nullptr;
        if (!Qc || p.recompute_model) {
            double* Qdst = Qc ? Qc : W[1];
            const double* M1d = Mq;
            if (p1 != 1) {
                if (power_rm(b->h, st, Mq, p1, N, C, W[3], W[0], W[1], W[2]) != rocblas_status_success)
                    return ctx_error(c, VERB_ERR_HIP, "rocBLAS dgemm failed");
                M1d = W[3];
            }
            hipLaunchKernelGGL(dc_scale_cols, dim3(grid_blocks(NN * C)), dim3(kDcThreads), 0, st, M1d, W[0], nP, nR,
                               p.model_error, cf, NN * C);
            if (gemm_rm(b->h, true, N, N, N, 1.0, W[0], NN, M1d, NN, 0.0, Qdst, NN, C) != rocblas_status_success)
                return ctx_error(c, VERB_ERR_HIP, "rocBLAS dgemm failed");
            hipLaunchKernelGGL(dc_q_diag, dim3(grid_blocks(NN * C)), dim3(kDcThreads), 0, st, Qdst, nP, nR,
                               p.model_error, cf, p.growing_Q_at_boundary, NN * C);
        }
        // Pf = (M Pa) M^T + Q into W[1] (Q read from the caller's array or from W[1])
        hipLaunchKernelGGL(dc_pa_init, dim3(C), dim3(kDcThreads), 0, st, Pa, N, p.model_error);
        if (gemm_rm(b->h, false, N, N, N, 1.0, Mq, NN, Pa, NN, 0.0, W[2], NN, C) != rocblas_status_success ||
            gemm_rm(b->h, true, N, N, N, 1.0, W[2], NN, Mq, NN, 1.0, Qc ? Qc : W[1], NN, W[1], NN, C) !=
                rocblas_status_success)
            return ctx_error(c, VERB_ERR_HIP, "rocBLAS dgemm failed");
        // observations
        hipLaunchKernelGGL(dc_observe, dim3(C), dim3(kDcThreads), 0, st, a, obs, p.use_log, d_idx, d_m, d_x, d_data,
                           d_innov);
        if (hipMemcpyAsync(m.data(), d_m, sizeof(int) * C, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return ctx_error(c, VERB_ERR_HIP, "observation count read-back failed");
        const int mm = std::max(1, *std::max_element(m.begin(), m.begin() + C));
        double* S = (double*)ctx_scratch(c, 119, sizeof(double) * (size_t)mm * mm * C);
        double* PH = (double*)ctx_scratch(c, 120, sizeof(double) * (size_t)N * mm * C);
        double* HP = (double*)ctx_scratch(c, 121, sizeof(double) * (size_t)mm * N * C);
        rocblas_int* piv = (rocblas_int*)ctx_scratch(c, 122, sizeof(rocblas_int) * (size_t)mm * C);
        rocblas_int* info = (rocblas_int*)ctx_scratch(c, 123, sizeof(rocblas_int) * C);
        if (!S || !PH || !HP || !piv || !info) return ctx_error(c, VERB_ERR_NOMEM, "device allocation failed");
        const int ysel = (int)std::min<size_t>(256, ((size_t)mm * mm + 2 * (size_t)N * mm + 4095) / 4096);
        hipLaunchKernelGGL(dc_select, dim3(C, ysel), dim3(kDcThreads), 0, st, W[1], d_idx, d_m, d_data, N, mm,
                           p.use_log, p.observation_error, S, PH, HP);
        // K = Pf H^T S^-1 (trans_solve: dgesv on S^T with the rows of Pf H^T
        // as right-hand sides). S row-major is S^T column-major: its LU with
        // partial pivoting and the solve with the identity give S^-T
        // column-major = S^-1 row-major; then K = (Pf H^T) S^-1 by one GEMM
        // (row-major N x mm in W[0], plane stride N*N)
        double* Si = (double*)ctx_scratch(c, 124, sizeof(double) * (size_t)mm * mm * C);
        if (!Si) return ctx_error(c, VERB_ERR_NOMEM, "device allocation failed");
        hipLaunchKernelGGL(dc_identity, dim3(C, std::min(64, (mm * mm + 4095) / 4096)), dim3(kDcThreads), 0, st, Si,
                           mm);
        if (rocsolver_dgetrf_strided_batched(b->h, mm, mm, S, mm, (rocblas_stride)mm * mm, piv, mm, info, C) !=
            rocblas_status_success)
            return ctx_error(c, VERB_ERR_HIP, "rocSOLVER getrf failed");
        std::vector<rocblas_int> hinfo(C);
        if (hipMemcpyAsync(hinfo.data(), info, sizeof(rocblas_int) * C, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return ctx_error(c, VERB_ERR_HIP, "solve status read-back failed");
        for (int i = 0; i < C; i++)
            if (hinfo[i] != 0) return ctx_error(c, VERB_ERR_SINGULAR, "trans_solve: zero pivot");
        if (rocsolver_dgetrs_strided_batched(b->h, rocblas_operation_none, mm, mm, S, mm, (rocblas_stride)mm * mm, piv,
                                             mm, Si, mm, (rocblas_stride)mm * mm, C) != rocblas_status_success)
            return ctx_error(c, VERB_ERR_HIP, "rocSOLVER getrs failed");
        if (gemm_rm(b->h, false, N, mm, mm, 1.0, PH, (long)N * mm, Si, (long)mm * mm, 0.0, W[0], NN, C) !=
            rocblas_status_success)
            return ctx_error(c, VERB_ERR_HIP, "rocBLAS dgemm failed");
        hipLaunchKernelGGL(dc_analysis, dim3(C, (N + 15) / 16), dim3(kDcThreads), 0, st, a, psd, d_x, W[0], NN, d_m,
                           d_innov, mm, p.use_log, status);
        // Pa = Pf - K (H Pf)
        if (gemm_rm(b->h, false, N, N, mm, -1.0, W[0], NN, HP, (long)mm * N, 1.0, W[1], NN, Pa, NN, C) !=
            rocblas_status_success)
            return ctx_error(c, VERB_ERR_HIP, "rocBLAS dgemm failed");
    }
    return VERB_OK;
}

}  // namespace verb
