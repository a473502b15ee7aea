// Pieces of the reference's time loop around the three operators
// (src/VERB4D_Solver.cpp:539-1282) that act on the whole device-resident
// grid: the positivity clamp PSD.max_of(1e-21) (:621, 980, 1072, 1218-1221;
// Matrix4D::max_of, src/Matrix.cpp:5948-5953), the count of negative points
// when no clamp is applied (:1223-1238), the V boundaries taken from the
// PSD after convection (:965-975, Vl_BC = PSD.ySlice(0), Vu_BC =
// PSD.ySlice(V_size-1)), the coefficient and boundary updates' scaling and
// limiting, and the stability check of the local diffusion coefficients.
#include "common.h"

namespace verb {

namespace {

constexpr int kDrvThreads = 256;

// std::max(v, floor) as Matrix4D::max_of computes it: (v < floor) ? floor : v
// (a NaN stays NaN)
__global__ __launch_bounds__(kDrvThreads) void clamp_floor(double* __restrict__ psd, size_t n, double floor) {
    const size_t stride = (size_t)gridDim.x * kDrvThreads;
    for (size_t i = (size_t)blockIdx.x * kDrvThreads + threadIdx.x; i < n; i += stride) {
        const double v = psd[i];
        psd[i] = (v < floor) ? floor : v;
    }
}

__global__ void zero_count(unsigned long long* c) { *c = 0ull; }

__global__ __launch_bounds__(kDrvThreads) void count_negative(const double* __restrict__ psd, size_t n,
                                                              unsigned long long* __restrict__ out) {
    const size_t stride = (size_t)gridDim.x * kDrvThreads;
    unsigned cnt = 0;
    for (size_t i = (size_t)blockIdx.x * kDrvThreads + threadIdx.x; i < n; i += stride) cnt += (psd[i] < 0.0) ? 1u : 0u;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

// v[p][r][k] = psd[p][r][iv][k] (iv = 0: Vl_BC, iv = nV-1: Vu_BC)
__global__ __launch_bounds__(kDrvThreads) void v_slice(int nP, int nR, int nV, int nK, int iv,
                                                       const double* __restrict__ psd, double* __restrict__ v) {
    const size_t n = (size_t)nP * nR * nK;
    const size_t i = (size_t)blockIdx.x * kDrvThreads + threadIdx.x;
    if (i >= n) return;
    const size_t pr = i / nK, k = i % nK;
    v[i] = psd[(pr * (size_t)nV + iv) * nK + k];
}

// UpdatableMatrix::update after its queues (src/UpdatableMatrix.cpp:672-
// 827): the current array is the original one, times the scaling
// coefficient when it is not 1 (operator*=, :705-712), then every point
// whose coordinate on one of the limited axes lies outside [from, to] is
// set to 0 (MatrixLimit, :196-300; an axis whose range covers all its
// coordinates excludes nothing, which is the reference's test for whether
// to limit at all).
__global__ __launch_bounds__(kDrvThreads) void field_update(double* __restrict__ dst, const double* __restrict__ src,
                                                            size_t n, double scale, verb_field_update_params prm) {
    const size_t stride = (size_t)gridDim.x * kDrvThreads;
    for (size_t i = (size_t)blockIdx.x * kDrvThreads + threadIdx.x; i < n; i += stride) {
        double v = src[i];
        if (scale != 1.0) v = v * scale;
        bool out = false;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            if (prm.grid[a]) {
                const double q = prm.grid[a][i];
                out = out || (q < prm.from[a]) || (q > prm.to[a]);
            }
        }
        dst[i] = out ? 0.0 : v;
    }
}

// min over points of DVV*DKK - DVK*DVK, as (DVV.times(DKK) -
// DVK.times(DVK)).min() (src/VERB4D_Solver.cpp:747; Matrix4D::min starts
// from +inf and takes std::min, so NaN entries are passed over,
// src/Matrix.cpp:5881-5888): per-block minima into part[]
__global__ __launch_bounds__(kDrvThreads) void stability_part(const double* __restrict__ dvv,
                                                              const double* __restrict__ dkk,
                                                              const double* __restrict__ dvk, size_t n,
                                                              double* __restrict__ part) {
    __shared__ double red[kDrvThreads];
    double m = __builtin_inf();
    const size_t stride = (size_t)gridDim.x * kDrvThreads;
    for (size_t i = (size_t)blockIdx.x * kDrvThreads + threadIdx.x; i < n; i += stride) {
        const double a = dvv[i] * dkk[i];
        const double b = dvk[i] * dvk[i];
        const double v = a - b;
        m = (v < m) ? v : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kDrvThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            red[threadIdx.x] = (o < red[threadIdx.x]) ? o : red[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kDrvThreads) void min_of(const double* __restrict__ part, int count,
                                                      double* __restrict__ out) {
    __shared__ double red[kDrvThreads];
    double m = __builtin_inf();
    for (int i = threadIdx.x; i < count; i += kDrvThreads) m = (part[i] < m) ? part[i] : m;
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kDrvThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            red[threadIdx.x] = (o < red[threadIdx.x]) ? o : red[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

int blocks_for(size_t n) {
    const size_t b = (n + kDrvThreads - 1) / kDrvThreads;
    return (int)(b < 4096 ? (b ? b : 1) : 4096);
}

}  // namespace

int launch_clamp_floor(double* psd, size_t n, double floor, hipStream_t stream) {
    hipLaunchKernelGGL(clamp_floor, dim3(blocks_for(n)), dim3(kDrvThreads), 0, stream, psd, n, floor);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_count_negative(const double* psd, size_t n, unsigned long long* count, hipStream_t stream) {
    hipLaunchKernelGGL(zero_count, dim3(1), dim3(1), 0, stream, count);
    hipLaunchKernelGGL(count_negative, dim3(blocks_for(n)), dim3(kDrvThreads), 0, stream, psd, n, count);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_v_slice(const verb_grid& g, const double* psd, int iv, double* v, hipStream_t stream) {
    const size_t n = (size_t)g.nP * g.nR * g.nK;
    const int blocks = (int)((n + kDrvThreads - 1) / kDrvThreads);
    if (blocks < 1) return VERB_OK;
    hipLaunchKernelGGL(v_slice, dim3(blocks), dim3(kDrvThreads), 0, stream, g.nP, g.nR, g.nV, g.nK, iv, psd, v);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_vl_from_psd(const verb_grid& g, const double* psd, double* vl, hipStream_t stream) {
    return launch_v_slice(g, psd, 0, vl, stream);
}

int launch_field_update(double* dst, const double* src, size_t n, const verb_field_update_params& p,
                        hipStream_t stream) {
    hipLaunchKernelGGL(field_update, dim3(blocks_for(n)), dim3(kDrvThreads), 0, stream, dst, src, n, p.scale, p);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

// part: kStabilityParts doubles of scratch; out: one double
int launch_stability_min(const verb_grid& g, double* part, double* out, hipStream_t stream) {
    const size_t n = (size_t)g.nP * g.nR * g.nV * g.nK;
    int blocks = blocks_for(n);
    if (blocks > kStabilityParts) blocks = kStabilityParts;
    hipLaunchKernelGGL(stability_part, dim3(blocks), dim3(kDrvThreads), 0, stream, g.DVV, g.DKK, g.DVK, n, part);
    hipLaunchKernelGGL(min_of, dim3(1), dim3(kDrvThreads), 0, stream, part, blocks, out);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace verb
