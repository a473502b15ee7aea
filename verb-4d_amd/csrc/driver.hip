// Pieces of the reference's time loop around the three operators
// (src/VERB4D_Solver.cpp:539-1282) that act on the whole device-resident
// grid: the positivity clamp PSD.max_of(1e-21) (:621, 980, 1072, 1218-1221;
// Matrix4D::max_of, src/Matrix.cpp:5948-5953), the count of negative points
// when no clamp is applied (:1223-1238), and the V-lower boundary taken from
// the PSD after convection (:965-966, Vl_BC = PSD.ySlice(0)).
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

// vl[p][r][k] = psd[p][r][0][k]
__global__ __launch_bounds__(kDrvThreads) void vl_from_psd(int nP, int nR, int nV, int nK,
                                                           const double* __restrict__ psd, double* __restrict__ vl) {
    const size_t n = (size_t)nP * nR * nK;
    const size_t i = (size_t)blockIdx.x * kDrvThreads + threadIdx.x;
    if (i >= n) return;
    const size_t pr = i / nK, k = i % nK;
    vl[i] = psd[pr * (size_t)nV * nK + k];
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

int launch_vl_from_psd(const verb_grid& g, const double* psd, double* vl, hipStream_t stream) {
    const size_t n = (size_t)g.nP * g.nR * g.nK;
    const int blocks = (int)((n + kDrvThreads - 1) / kDrvThreads);
    if (blocks < 1) return VERB_OK;
    hipLaunchKernelGGL(vl_from_psd, dim3(blocks), dim3(kDrvThreads), 0, stream, g.nP, g.nR, g.nV, g.nK, psd, vl);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace verb
