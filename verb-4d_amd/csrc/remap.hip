// Adiabatic remapping: when L* changes, every (P,V,K) line of the PSD is
// interpolated from the old L* values of its points to the new ones
// (src/VERB4D_Solver.cpp:557-619) with Cubic1D (src/Interpolation.cpp:25-51),
// a monotone cubic Hermite interpolator (src/MonotCubicInterpolator.cpp).
//
// Per line, in the reference's arithmetic (this file is compiled with
// -ffp-contract=off):
//   points   (0, 0) and then (L*_old[i], PSD[i]) in order, kept sorted and
//            unique, a repeated key taking the value inserted last: Cubic1D
//            sizes its vectors n and then push_backs the grid (:33-40), and
//            the interpolator stores the pairs in a std::map (:92-106);
//   monotone the direction of the first unequal pair, then every later pair
//            checked against it (:378-497);
//   slopes   one-sided at the ends, the mean of the two half-slopes inside
//            (:612-664), and for monotone data the Fritsch-Carlson limit
//            applied interval by interval, left to right (:668-704);
//   values   constant beyond both ends, the cubic Hermite form inside
//            (:175-233; bases src/MonotCubicInterpolator.hpp:522-537).
//
// Mapping: one thread per line, like the radial solve: the lines with
// consecutive (V,K) index are adjacent in HBM, so every R-level load and
// store of a wavefront is one contiguous 512-byte segment. A thread's
// sorted points and slopes live in an LDS column of its own (stride 64
// doubles: the lanes of a wavefront that touch the same entry hit 64
// consecutive addresses). Work per point is a few dozen flops; the kernel
// reads L*_old, L*_new and the PSD once and writes the PSD once (32 B per
// point).
#include "common.h"

namespace verb {
namespace {

constexpr int kRemapThreads = 64;

__device__ inline double h00(double t) { return 2 * t * t * t - 3 * t * t + 1; }
__device__ inline double h10(double t) { return t * t * t - 2 * t * t + t; }
__device__ inline double h01(double t) { return -2 * t * t * t + 3 * t * t; }
__device__ inline double h11(double t) { return t * t * t - t * t; }

__global__ __launch_bounds__(kRemapThreads) void remap_kernel(double* __restrict__ psd,
                                                              const double* __restrict__ lold,
                                                              const double* __restrict__ lnew, int n, long NQ,
                                                              long nlines, int* __restrict__ status) {
    extern __shared__ double sh[];
    const int tid = threadIdx.x;
    const long line = (long)blockIdx.x * kRemapThreads + tid;
    if (line >= nlines || n <= 0) return;  // no barriers below
    const long iP = line / NQ;
    const long q = line - iP * NQ;
    const size_t base = (size_t)iP * n * NQ + q;
    const int M = n + 1;
    double* K = sh + tid;
    double* V = K + (size_t)M * kRemapThreads;
    double* D = V + (size_t)M * kRemapThreads;
#define AT(A, s) A[(size_t)(s) * kRemapThreads]

    // a NaN key has no place in the map's order; the interpolator throws on
    // a non-finite point to evaluate: reject the line before touching it
    for (int i = 0; i < n; i++) {
        const double xo = lold[base + (size_t)i * NQ], xn = lnew[base + (size_t)i * NQ];
        if (isnan(xo) || !isfinite(xn)) {
            raise_error(status, VERB_ERR_DOMAIN);
            return;
        }
    }
    AT(K, 0) = 0.0;
    AT(V, 0) = 0.0;
    int m = 1;
    for (int i = 0; i < n; i++) {
        const double x = lold[base + (size_t)i * NQ];
        const double f = psd[base + (size_t)i * NQ];
        int pos = m;
        while (pos > 0 && x < AT(K, pos - 1)) pos--;
        if (pos > 0 && AT(K, pos - 1) == x) {
            AT(V, pos - 1) = f;
        } else {
            for (int s = m; s > pos; s--) {
                AT(K, s) = AT(K, s - 1);
                AT(V, s) = AT(V, s - 1);
            }
            AT(K, pos) = x;
            AT(V, pos) = f;
            m++;
        }
    }
    if (m >= 2) {
        bool mono = true;
        int s = 0;
        while (s + 1 < m && AT(V, s) == AT(V, s + 1)) s++;
        if (s + 1 < m) {
            const double a = AT(V, s), b = AT(V, s + 1);
            if (a > b) {
                for (s++; s + 1 < m; s++)
                    if (AT(V, s) < AT(V, s + 1)) { mono = false; break; }
            } else if (a < b) {
                for (s++; s + 1 < m; s++)
                    if (AT(V, s) > AT(V, s + 1)) { mono = false; break; }
            } else {
                // NaN in the pair that sets the direction: the interpolator
                // throws ("Programming logic error")
                raise_error(status, VERB_ERR_DOMAIN);
                return;
            }
        }
        AT(D, 0) = (AT(V, 1) - AT(V, 0)) / (AT(K, 1) - AT(K, 0));
        AT(D, m - 1) = (AT(V, m - 1) - AT(V, m - 2)) / (AT(K, m - 1) - AT(K, m - 2));
        for (s = 1; s < m - 1; s++)
            AT(D, s) = (AT(V, s + 1) - AT(V, s)) / (2 * (AT(K, s + 1) - AT(K, s))) +
                       (AT(V, s) - AT(V, s - 1)) / (2 * (AT(K, s) - AT(K, s - 1)));
        if (mono) {
            for (s = 0; s < m - 1; s++) {
                const double delta = (AT(V, s + 1) - AT(V, s)) / (AT(K, s + 1) - AT(K, s));
                if (fabs(delta) < 1e-14) {
                    AT(D, s) = 0.0;
                    AT(D, s + 1) = 0.0;
                } else {
                    const double alpha = AT(D, s) / delta, beta = AT(D, s + 1) / delta;
                    if (!((alpha * alpha + beta * beta) <= 9)) {
                        const double tau = 3 / sqrt(alpha * alpha + beta * beta);
                        AT(D, s) = tau * alpha * delta;
                        AT(D, s + 1) = tau * beta * delta;
                    }
                }
            }
        }
    }
    for (int i = 0; i < n; i++) {
        const double x = lnew[base + (size_t)i * NQ];
        int lo = 0, hi = m;  // first key not below x (std::map::lower_bound)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (AT(K, mid) < x) lo = mid + 1;
            else hi = mid;
        }
        double r;
        if (lo == 0) {
            r = AT(V, 0);
        } else if (lo == m) {
            r = AT(V, m - 1);
        } else {
            const double x1 = AT(K, lo - 1), x2 = AT(K, lo);
            const double t = (x - x1) / (x2 - x1);
            const double h = x2 - x1;
            r = AT(V, lo - 1) * h00(t) + AT(D, lo - 1) * h10(t) * h + AT(V, lo) * h01(t) + AT(D, lo) * h11(t) * h;
        }
        psd[base + (size_t)i * NQ] = r;
    }
#undef AT
}

}  // namespace

size_t remap_lds_bytes(int n) { return (size_t)3 * (n + 1) * kRemapThreads * sizeof(double); }

int launch_remap(double* psd, const double* lold, const double* lnew, int nP, int nR, int nV, int nK, int* status,
                 hipStream_t stream) {
    if (nP < 0 || nR < 0 || nV < 0 || nK < 0) return VERB_ERR_ARG;
    const long NQ = (long)nV * nK;
    const long nlines = (long)nP * NQ;
    if (nlines == 0 || nR == 0) return VERB_OK;
    int dev = 0, max_lds = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    const size_t lds = remap_lds_bytes(nR);
    if (lds > (size_t)max_lds) return VERB_ERR_ARG;
    hipFuncSetAttribute((const void*)remap_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const long blocks = (nlines + kRemapThreads - 1) / kRemapThreads;
    hipLaunchKernelGGL(remap_kernel, dim3((unsigned)blocks), dim3(kRemapThreads), lds, stream, psd, lold, lnew, nR,
                       NQ, nlines, status);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace verb
