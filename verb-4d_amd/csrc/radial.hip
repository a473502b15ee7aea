// Radial diffusion: implicit 1-D diffusion along L for every (P,V,K) line.
//
// Restates, per line, getCalculationMatrix1DDiffusion
// (src/Diffusion_1D.cpp:91-184) with its two SecondDerivativeApproximation_1D
// calls (src/MatrixSolver.cpp:353-404), the right-hand side C + B*f
// (src/Diffusion_1D.cpp:66-76) and the Thomas solve tridag
// (src/MatrixSolver.cpp:689-720), in the reference's operation order. This
// file is compiled with -ffp-contract=off, so results are bit-identical to
// the reference's CPU arithmetic.
//
// Mapping: one thread owns one (P,V,K) line. Lines with consecutive (V,K)
// index are adjacent in HBM ([P][R][V][K], K fastest), so every R-level
// load and store of a wavefront is one contiguous 512-byte segment. The
// assembly is fused into the forward sweep (a sliding 3-point window of
// L, D_LL and G), so each input value is read exactly once and the
// solution is written once: 56 bytes of HBM traffic per grid point.
#include "common.h"

namespace verb {
namespace {

template <int MAXR>
__global__ __launch_bounds__(256) void radial_kernel(const verb_grid g, double* __restrict__ psd,
                                                     int* __restrict__ status) {
    const long VK = (long)g.nV * g.nK;
    const long total = (long)g.nP * VK;
    const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= total) return;
    const int n = g.nR;
    const long iP = line / VK;
    const long q = line - iP * VK;
    const size_t base = (size_t)iP * n * VK + q;
    const size_t bidx = (size_t)iP * VK + q;  // [P][V][K] boundary arrays
    const double dt = g.dt;
    const double rs = g.radial_sources;

    double gam[MAXR];
    double u[MAXR];

    // Sliding window over x (L grid), D (D_LL), G (G_radial).
    double xm = 0, x0 = g.Lg[base], xp = g.Lg[base + VK];
    double Dm = 0, D0 = g.DLL[base], Dp = g.DLL[base + VK];
    double Gm = 0, G0 = g.G_radial[base], Gp = g.G_radial[base + VK];

    double bet = 0.0, aprev_c = 0.0;  // c[j-1] carried into row j
    bool bc_ok = true;
    // Per-point inputs loaded without a branch, from a clamped row (the
    // values of rows past the line are never used): the loads of later rows
    // are free to be issued ahead of the recurrence, instead of each row
    // waiting for its own.
    double ps[MAXR], sr[MAXR], lr[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; j++) {
        const size_t pj = base + (size_t)min(j, n - 1) * VK;
        ps[j] = psd[pj];
        sr[j] = g.sources[pj];
        lr[j] = g.losses_radial[pj];
    }
#pragma unroll
    for (int j = 0; j < MAXR; j++) {
        if (j < n) {
            double a = 0.0, b = 0.0, c = 0.0, B0 = 0.0, C0 = 0.0;
            if (j == 0) {
                C0 = g.Ll[bidx];
                const double dh = xp - x0;
                if (g.Ll_type == VERB_BC_CONSTANT_VALUE) { b = 1; c = 0; }
                else if (g.Ll_type == VERB_BC_CONSTANT_DERIVATIVE) { b = 1 / dh; c = -1 / dh; }
                else bc_ok = false;
            } else if (j == n - 1) {
                C0 = g.Lu[bidx];
                const double dh = x0 - xm;
                if (g.Lu_type == VERB_BC_CONSTANT_VALUE) { b = 1; a = 0; }
                else if (g.Lu_type == VERB_BC_CONSTANT_DERIVATIVE) { b = 1 / dh; a = -1 / dh; }
                else bc_ok = false;
            } else {
                b += 1.0 / dt;
                B0 += 1.0 / dt;
                C0 += sr[j] * rs;
                b -= lr[j];
                {   // ("x_left", "x_right"): dx1 = -1, dx2 = +1
                    const double dh1 = (x0) - (xm);
                    const double dh21 = (x0) - (xp);
                    const double dh22 = (xm) - (x0);
                    const double cp = -0.5 / G0 / dh1;
                    b += +cp / dh21 * D0 * G0;
                    c += -cp / dh21 * D0 * G0;
                    a += -cp / dh22 * Dm * Gm;
                    b += +cp / dh22 * Dm * Gm;
                }
                {   // ("x_right", "x_left"): dx1 = +1, dx2 = -1
                    const double dh1 = (x0) - (xp);
                    const double dh21 = (x0) - (xm);
                    const double dh22 = (xp) - (x0);
                    const double cp = -0.5 / G0 / dh1;
                    b += +cp / dh21 * D0 * G0;
                    a += -cp / dh21 * D0 * G0;
                    c += -cp / dh22 * Dp * Gp;
                    b += +cp / dh22 * Dp * Gp;
                }
            }
            double r = C0;
            r += B0 * ps[j];
            // Thomas forward sweep (tridag)
            if (j == 0) {
                u[0] = r / (bet = b);
            } else {
                gam[j] = aprev_c / bet;
                bet = b - a * gam[j];
                u[j] = (r - a * u[j - 1]) / bet;
            }
            aprev_c = c;
            // slide the window
            xm = x0; x0 = xp;
            Dm = D0; D0 = Dp;
            Gm = G0; G0 = Gp;
            {
                const size_t pn = base + (size_t)min(j + 2, n - 1) * VK;
                xp = g.Lg[pn];
                Dp = g.DLL[pn];
                Gp = g.G_radial[pn];
            }
        }
    }
#pragma unroll
    for (int j = MAXR - 1; j >= 1; j--) {
        if (j <= n - 1) u[j - 1] -= gam[j] * u[j];
    }
#pragma unroll
    for (int j = 0; j < MAXR; j++)
        if (j < n) psd[base + (size_t)j * VK] = u[j];
    if (!bc_ok) raise_error(status, VERB_ERR_BC);
}

// Stand-alone batched tridag on line-major arrays (the drop-in of tridag).
template <int MAXN>
__global__ __launch_bounds__(256) void tridag_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                     const double* __restrict__ c, const double* __restrict__ r,
                                                     double* __restrict__ u_out, int n, int lines) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= lines) return;
    const size_t o = (size_t)l * n;
    double gam[MAXN], u[MAXN];
    double bet;
    u[0] = r[o] / (bet = b[o]);
#pragma unroll
    for (int j = 1; j < MAXN; j++) {
        if (j < n) {
            gam[j] = c[o + j - 1] / bet;
            bet = b[o + j] - a[o + j] * gam[j];
            u[j] = (r[o + j] - a[o + j] * u[j - 1]) / bet;
        }
    }
#pragma unroll
    for (int j = MAXN - 1; j >= 1; j--)
        if (j <= n - 1) u[j - 1] -= gam[j] * u[j];
#pragma unroll
    for (int j = 0; j < MAXN; j++)
        if (j < n) u_out[o + j] = u[j];
}

}  // namespace

int launch_radial(const verb_grid& g, double* psd, int* status, hipStream_t stream) {
    if (g.nR < 3 || g.nR > 64) return VERB_ERR_ARG;
    const long total = (long)g.nP * g.nV * g.nK;
    const int threads = 256;
    const long blocks = (total + threads - 1) / threads;
    if (g.nR <= 32)
        hipLaunchKernelGGL(radial_kernel<32>, dim3((unsigned)blocks), dim3(threads), 0, stream, g, psd, status);
    else
        hipLaunchKernelGGL(radial_kernel<64>, dim3((unsigned)blocks), dim3(threads), 0, stream, g, psd, status);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_tridag(const double* a, const double* b, const double* c, const double* r, double* u, int n, int lines,
                  hipStream_t stream) {
    if (n < 1 || n > 64 || lines < 0) return VERB_ERR_ARG;
    if (lines == 0) return VERB_OK;
    const int threads = 256;
    const int blocks = (lines + threads - 1) / threads;
    if (n <= 32)
        hipLaunchKernelGGL(tridag_kernel<32>, dim3(blocks), dim3(threads), 0, stream, a, b, c, r, u, n, lines);
    else
        hipLaunchKernelGGL(tridag_kernel<64>, dim3(blocks), dim3(threads), 0, stream, a, b, c, r, u, n, lines);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace verb
