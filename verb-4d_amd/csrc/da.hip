// Data assimilation along L: the Kalman filter that the reference runs on
// every (V,K) line of the P = iP slice of the PSD after a radial step
// (DataAssimilationManagerDiffusion::assimilate, src/DataAssimilationDiffusion.cpp:
// 70-149, runKalmanFilterDiffusion1D :151-249, runKalmanFilter
// src/DataAssimilation.cpp:18-73):
//
//   F  = B A^-1, A and B the implicit 1-D diffusion matrices of the line
//        with zero sources and losses (getModelMatrixDiffusion1D,
//        src/DataAssimilationHelper.cpp:168-189: trans_solve(A, B), a dgesv
//        on A^T with B^T as right-hand sides, src/MatrixOperations.cpp:761-787)
//   Pa = 5 modelError I when the line's covariance is still all zero
//   Pf = F Pa F^T + modelError I
//   S  = H Pf H^T + observationError I,  K = Pf H^T S^-1
//   x += K (obs - H x),  Pa = Pf - K H Pf
//
// H selects the line's finite observations (convertToObservationSpace,
// src/DataAssimilationHelper.cpp:271-305): H X and X H^T are row and column
// selections, exact like the reference's products with the 0/1 matrix. A
// line without observations keeps x and takes Pa = Pf (its one all-zero
// pseudo-observation gives K = 0).
//
// Mapping: one wavefront per line; F, Pa, Pf and the work matrices stay in
// LDS (n <= 32 points, padded rows of 33); every product is a 4x4 register
// block per lane over LDS operands; the two dense solves are LU with
// partial pivoting (the pivot order of dgetf2) by the wavefront. FP64
// throughout. The arithmetic differs from OpenBLAS's in summation order, so
// results agree with the reference to rounding, not bit for bit.
#include "common.h"

namespace verb {
namespace {

constexpr int kDaN = 32;       // largest line
constexpr int kDaLd = 33;      // LDS row stride (bank spread)
constexpr int kDaMat = kDaN * kDaLd;

struct DaArgs {
    const double* Lg;
    const double* DLL;
    const double* G;
    int Ll_type, Lu_type;
    double* psd;        // [P][R][V][K]
    const double* obs;  // [P][R][V][K], non-finite = no observation
    double* Pa;         // [V][K][n][n]
    double* F_out;      // [V][K][n][n] or null
    int iP, nR;
    long VK;
    double dt, model_error, observation_error;
    int* status;
};

// C[i][j] = sum_k A[i][k] B[k][j] (tb: B given as B^T, i.e. read B[j][k]),
// i < rows, j < cols, k < inner; one 4x4 block per lane over 32x32
__device__ inline void gemm_lds(double* C, const double* A, const double* B, int rows, int cols, int inner, bool tb,
                                int lane) {
    const int i0 = (lane >> 3) * 4, j0 = (lane & 7) * 4;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = 0.0;
    for (int k = 0; k < inner; k++) {
        double a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; r++) a[r] = A[(i0 + r) * kDaLd + k];
#pragma unroll
        for (int c = 0; c < 4; c++) b[c] = tb ? B[(j0 + c) * kDaLd + k] : B[k * kDaLd + j0 + c];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[r][c] = fma(a[r], b[c], acc[r][c]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (i0 + r < rows && j0 + c < cols) C[(i0 + r) * kDaLd + j0 + c] = acc[r][c];
    __builtin_amdgcn_wave_barrier();
}

// In place LU with partial pivoting of the m x m matrix M (dgetf2's pivot
// choice: the first largest |value| in the column), then X (rhs rows of
// length m, each row one right-hand side b^T) replaced by the solutions of
// M y = b. m, rhs <= 32. Lane (c, h) = (lane % 32, lane / 32) works on
// column / right-hand side c and every other row from h, so each
// elimination step is one short loop per lane with no index division.
// Returns false on a zero pivot.
__device__ inline bool lu_solve_rows(double* M, double* X, int m, int rhs, int lane, int* piv) {
    const int c = lane & 31, h = lane >> 5;
    bool ok = true;
    for (int j = 0; j < m; j++) {
        // pivot search over rows j..m-1 of column j, then the first index of
        // the maximum
        double best = -1.0;
        int bi = m;
        const int i0 = j + lane;
        if (i0 < m) {
            best = fabs(M[i0 * kDaLd + j]);
            bi = i0;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            const double ob = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        const int p = bi;
        if (lane == 0) piv[j] = p;
        if (!(best > 0.0)) ok = false;
        if (p != j && p < m && h == 0 && c < m) {
            const double t = M[j * kDaLd + c];
            M[j * kDaLd + c] = M[p * kDaLd + c];
            M[p * kDaLd + c] = t;
        }
        __builtin_amdgcn_wave_barrier();
        const double d = M[j * kDaLd + j];
        const double rinv = d != 0.0 ? 1.0 / d : 0.0;
        if (h == 0 && j + 1 + c < m && d != 0.0) M[(j + 1 + c) * kDaLd + j] *= rinv;
        __builtin_amdgcn_wave_barrier();
        // rank-1 update of the trailing block: lane (c, h) takes column j+1+c,
        // rows j+1+h, j+3+h, ...
        const int col = j + 1 + c;
        if (col < m) {
            const double u = M[j * kDaLd + col];
            for (int i = j + 1 + h; i < m; i += 2) M[i * kDaLd + col] -= M[i * kDaLd + j] * u;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // right-hand side c: interchanges, then forward and back substitution,
    // rows split over the two lane halves
    double* x = X + c * kDaLd;
    const bool mine = c < rhs;
    for (int j = 0; j < m; j++) {
        const int p = piv[j];
        if (mine && h == 0 && p != j) {
            const double t = x[j];
            x[j] = x[p];
            x[p] = t;
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int j = 0; j < m; j++) {
        if (mine) {
            const double v = x[j];
            for (int i = j + 1 + h; i < m; i += 2) x[i] -= M[i * kDaLd + j] * v;
        }
        __builtin_amdgcn_wave_barrier();
    }
    for (int j = m - 1; j >= 0; j--) {
        if (mine && h == 0) x[j] = x[j] / M[j * kDaLd + j];
        __builtin_amdgcn_wave_barrier();
        if (mine) {
            const double v = x[j];
            for (int i = h; i < j; i += 2) x[i] -= M[i * kDaLd + j] * v;
        }
        __builtin_amdgcn_wave_barrier();
    }
    return ok;
}

__global__ __launch_bounds__(64) void da_radial_kernel(DaArgs a) {
    extern __shared__ double sh[];
    const int lane = threadIdx.x;
    const long q = blockIdx.x;  // (V,K) line
    if (q >= a.VK) return;
    const int n = a.nR;
    double* M0 = sh;                 // A^T and its LU
    double* M1 = sh + kDaMat;        // F; later S^T and its LU, then K H Pf
    double* M2 = sh + 2 * kDaMat;    // Pa, later Pf H^T and K
    double* M3 = sh + 3 * kDaMat;    // F Pa, later H Pf
    double* M4 = M0;                 // Pf (A's factors are done with by then)
    double* x = sh + 4 * kDaMat;     // the line's PSD [32]
    double* ob = x + kDaN;           // observations at the selected points
    int* idx = (int*)(ob + kDaN);    // selected points
    int* piv = idx + kDaN;
    int* nobs = piv + kDaN;
    const size_t base = ((size_t)a.iP * n) * a.VK + q;  // element iR at base + iR*VK
    double* Pa = a.Pa + (size_t)q * n * n;

    // zero the work matrices (padding included)
    for (int t = lane; t < 4 * kDaMat; t += 64) sh[t] = 0.0;
    __builtin_amdgcn_wave_barrier();

    // 1. the 1-D implicit diffusion matrices (getCalculationMatrix1DDiffusion,
    // src/Diffusion_1D.cpp:91-184, zero sources and losses): A^T into M0,
    // the right-hand sides B^T (diagonal) into M1 rows: row j = B[j][j] e_j
    bool bc_ok = true;
    if (lane < n) {
        const int j = lane;
        const size_t pj = base + (size_t)j * a.VK;
        double am = 0.0, b = 0.0, c = 0.0, B0 = 0.0;
        const double x0 = a.Lg[pj];
        const double xm = j > 0 ? a.Lg[pj - a.VK] : 0.0, xp = j + 1 < n ? a.Lg[pj + a.VK] : 0.0;
        if (j == 0) {
            const double dh = xp - x0;
            if (a.Ll_type == VERB_BC_CONSTANT_VALUE) { b = 1; c = 0; }
            else if (a.Ll_type == VERB_BC_CONSTANT_DERIVATIVE) { b = 1 / dh; c = -1 / dh; }
            else bc_ok = false;
        } else if (j == n - 1) {
            const double dh = x0 - xm;
            if (a.Lu_type == VERB_BC_CONSTANT_VALUE) { b = 1; am = 0; }
            else if (a.Lu_type == VERB_BC_CONSTANT_DERIVATIVE) { b = 1 / dh; am = -1 / dh; }
            else bc_ok = false;
        } else {
            const double D0 = a.DLL[pj], Dm = a.DLL[pj - a.VK], Dp = a.DLL[pj + a.VK];
            const double G0 = a.G[pj], Gm = a.G[pj - a.VK], Gp = a.G[pj + a.VK];
            b += 1.0 / a.dt;
            B0 += 1.0 / a.dt;
            b -= 0.0;
            {
                const double dh1 = (x0) - (xm);
                const double dh21 = (x0) - (xp);
                const double dh22 = (xm) - (x0);
                const double cp = -0.5 / G0 / dh1;
                b += +cp / dh21 * D0 * G0;
                c += -cp / dh21 * D0 * G0;
                am += -cp / dh22 * Dm * Gm;
                b += +cp / dh22 * Dm * Gm;
            }
            {
                const double dh1 = (x0) - (xp);
                const double dh21 = (x0) - (xm);
                const double dh22 = (xp) - (x0);
                const double cp = -0.5 / G0 / dh1;
                b += +cp / dh21 * D0 * G0;
                am += -cp / dh21 * D0 * G0;
                c += -cp / dh22 * Dp * Gp;
                b += +cp / dh22 * Dp * Gp;
            }
        }
        // row j of A: A[j][j-1] = am, A[j][j] = b, A[j][j+1] = c; stored as A^T
        M0[j * kDaLd + j] = b;
        if (j > 0) M0[(j - 1) * kDaLd + j] = am;
        if (j + 1 < n) M0[(j + 1) * kDaLd + j] = c;
        M1[j * kDaLd + j] = B0;
        x[j] = a.psd[pj];
        const double o = a.obs[pj];
        ob[j] = o;
    }
    if (__builtin_amdgcn_ballot_w64(!bc_ok) != 0) {
        if (lane == 0) raise_error(a.status, VERB_ERR_BC);
        return;
    }
    __builtin_amdgcn_wave_barrier();
    // selected points (finite observations, in order)
    {
        const bool fin = lane < n && isfinite(ob[lane]);
        const unsigned long long bal = __builtin_amdgcn_ballot_w64(fin);
        const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
        const double o = lane < n ? ob[lane] : 0.0;
        __builtin_amdgcn_wave_barrier();
        if (fin) {
            idx[pos] = lane;
            ob[pos] = o;
        }
        if (lane == 0) *nobs = __builtin_popcountll(bal);
    }
    __builtin_amdgcn_wave_barrier();
    const int m = *nobs;

    // F: solve A^T X = B^T (rows of M1 are the right-hand sides), F = X^T
    if (!lu_solve_rows(M0, M1, n, n, lane, piv)) {
        if (lane == 0) raise_error(a.status, VERB_ERR_SINGULAR);
        return;
    }
    // row i of F = B A^-1 is B[i][i] e_i^T A^-1: the solution of
    // A^T y = B[i][i] e_i, now row i of M1
    if (a.F_out)
        for (int t = lane; t < n * n; t += 64) a.F_out[(size_t)q * n * n + t] = M1[(t / n) * kDaLd + t % n];

    // 2. Pa (all zero: 5 modelError on the diagonal), Pf = F Pa F^T + Q
    double amax = 0.0;
    for (int t = lane; t < n * n; t += 64) {
        const double v = Pa[t];
        M2[(t / n) * kDaLd + t % n] = v;
        amax = fmax(amax, fabs(v));
    }
    for (int o = 32; o >= 1; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o));
    __builtin_amdgcn_wave_barrier();
    if (amax == 0.0 && lane < n) M2[lane * kDaLd + lane] = 5 * a.model_error;
    __builtin_amdgcn_wave_barrier();
    gemm_lds(M3, M1, M2, n, n, n, false, lane);  // F Pa
    gemm_lds(M4, M3, M1, n, n, n, true, lane);   // (F Pa) F^T
    for (int t = lane; t < n * n; t += 64) {
        const int i = t / n, j = t % n;
        M4[i * kDaLd + j] += (i == j ? a.model_error : 0.0);
    }
    __builtin_amdgcn_wave_barrier();

    if (m > 0) {
        // 3. S = H Pf H^T + R, stored transposed in M1; Pf H^T (n x m) in M2
        for (int t = lane; t < m * m; t += 64) {
            const int i = t / m, j = t % m;
            M1[j * kDaLd + i] = M4[idx[i] * kDaLd + idx[j]] + (i == j ? a.observation_error : 0.0);
        }
        for (int t = lane; t < n * m; t += 64) {
            const int i = t / m, r = t % m;
            M2[i * kDaLd + r] = M4[i * kDaLd + idx[r]];
        }
        __builtin_amdgcn_wave_barrier();
        // 4. K = Pf H^T S^-1: row i of K solves S^T z = (row i of Pf H^T)^T
        if (!lu_solve_rows(M1, M2, m, n, lane, piv)) {
            if (lane == 0) raise_error(a.status, VERB_ERR_SINGULAR);
            return;
        }
        // M2 row i = row i of K (n x m)
        // 5. x += K (obs - H x)
        if (lane < n) {
            double corr = 0.0;
            for (int t = 0; t < m; t++) corr = fma(M2[lane * kDaLd + t], ob[t] - x[idx[t]], corr);
            const double xn = x[lane] + corr;
            if (isnan(xn)) raise_error(a.status, VERB_ERR_DOMAIN);
            a.psd[base + (size_t)lane * a.VK] = xn;
        }
        // 6. Pa = Pf - K (H Pf): H Pf = rows idx of Pf into M3 (m x n)
        for (int t = lane; t < m * n; t += 64) {
            const int r = t / n, j = t % n;
            M3[r * kDaLd + j] = M4[idx[r] * kDaLd + j];
        }
        __builtin_amdgcn_wave_barrier();
        gemm_lds(M1, M2, M3, n, n, m, false, lane);
        for (int t = lane; t < n * n; t += 64) {
            const int i = t / n, j = t % n;
            Pa[t] = M4[i * kDaLd + j] - M1[i * kDaLd + j];
        }
    } else {
        for (int t = lane; t < n * n; t += 64) Pa[t] = M4[(t / n) * kDaLd + t % n];
    }
}

}  // namespace

size_t da_lds_bytes() {
    return (size_t)4 * kDaMat * sizeof(double) + 2 * kDaN * sizeof(double) + (2 * kDaN + 1) * sizeof(int);
}

int launch_da_radial(const verb_grid& g, double* psd, const double* obs, double* Pa, double* F_out, int iP, double dt,
                     double model_error, double observation_error, int* status, hipStream_t stream) {
    if (g.nR < 3 || g.nR > kDaN || iP < 0 || iP >= g.nP) return VERB_ERR_ARG;
    DaArgs a;
    a.Lg = g.Lg;
    a.DLL = g.DLL;
    a.G = g.G_radial;
    a.Ll_type = g.Ll_type;
    a.Lu_type = g.Lu_type;
    a.psd = psd;
    a.obs = obs;
    a.Pa = Pa;
    a.F_out = F_out;
    a.iP = iP;
    a.nR = g.nR;
    a.VK = (long)g.nV * g.nK;
    a.dt = dt;
    a.model_error = model_error;
    a.observation_error = observation_error;
    a.status = status;
    hipLaunchKernelGGL(da_radial_kernel, dim3((unsigned)a.VK), dim3(64), da_lds_bytes(), stream, a);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace verb
