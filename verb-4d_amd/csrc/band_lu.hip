// Batched banded LU with LAPACK's partial pivoting, one workgroup per
// (P,R) node, window resident in MFMA accumulators.
//
// Semantics: the factorisation of dgbtrf/dgbtf2 as the reference calls it
// (src/MatrixSolver.cpp:272/302): pivot = first largest |a| among rows
// j..j+kl of column j; ipiv 1-based; U stored in rows 0..kl+ku of the LAPACK
// band column, the multipliers in rows kl+ku+1.. with the dgbtf2 convention
// (each column's multipliers in the row order of its own step). The
// right-hand side B0*psd + C0 is forward-eliminated on the fly, in
// dgbtrs order (swap, then subtract the column's multipliers), so the plain
// path never stores L.
//
// Blocking. The matrix is energy-fast (kl = ku = nV+1). Panels are NB = 16
// columns wide. At panel p the live window is RB row blocks x CB column
// blocks of 16x16 FP64 tiles (RB*16 >= kl+16, CB*16 >= kl+ku+16); row block
// b lives in wavefront b mod RB and column block c in register slot c mod CB
// of every wavefront, so the window slides by one block per panel without
// moving data: a finished row block's wavefront reloads the row block
// entering at the bottom, every wavefront reloads the slot of the finished
// column block with the column block entering on the right, both straight
// from the assembled 9-diagonal operator (entries that enter the window have
// not yet been touched by any elimination step).
//
// Per panel:
//   1. every wavefront copies its tile of the panel column block to LDS;
//   2. wavefront 0 factors the panel (118 x 16 at kl = 102) in LDS with
//      full-row interchanges and eliminates the right-hand side window;
//   3. the panel's 16 interchanges are applied to the rest of the window as
//      one net permutation through LDS, and U12 = L11^-1 P A12 is formed by
//      one lane per column;
//   4. A22 -= L21 U12 with v_mfma_f64_16x16x4f64 on the register tiles,
//      skipping column blocks beyond the running fill extent ju; the panel
//      wavefront meanwhile streams the U rows (and, for the cached path, the
//      multipliers re-ordered to the dgbtf2 convention) to HBM.
#include "common.h"
#include "band_lu.h"

namespace verb {

typedef double v4d __attribute__((ext_vector_type(4)));

int local_lu_variant = 0;

namespace {

constexpr int NB = 16;

__device__ inline double a_orig(const double* dg, int n, int nx, int r, int c) {
    if (r >= n || c >= n) return 0.0;
    const int off = c - r;
    const int dy = (off >= nx - 1) ? 1 : ((off <= 1 - nx) ? -1 : 0);
    const int dx = off - dy * nx;
    if (dx < -1 || dx > 1) return 0.0;
    return dg[(size_t)((dy + 1) * 3 + dx + 1) * n + r];
}

// Accumulator layout of v_mfma_f64_16x16x4f64: lane j + 16*g, element r
// holds row 4*r + g, column j of the 16x16 tile (rows interleave across the
// four 16-lane groups).
__device__ inline int crow(int g, int r) { return 4 * r + g; }

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP move of a double within 16-lane rows (ctrl: a row-local DPP pattern).
template <int CTRL>
__device__ inline double dpp_row(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Wavefront maximum: quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror and row_mirror leave every lane of a 16-lane row with the
// row maximum; the four row maxima are then combined.
__device__ inline double wave_max(double v) {
    v = fmax(v, dpp_row<0xB1>(v));
    v = fmax(v, dpp_row<0x4E>(v));
    v = fmax(v, dpp_row<0x141>(v));
    v = fmax(v, dpp_row<0x140>(v));
    const double r0 = __shfl(v, 0), r1 = __shfl(v, 16), r2 = __shfl(v, 32), r3 = __shfl(v, 48);
    return fmax(fmax(r0, r1), fmax(r2, r3));
}

template <int RB, int CB>
struct Lds {
    static constexpr int W = NB * (CB - 1);  // window columns right of the panel
    static constexpr int PR = NB * RB;       // window rows
    double pnl[PR * NB];                     // panel (full-row interchanges)
    double pnl2[PR * NB];                    // copy for the dgbtf2-order multipliers
    double u12[NB * W];
    double sb[2 * NB * W];                   // rows touched by the panel's interchanges
    double rw[PR];                           // right-hand-side window (index mod PR)
    int jp[NB];
    int uext[NB];                            // U row extents of the panel rows
    int srow[2 * NB];
    int ssrc[2 * NB];
    int ns, ju, info;
};

}  // namespace

template <int RB, int CB>
__global__ __launch_bounds__(RB * 64, 1) void band_lu_window(LuArgs a) {
    using L = Lds<RB, CB>;
    constexpr int W = L::W;
    constexpr int PR = L::PR;
    extern __shared__ double lds_raw[];
    L& S = *reinterpret_cast<L*>(lds_raw);

    const int node = a.order ? a.order[blockIdx.x] : blockIdx.x;
    if (a.nR > 0 && !a.include_boundary && (node % a.nR == 0 || node % a.nR == a.nR - 1)) return;
    const int stride = a.n;
    const int n = a.nn ? a.nn[node] : a.n, nx = a.nx;
    const int kl = nx + 1, ku = nx + 1, kv = kl + ku, ldab = 2 * kl + ku + 1;
    const double* dg = a.diag + (size_t)node * 9 * stride;
    const double* rhs = a.rhs ? a.rhs + (size_t)node * stride : nullptr;
    double* yv = a.y ? a.y + (size_t)node * stride : nullptr;
    double* ab = a.ab + (size_t)node * ldab * stride;
    int64_t* ipiv = a.ipiv + (size_t)node * stride;
    int* uext = a.uext ? a.uext + (size_t)node * stride : nullptr;

    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63, g = l >> 4, jj = l & 15;

    v4d acc[CB];
    // initial window: row block w, column blocks 0..CB-1 in slots 0..CB-1
#pragma unroll
    for (int s = 0; s < CB; s++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[s][r] = a_orig(dg, n, nx, NB * w + crow(g, r), NB * s + jj);
    for (int i = tid; i < PR; i += RB * 64) S.rw[i] = (rhs && i < n) ? rhs[i] : 0.0;
    if (tid == 0) {
        S.ju = 0;
        S.info = 0;
    }
    __syncthreads();

    const int npan = (n + NB - 1) / NB;
    for (int p = 0; p < npan; p++) {
        const int j0 = NB * p;
        const int s0 = p % CB;
        const int lb = (w - p % RB + RB) % RB;  // local row block of this wavefront
        const int nc = min(NB, n - j0);

        // 1. panel column block -> LDS
#pragma unroll
        for (int s = 0; s < CB; s++)
            if (s == s0) {
#pragma unroll
                for (int r = 0; r < 4; r++) S.pnl[(NB * lb + crow(g, r)) * NB + jj] = acc[s][r];
            }
        __syncthreads();

        // 2. panel factorisation (wavefront 0)
        if (w == 0) {
            int ju = S.ju;
            int info = S.info;
            for (int c = 0; c < nc; c++) {
                const int j = j0 + c;
                const int km = min(kl, n - 1 - j);
                const double v1 = (l <= km) ? fabs(S.pnl[(c + l) * NB + c]) : -1.0;
                const double v2 = (l + 64 <= km) ? fabs(S.pnl[(c + l + 64) * NB + c]) : -1.0;
                const double m = wave_max(fmax(v1, v2));
                const unsigned long long b1 = __ballot(v1 == m && l <= km);
                const unsigned long long b2 = __ballot(v2 == m && l + 64 <= km);
                int jp = 0;
                if (b1) jp = __ffsll(b1) - 1;
                else if (b2) jp = 64 + __ffsll(b2) - 1;
                if (l == 0) ipiv[j] = (int64_t)(j + jp + 1);
                const double piv = S.pnl[(c + jp) * NB + c];
                if (piv != 0.0) {
                    ju = max(ju, min(j + ku + jp, n - 1));
                    if (jp != 0) {
                        if (l < NB) {
                            const double t = S.pnl[c * NB + l];
                            S.pnl[c * NB + l] = S.pnl[(c + jp) * NB + l];
                            S.pnl[(c + jp) * NB + l] = t;
                        }
                        if (l == 0) {
                            const double t = S.rw[j % PR];
                            S.rw[j % PR] = S.rw[(j + jp) % PR];
                            S.rw[(j + jp) % PR] = t;
                        }
                    } else {
                        jp = 0;
                    }
                    wave_sync();
                    const double rinv = 1.0 / S.pnl[c * NB + c];
                    double L0 = 0.0, L1 = 0.0;
                    const bool h0 = (l + 1 <= km), h1 = (l + 65 <= km);
                    if (h0) {
                        L0 = S.pnl[(c + 1 + l) * NB + c] * rinv;
                        S.pnl[(c + 1 + l) * NB + c] = L0;
                    }
                    if (h1) {
                        L1 = S.pnl[(c + 65 + l) * NB + c] * rinv;
                        S.pnl[(c + 65 + l) * NB + c] = L1;
                    }
                    const double tb = -S.rw[j % PR];
                    if (tb != 0.0) {
                        if (h0) S.rw[(j + 1 + l) % PR] += L0 * tb;
                        if (h1) S.rw[(j + 65 + l) % PR] += L1 * tb;
                    }
                    for (int cc = c + 1; cc < NB; cc++) {
                        const double temp = -S.pnl[c * NB + cc];
                        if (temp != 0.0) {
                            if (h0) S.pnl[(c + 1 + l) * NB + cc] += L0 * temp;
                            if (h1) S.pnl[(c + 65 + l) * NB + cc] += L1 * temp;
                        }
                    }
                } else {
                    jp = 0;
                    if (info == 0) info = j + 1;
                }
                if (l == 0) {
                    S.jp[c] = jp;
                    S.uext[c] = ju;
                    if (uext) uext[j] = ju;
                }
                wave_sync();
            }
            for (int c = nc + l; c < NB; c += 64) S.jp[c] = 0;
            if (l == 0) {
                S.ju = ju;
                S.info = info;
                // net permutation of the rows the interchanges touched
                int rows[2 * NB], cont[2 * NB];
                int cnt = NB;
                for (int c = 0; c < NB; c++) { rows[c] = c; cont[c] = c; }
                for (int c = 0; c < nc; c++) {
                    const int t = c + S.jp[c];
                    int k = -1;
                    for (int q = 0; q < cnt; q++)
                        if (rows[q] == t) k = q;
                    if (k < 0) { rows[cnt] = t; cont[cnt] = t; k = cnt++; }
                    const int tmp = cont[c]; cont[c] = cont[k]; cont[k] = tmp;
                }
                for (int q = 0; q < cnt; q++) {
                    int src = 0;
                    for (int m = 0; m < cnt; m++)
                        if (rows[m] == cont[q]) src = m;
                    S.srow[q] = rows[q];
                    S.ssrc[q] = src;
                }
                S.ns = cnt;
            }
        }
        __syncthreads();
        const int ju = S.ju;
        const int ns = S.ns;

        // 3a. rows touched by the interchanges -> LDS (columns right of the panel)
        for (int q = 0; q < ns; q++) {
            const int row = S.srow[q];
            if (row / NB == lb && (row & 3) == g) {
                const int r = (row & 15) >> 2;
#pragma unroll
                for (int s = 0; s < CB; s++) {
                    if (s == s0) continue;
                    const int lc = (s - s0 + CB) % CB;
                    double v = (r == 0) ? acc[s][0] : (r == 1) ? acc[s][1] : (r == 2) ? acc[s][2] : acc[s][3];
                    S.sb[q * W + NB * (lc - 1) + jj] = v;
                }
            }
        }
        if (w == 0) {
            for (int i = l; i < PR * NB; i += 64) S.pnl2[i] = S.pnl[i];
        }
        __syncthreads();

        // 3b. U12 = L11^-1 P A12, one lane per column
        for (int t = tid; t < W; t += RB * 64) {
            double u[NB];
#pragma unroll
            for (int c = 0; c < NB; c++) u[c] = (c < nc) ? S.sb[S.ssrc[c] * W + t] : 0.0;
#pragma unroll
            for (int c = 1; c < NB; c++)
#pragma unroll
                for (int m = 0; m < c; m++) {
                    const double temp = -u[m];
                    if (temp != 0.0) u[c] += S.pnl[c * NB + m] * temp;
                }
#pragma unroll
            for (int c = 0; c < NB; c++) S.u12[c * W + t] = u[c];
        }
        // 3c. displaced rows below the panel take their new contents
        for (int q = NB; q < ns; q++) {
            const int row = S.srow[q];
            if (S.ssrc[q] == q) continue;
            if (row / NB == lb && (row & 3) == g) {
                const int r = (row & 15) >> 2;
                const int src = S.ssrc[q];
#pragma unroll
                for (int s = 0; s < CB; s++) {
                    if (s == s0) continue;
                    const int lc = (s - s0 + CB) % CB;
                    const double v = S.sb[src * W + NB * (lc - 1) + jj];
                    if (r == 0) acc[s][0] = v;
                    else if (r == 1) acc[s][1] = v;
                    else if (r == 2) acc[s][2] = v;
                    else acc[s][3] = v;
                }
            }
        }
        __syncthreads();

        // 4. trailing update / streaming out the finished panel
        if (lb != 0) {
            const int rowbase = NB * lb;
            double an[4];
#pragma unroll
            for (int kk = 0; kk < 4; kk++) an[kk] = -S.pnl[(rowbase + jj) * NB + 4 * kk + g];
#pragma unroll
            for (int s = 0; s < CB; s++) {
                if (s == s0) continue;
                const int lc = (s - s0 + CB) % CB;
                if (NB * (p + lc) > ju) continue;
                v4d c4 = acc[s];
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const double bv = S.u12[(4 * kk + g) * W + NB * (lc - 1) + jj];
                    c4 = __builtin_amdgcn_mfma_f64_16x16x4f64(an[kk], bv, c4, 0, 0, 0);
                }
                acc[s] = c4;
            }
            // column block p+CB enters the window in slot s0
            const int rb = p + lb;
#pragma unroll
            for (int s = 0; s < CB; s++)
                if (s == s0) {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        acc[s][r] = a_orig(dg, n, nx, NB * rb + crow(g, r), NB * (p + CB) + jj);
                }
        } else {
            // U rows j0..j0+nc-1: columns r..ext(r); beyond ext(r) the row is
            // zero (a zero-filled band stays exact)
            for (int e = l; e < NB * (NB + W); e += 64) {
                const int c = e % NB;
                const int col_off = e / NB;  // 0..NB+W-1, relative to j0
                if (c >= nc) continue;
                const int r = j0 + c;
                const int col = j0 + col_off;
                if (col < r || col > S.uext[c]) continue;
                const double v = (col_off < NB) ? S.pnl[c * NB + col_off] : S.u12[c * W + (col_off - NB)];
                ab[(size_t)col * ldab + kv + r - col] = v;
            }
            if (yv)
                for (int c = l; c < nc; c += 64) yv[j0 + c] = S.rw[(j0 + c) % PR];
            if (a.write_L) {
                // multipliers in dgbtf2 order: undo the later interchanges of
                // the panel on the columns left of each pivot step
                for (int c = nc - 1; c >= 0; c--) {
                    const int jp = S.jp[c];
                    if (jp != 0 && l < c) {
                        const double t = S.pnl2[c * NB + l];
                        S.pnl2[c * NB + l] = S.pnl2[(c + jp) * NB + l];
                        S.pnl2[(c + jp) * NB + l] = t;
                    }
                    wave_sync();
                }
                for (int c = 0; c < nc; c++) {
                    const int km = min(kl, n - 1 - (j0 + c));
                    for (int i = l; i < kl; i += 64)
                        ab[(size_t)(j0 + c) * ldab + kv + 1 + i] = (i < km) ? S.pnl2[(c + 1 + i) * NB + c] : 0.0;
                }
            }
            // the row block entering at the bottom, for panel p+1's window
            const int rb = p + RB;
#pragma unroll
            for (int s = 0; s < CB; s++) {
                const int cb = (p + 1) + ((s - (p + 1) % CB) + CB) % CB;
#pragma unroll
                for (int r = 0; r < 4; r++) acc[s][r] = a_orig(dg, n, nx, NB * rb + crow(g, r), NB * cb + jj);
            }
            for (int c = l; c < NB; c += 64) {
                const int i = NB * rb + c;
                S.rw[i % PR] = (rhs && i < n) ? rhs[i] : 0.0;
            }
        }
        __syncthreads();
    }
    if (tid == 0 && S.info != 0) raise_pivot(a.status, S.info, node);
}

// ---------------------------------------------------------------------------
// x = U^-1 y, one workgroup per node. U rows are read in blocks of 16 rows
// (only up to each row's extent), double-buffered through LDS one block
// ahead of the substitution; x stays in LDS. Within a block: the part of
// each row right of the block is reduced by 16 lanes per row, then the
// 16x16 triangle is solved by one wavefront, last row first.
// ---------------------------------------------------------------------------
constexpr int kBsThreads = 256;

__global__ __launch_bounds__(kBsThreads) void band_backsolve(const double* __restrict__ ab_all,
                                                             const double* __restrict__ y_all,
                                                             const int* __restrict__ uext_all,
                                                             double* __restrict__ x_all, int n, int kl, int ku,
                                                             int ldab, int nR, int include_boundary,
                                                             const int* __restrict__ nn,
                                                             const int* __restrict__ order) {
    const int node = order ? order[blockIdx.x] : blockIdx.x;
    if (nR > 0 && !include_boundary && (node % nR == 0 || node % nR == nR - 1)) return;
    const int stride = n;
    if (nn) n = nn[node];
    const int kv = kl + ku;
    const int WS = kv + NB;  // staged columns per block (relative to the block start)
    const double* ab = ab_all + (size_t)node * ldab * stride;
    const double* yv = y_all + (size_t)node * stride;
    const int* ext = uext_all + (size_t)node * stride;
    double* xo = x_all + (size_t)node * stride;
    extern __shared__ double sh[];
    double* x = sh;                       // n
    double* stg = sh + n;                 // 2 x [NB][WS]
    double* part = stg + 2 * NB * WS;     // [NB][16] partial sums
    __shared__ int sext[2][NB];
    const int tid = threadIdx.x;
    const int nblk = (n + NB - 1) / NB;

    auto stage = [&](int blk, int buf) {
        const int jb = NB * blk;
        double* st = stg + buf * NB * WS;
        for (int e = tid; e < NB * WS; e += kBsThreads) {
            const int ri = e % NB, co = e / NB;  // 16 consecutive lanes: 16 rows of one column
            const int r = jb + ri, col = jb + co;
            double v = 0.0;
            if (r < n && col >= r && col <= ext[r]) v = ab[(size_t)col * ldab + kv + r - col];
            st[ri * WS + co] = v;
        }
        if (tid < NB) sext[buf][tid] = (jb + tid < n) ? ext[jb + tid] : 0;
    };

    for (int i = tid; i < n; i += kBsThreads) x[i] = yv[i];
    stage(nblk - 1, (nblk - 1) & 1);
    __syncthreads();
    for (int blk = nblk - 1; blk >= 0; blk--) {
        const int buf = blk & 1;
        const int jb = NB * blk;
        const double* st = stg + buf * NB * WS;
        if (blk > 0) stage(blk - 1, buf ^ 1);
        // right-of-block part: row ri, lanes t = 0..15 stride over columns
        {
            const int ri = tid % NB, t = tid / NB;
            double s = 0.0;
            const int r = jb + ri;
            if (r < n) {
                const int hi = min(sext[buf][ri], n - 1) - jb;
                for (int co = NB + t; co <= hi; co += NB) s += st[ri * WS + co] * x[jb + co];
            }
            part[ri * 16 + t] = s;
        }
        __syncthreads();
        if (tid < 64) {
            const int ri = tid & 15;
            double b = 0.0;
            double urow[NB];
#pragma unroll
            for (int k = 0; k < NB; k++) urow[k] = 0.0;
            const int r = jb + ri;
            if (tid < NB && r < n) {
                double s = 0.0;
#pragma unroll
                for (int t = 0; t < 16; t++) s += part[ri * 16 + t];
                b = x[r] - s;
#pragma unroll
                for (int k = 0; k < NB; k++) urow[k] = st[ri * WS + k];
            }
#pragma unroll
            for (int k = NB - 1; k >= 0; k--) {
                double xk = 0.0;
                if (tid == k && jb + k < n) xk = b / urow[k];
                xk = __shfl(xk, k);
                if (tid < k && jb + tid < n) b -= urow[k] * xk;
                if (tid == k && jb + k < n) x[jb + k] = xk;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += kBsThreads) xo[i] = x[i];
}

int launch_backsolve(const double* ab, const double* y, const int* uext, double* x, int n, int kl, int ku, int ldab,
                     int nodes, int nR, int include_boundary, hipStream_t stream, const int* nn, const int* order) {
    const int WS = kl + ku + NB;
    const size_t lds = sizeof(double) * ((size_t)n + 2 * NB * WS + NB * 16);
    int dev = 0, max_lds = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (lds + 256 > (size_t)max_lds) return VERB_ERR_ARG;
    hipFuncSetAttribute((const void*)band_backsolve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(band_backsolve, dim3(nodes), dim3(kBsThreads), lds, stream, ab, y, uext, x, n, kl, ku, ldab, nR,
                       include_boundary, nn, order);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

template <int RB, int CB>
size_t band_lu_lds_bytes() {
    return sizeof(Lds<RB, CB>);
}

template <int RB, int CB>
int launch_band_lu_t(const LuArgs& a, int nodes, hipStream_t stream) {
    const size_t lds = band_lu_lds_bytes<RB, CB>();
    hipFuncSetAttribute((const void*)band_lu_window<RB, CB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((band_lu_window<RB, CB>), dim3(nodes), dim3(RB * 64), lds, stream, a);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

// Returns VERB_ERR_ARG when no register-window instantiation covers kl.
int launch_band_lu(const LuArgs& a, int nodes, hipStream_t stream) {
    const int kl = a.nx + 1;
    // RB*16 >= kl+16, CB*16 >= 2kl+16
    if (kl <= 32) return launch_band_lu_t<3, 5>(a, nodes, stream);
    if (kl <= 48) return launch_band_lu_t<4, 7>(a, nodes, stream);
    if (kl <= 64) return launch_band_lu_t<5, 9>(a, nodes, stream);
    if (kl <= 80) return launch_band_lu_t<6, 11>(a, nodes, stream);
    if (kl <= 96) return launch_band_lu_t<7, 13>(a, nodes, stream);
    if (kl <= 104) return launch_band_lu_t<8, 14>(a, nodes, stream);
    if (kl <= 112) return launch_band_lu_t<8, 15>(a, nodes, stream);
    return VERB_ERR_ARG;
}

}  // namespace verb
