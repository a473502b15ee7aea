// Local (E, alpha) diffusion: Diffusion_2D over every (P,R) node.
//
// Three kernels per call:
//   local_assemble  - the 9-point operator of src/Diffusion_2D.cpp:49-130
//                     (AddBoundaries_2D, src/MatrixSolver.cpp:118-213, and
//                     twelve SecondDerivativeApproximation_2D calls,
//                     src/MatrixSolver.cpp:439-494), written straight into
//                     LAPACK band storage (src/MatrixSolver.cpp:283-299).
//                     Compiled without FMA contraction: bit-identical to the
//                     reference assembly.
//   band_factor     - banded LU with LAPACK's partial pivoting (dgbtf2:
//                     first-largest |a| among rows j..j+kl of column j,
//                     row swap over columns j..ju, reciprocal scaling,
//                     rank-1 update), one workgroup per node.
//   band_solve      - num_substeps x { rhs = B0*rhs + C0; dgbtrs('N') }
//                     (src/MatrixSolver.cpp:309-326), one workgroup per node,
//                     the right-hand side resident in LDS.
// The unknown vector is energy-fast (in = iK*nV + iV, src/Matrix.cpp:6479),
// so kl = ku = nV+1 as in the reference.
#include "common.h"
#include "band_lu.h"
#include "gmres.h"

namespace verb {
namespace {

constexpr int kAsmThreads = 256;

// Node-level view of the grid slices that Diffusion_2D receives.
struct NodeArgs {
    int nx, ny;          // V, active K extent of this node
    int ld;              // row stride of the [V][K] node arrays (full K)
    double dt;           // sub-step dt
    const double *x, *y, *Dxx, *Dyy, *Dxy, *G, *S, *L;  // [nx][ld], first active K column
    const double *xl, *xu, *yl, *yu;                     // ny, ny, nx, nx
    int xlt, xut, ylt, yut;
    double src_mult;
};

__device__ inline void sda2(double A[9], const NodeArgs& a, int ix, int iy, int dx1, int dy1, int dx2, int dy2,
                            const double* D, double mult) {
    const int ld = a.ld;
#define AT(arr, i, j) (arr)[(size_t)(i) * ld + (j)]
    const double dh1 = (AT(a.x, ix, iy) + AT(a.y, ix, iy)) - (AT(a.x, ix + dx1, iy + dy1) + AT(a.y, ix + dx1, iy + dy1));
    const double dh21 = (AT(a.x, ix, iy) + AT(a.y, ix, iy)) - (AT(a.x, ix + dx2, iy + dy2) + AT(a.y, ix + dx2, iy + dy2));
    const double dh22 = (AT(a.x, ix + dx1, iy + dy1) + AT(a.y, ix + dx1, iy + dy1)) -
                        (AT(a.x, ix + dx1 + dx2, iy + dy1 + dy2) + AT(a.y, ix + dx1 + dx2, iy + dy1 + dy2));
    const double common_part = mult / AT(a.G, ix, iy) / dh1;
    A[4] += +common_part / dh21 * AT(D, ix, iy) * AT(a.G, ix, iy);
    A[(dy2 + 1) * 3 + (dx2 + 1)] += -common_part / dh21 * AT(D, ix, iy) * AT(a.G, ix, iy);
    A[(dy1 + 1) * 3 + (dx1 + 1)] += -common_part / dh22 * AT(D, ix + dx1, iy + dy1) * AT(a.G, ix + dx1, iy + dy1);
    A[(dy1 + dy2 + 1) * 3 + (dx1 + dx2 + 1)] +=
        +common_part / dh22 * AT(D, ix + dx1, iy + dy1) * AT(a.G, ix + dx1, iy + dy1);
#undef AT
}

// Row `in` of A (9 slots, slot (dy+1)*3+(dx+1) for offset dy*nx+dx), B0, C0.
// Returns false on an unknown boundary type.
__device__ inline bool assemble_row(const NodeArgs& a, int ix, int iy, double A[9], double& B0, double& C0) {
#pragma unroll
    for (int s = 0; s < 9; s++) A[s] = 0.0;
    B0 = 0.0;
    C0 = 0.0;
    const int nx = a.nx, ny = a.ny, ld = a.ld;
    if (ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1) {
        int type, s1;
        double dh;
        if (ix == 0) {
            C0 = a.xl[iy]; s1 = 5; dh = a.x[(size_t)(ix + 1) * ld + iy] - a.x[(size_t)ix * ld + iy]; type = a.xlt;
        } else if (ix == nx - 1) {
            C0 = a.xu[iy]; s1 = 3; dh = a.x[(size_t)ix * ld + iy] - a.x[(size_t)(ix - 1) * ld + iy]; type = a.xut;
        } else if (iy == 0) {
            C0 = a.yl[ix]; s1 = 7; dh = a.y[(size_t)ix * ld + iy + 1] - a.y[(size_t)ix * ld + iy]; type = a.ylt;
        } else {
            C0 = a.yu[ix]; s1 = 1; dh = a.y[(size_t)ix * ld + iy] - a.y[(size_t)ix * ld + iy - 1]; type = a.yut;
        }
        if (type == VERB_BC_CONSTANT_VALUE) { A[4] = 1; A[s1] = 0; }
        else if (type == VERB_BC_CONSTANT_DERIVATIVE) { A[4] = 1 / dh; A[s1] = -1 / dh; }
        else return false;
        return true;
    }
    const size_t p = (size_t)ix * ld + iy;
    A[4] += 1.0 / a.dt;
    B0 += 1.0 / a.dt;
    C0 += a.S[p] * a.src_mult;
    A[4] -= a.L[p];
    sda2(A, a, ix, iy, -1, 0, +1, 0, a.Dxx, -0.5);
    sda2(A, a, ix, iy, +1, 0, -1, 0, a.Dxx, -0.5);
    sda2(A, a, ix, iy, 0, -1, 0, +1, a.Dyy, -0.5);
    sda2(A, a, ix, iy, 0, +1, 0, -1, a.Dyy, -0.5);
    sda2(A, a, ix, iy, 0, -1, +1, 0, a.Dxy, -0.25);
    sda2(A, a, ix, iy, 0, +1, -1, 0, a.Dxy, -0.25);
    sda2(A, a, ix, iy, 0, -1, -1, 0, a.Dxy, -0.25);
    sda2(A, a, ix, iy, 0, +1, +1, 0, a.Dxy, -0.25);
    sda2(A, a, ix, iy, -1, 0, 0, +1, a.Dxy, -0.25);  // Dyx is DVK too (VERB4D_Solver.cpp:1150)
    sda2(A, a, ix, iy, +1, 0, 0, -1, a.Dxy, -0.25);
    sda2(A, a, ix, iy, -1, 0, 0, -1, a.Dxy, -0.25);
    sda2(A, a, ix, iy, +1, 0, 0, +1, a.Dxy, -0.25);
    return true;
}

// k0: first active pitch-angle column of the node; the node's system is the
// reference's Diffusion_2D on the slice [V][k0:K] (y_size = K - k0).
__device__ inline NodeArgs node_args(const verb_grid& g, int iP, int iR, double dt_sub, int k0) {
    NodeArgs a;
    a.nx = g.nV; a.ny = g.nK - k0; a.ld = g.nK; a.dt = dt_sub;
    const size_t o = ((size_t)iP * g.nR + iR) * g.nV * g.nK + k0;
    a.x = g.Vg + o; a.y = g.Kg + o; a.Dxx = g.DVV + o; a.Dyy = g.DKK + o; a.Dxy = g.DVK + o;
    a.G = g.G_local + o; a.S = g.sources + o; a.L = g.losses_local + o;
    a.xl = g.Vl + ((size_t)iP * g.nR + iR) * g.nK + k0;
    a.xu = g.Vu + ((size_t)iP * g.nR + iR) * g.nK + k0;
    a.yl = g.Kl + ((size_t)iP * g.nR + iR) * g.nV;
    a.yu = g.Ku + ((size_t)iP * g.nR + iR) * g.nV;
    a.xlt = g.Vl_type; a.xut = g.Vu_type; a.ylt = g.Kl_type; a.yut = g.Ku_type;
    a.src_mult = g.local_sources;
    return a;
}

__device__ inline bool node_skipped(const verb_grid& g, int iR) {
    return !g.include_boundary && (iR == 0 || iR == g.nR - 1);
}

// One thread per grid point of one node (blockIdx.y = node). Writes the
// nonzero band entries of row `in` (the caller zeroed the band), B0, C0
// and the right-hand-side start vector psd in energy-fast order.
__global__ __launch_bounds__(kAsmThreads) void local_assemble(const verb_grid g, const double* __restrict__ psd,
                                                             double* __restrict__ ab, double* __restrict__ b0,
                                                             double* __restrict__ c0, double* __restrict__ rhs,
                                                             double* __restrict__ diag, double dt_sub,
                                                             int write_band, int premul, int* status,
                                                             const int* __restrict__ k0s) {
    const int node = blockIdx.y;
    const int iP = node / g.nR, iR = node % g.nR;
    if (node_skipped(g, iR)) return;
    const int k0 = k0s ? k0s[node] : 0;
    const int nx = g.nV, ny = g.nK - k0, n = nx * ny;
    const int nfull = g.nV * g.nK;  // per-node stride of every workspace buffer
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int ix = p / ny, iy = p % ny;
    const int in = iy * nx + ix;
    const NodeArgs a = node_args(g, iP, iR, dt_sub, k0);
    double A[9], B0, C0;
    if (!assemble_row(a, ix, iy, A, B0, C0)) raise_error(status, VERB_ERR_BC);
    const size_t nn = (size_t)node * nfull;
    const size_t pp = nn + (size_t)ix * g.nK + k0 + iy;  // this point in the [V][K] node block
    b0[nn + in] = B0;
    c0[nn + in] = C0;
    if (premul) {
        // first sub-step's right-hand side: rhs = rhs*B0 + C0 (MatrixSolver.cpp:318-319)
        double v = psd[pp] * B0;
        v += C0;
        rhs[nn + in] = v;
    } else {
        rhs[nn + in] = psd[pp];
    }
    if (diag) {
        double* dn = diag + (size_t)node * 9 * nfull;
#pragma unroll
        for (int sl = 0; sl < 9; sl++) dn[(size_t)sl * n + in] = A[sl];
    }
    if (write_band) {
        const int kl = nx + 1, ku = nx + 1, ldab = 2 * kl + ku + 1;
        double* abn = ab + (size_t)node * ldab * nfull;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int off = dy * nx + dx;
                const int col = in + off;
                if (col >= 0 && col < n) abn[(size_t)col * ldab + (kl + ku - off)] = A[(dy + 1) * 3 + (dx + 1)];
            }
    }
}

// The same assembly when no LAPACK band is written, by 16x16 tiles of
// (ix, iy): each point is computed by the lane whose [V][K] reads are
// consecutive in iy and written by the lane whose energy-fast outputs are
// consecutive in ix, the tile being transposed through LDS.
constexpr int kAsmTile = 16;

__global__ __launch_bounds__(kAsmTile* kAsmTile) void local_assemble_tiled(
    const verb_grid g, const double* __restrict__ psd, double* __restrict__ b0, double* __restrict__ c0,
    double* __restrict__ rhs, double* __restrict__ diag, double dt_sub, int premul, int* status,
    const int* __restrict__ k0s, int tiles, int per_xcd) {
    // XCD-aware order: consecutive workgroup ids go to different XCDs in
    // turn, so each XCD takes a contiguous run of (node, tile) pairs and the
    // stencil halos of neighbouring tiles of a node meet in one L2
    const int lin = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (lin >= tiles * g.nP * g.nR) return;
    const int node = lin / tiles, tile = lin - node * tiles;
    const int iP = node / g.nR, iR = node % g.nR;
    if (node_skipped(g, iR)) return;
    const int k0 = k0s ? k0s[node] : 0;
    const int nx = g.nV, ny = g.nK - k0, n = nx * ny;
    const int nfull = g.nV * g.nK;
    const int tiles_y = (g.nK + kAsmTile - 1) / kAsmTile;
    const int ix0 = (tile / tiles_y) * kAsmTile, iy0 = (tile % tiles_y) * kAsmTile;
    if (ix0 >= nx || iy0 >= ny) return;
    const int tx = threadIdx.x % kAsmTile, ty = threadIdx.x / kAsmTile;
    __shared__ double out[12][kAsmTile][kAsmTile + 1];
    const size_t nn = (size_t)node * nfull;
    {
        const int ix = ix0 + ty, iy = iy0 + tx;
        if (ix < nx && iy < ny) {
            const NodeArgs a = node_args(g, iP, iR, dt_sub, k0);
            double A[9], B0, C0;
            if (!assemble_row(a, ix, iy, A, B0, C0)) raise_error(status, VERB_ERR_BC);
            const size_t pp = nn + (size_t)ix * g.nK + k0 + iy;
            double v;
            if (premul) {
                // first sub-step's right-hand side: rhs = rhs*B0 + C0 (MatrixSolver.cpp:318-319)
                v = psd[pp] * B0;
                v += C0;
            } else {
                v = psd[pp];
            }
#pragma unroll
            for (int sl = 0; sl < 9; sl++) out[sl][ty][tx] = A[sl];
            out[9][ty][tx] = B0;
            out[10][ty][tx] = C0;
            out[11][ty][tx] = v;
        }
    }
    __syncthreads();
    const int ix = ix0 + tx, iy = iy0 + ty;
    if (ix < nx && iy < ny) {
        const int in = iy * nx + ix;
        if (b0) b0[nn + in] = out[9][tx][ty];
        if (c0) c0[nn + in] = out[10][tx][ty];
        rhs[nn + in] = out[11][tx][ty];
        if (diag) {
            double* dn = diag + (size_t)node * 9 * nfull;
#pragma unroll
            for (int sl = 0; sl < 9; sl++) dn[(size_t)sl * n + in] = out[sl][tx][ty];
        }
    }
}

void launch_assemble(const verb_grid& g, const double* psd, double* ab, double* b0, double* c0, double* rhs,
                     double* diag, double dt_sub, int write_band, int premul, int* status, const int* k0s,
                     hipStream_t stream) {
    const int nodes = g.nP * g.nR;
    if (write_band) {
        const int n = g.nV * g.nK;
        hipLaunchKernelGGL(local_assemble, dim3((n + kAsmThreads - 1) / kAsmThreads, nodes), dim3(kAsmThreads), 0,
                           stream, g, psd, ab, b0, c0, rhs, diag, dt_sub, write_band, premul, status, k0s);
    } else {
        const int tiles = ((g.nV + kAsmTile - 1) / kAsmTile) * ((g.nK + kAsmTile - 1) / kAsmTile);
        const int per_xcd = (tiles * nodes + 7) / 8;
        hipLaunchKernelGGL(local_assemble_tiled, dim3(8 * per_xcd), dim3(kAsmTile * kAsmTile), 0, stream, g, psd,
                           b0, c0, rhs, diag, dt_sub, premul, status, k0s, tiles, per_xcd);
    }
}

__global__ void fill_zero(double* __restrict__ p, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Banded LU (dgbtf2 semantics), one workgroup per system.
// ---------------------------------------------------------------------------
constexpr int kLuThreads = 256;

__global__ __launch_bounds__(kLuThreads) void band_factor_kernel(double* __restrict__ ab_all,
                                                                 int64_t* __restrict__ ipiv_all, size_t ab_stride,
                                                                 int n, int kl, int ku, int ldab, int* status,
                                                                 int nR, int include_boundary) {
    const int node = blockIdx.x;
    if (nR > 0 && !include_boundary && (node % nR == 0 || node % nR == nR - 1)) return;
    double* ab = ab_all + (size_t)node * ab_stride;
    int64_t* ipiv = ipiv_all + (size_t)node * n;
    const int kv = ku + kl;
    const int tid = threadIdx.x;
    extern __shared__ double sh[];
    double* sL = sh;                  // kl + 1
    double* sU = sh + (kl + 1);       // kl + ku + 1
    __shared__ int s_jp;
    __shared__ double s_val[kLuThreads / 64];
    __shared__ int s_idx[kLuThreads / 64];
#define AB(i, j) ab[(size_t)(j) * ldab + (i)]
    // zero the fill-in rows of columns ku+1 .. min(kv,n)-1 (dgbtf2 prologue)
    for (int j = ku + 1; j <= min(kv, n) - 1; j++)
        for (int i = kv - j + tid; i < kl; i += kLuThreads) AB(i, j) = 0.0;
    int ju = 0;
    int info = 0;
    __syncthreads();
    for (int j = 0; j < n; j++) {
        if (j + kv < n)
            for (int i = tid; i < kl; i += kLuThreads) AB(i, j + kv) = 0.0;
        const int km = min(kl, n - 1 - j);
        // first index of max |AB(kv+i, j)|, i in [0, km]
        double best = -1.0;
        int bidx = 0x7fffffff;
        for (int i = tid; i <= km; i += kLuThreads) {
            const double v = fabs(AB(kv + i, j));
            if (v > best) { best = v; bidx = i; }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double ov = __shfl_xor(best, off);
            const int oi = __shfl_xor(bidx, off);
            if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
        }
        if ((tid & 63) == 0) { s_val[tid >> 6] = best; s_idx[tid >> 6] = bidx; }
        __syncthreads();
        if (tid == 0) {
            double b = s_val[0];
            int bi = s_idx[0];
            for (int w = 1; w < kLuThreads / 64; w++)
                if (s_val[w] > b || (s_val[w] == b && s_idx[w] < bi)) { b = s_val[w]; bi = s_idx[w]; }
            // no comparable entry (all NaN): idamax returns the first index
            s_jp = (bi > km) ? 0 : bi;
        }
        __syncthreads();
        const int jp = s_jp;
        if (tid == 0) ipiv[j] = (int64_t)(jp + j + 1);
        const double piv = AB(kv + jp, j);
        if (piv != 0.0) {
            ju = max(ju, min(j + ku + jp, n - 1));
            if (jp != 0) {
                for (int c = j + tid; c <= ju; c += kLuThreads) {
                    const double t = AB(kv + jp + j - c, c);
                    AB(kv + jp + j - c, c) = AB(kv + j - c, c);
                    AB(kv + j - c, c) = t;
                }
            }
            __syncthreads();
            if (km > 0) {
                const double r = 1.0 / AB(kv, j);
                for (int i = 1 + tid; i <= km; i += kLuThreads) {
                    const double v = AB(kv + i, j) * r;
                    AB(kv + i, j) = v;
                    sL[i] = v;
                }
                for (int c = j + 1 + tid; c <= ju; c += kLuThreads) sU[c - j] = AB(kv + j - c, c);
                __syncthreads();
                const int w = ju - j;
                const int total = w * km;
                for (int e = tid; e < total; e += kLuThreads) {
                    const int cc = e / km;
                    const int i = 1 + (e - cc * km);
                    const int c = j + 1 + cc;
                    const double temp = -sU[cc + 1];
                    if (temp != 0.0) AB(kv + j + i - c, c) += sL[i] * temp;
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        __syncthreads();
    }
#undef AB
    if (tid == 0 && info != 0) raise_pivot(status, info, node);
}

// ---------------------------------------------------------------------------
// Band solve: per node num_substeps x { x = B0*x + C0; dgbtrs('N') }.
// One wavefront per system; the solution vector lives in LDS, band
// columns are staged through LDS in chunks of kChunk columns.
// ---------------------------------------------------------------------------
constexpr int kSolveThreads = 64;

__global__ __launch_bounds__(kSolveThreads) void band_solve_kernel(
    const double* __restrict__ ab_all, const int64_t* __restrict__ ipiv_all, size_t ab_stride, int n, int kl,
    int ku, int ldab, const double* __restrict__ b0_all, const double* __restrict__ c0_all,
    double* __restrict__ rhs_all, int nsub, int chunk, int nR, int include_boundary,
    const int* __restrict__ nn) {
    const int node = blockIdx.x;
    if (nR > 0 && !include_boundary && (node % nR == 0 || node % nR == nR - 1)) return;
    const int stride = n;  // per-node buffer stride (the full node size)
    if (nn) n = nn[node];
    const double* ab = ab_all + (size_t)node * ab_stride;
    const int64_t* ipiv = ipiv_all + (size_t)node * stride;
    double* rhs = rhs_all + (size_t)node * stride;
    const double* B0 = b0_all ? b0_all + (size_t)node * stride : nullptr;
    const double* C0 = c0_all ? c0_all + (size_t)node * stride : nullptr;
    const int tid = threadIdx.x;
    extern __shared__ double sh[];
    double* x = sh;           // n
    double* stage = sh + n;   // chunk * (kl+ku+1)
    const int kd = ku + kl;   // 0-based row of the diagonal
    const int k = kl + ku;
    for (int i = tid; i < n; i += kSolveThreads) x[i] = rhs[i];
    __syncthreads();
    for (int s = 0; s < nsub; s++) {
        if (B0) {
            for (int i = tid; i < n; i += kSolveThreads) {
                double v = x[i] * B0[i];
                v += C0[i];
                x[i] = v;
            }
        }
        __syncthreads();
        // forward: L with row interchanges (dgbtrs lines 'DO 10 J = 1, N-1')
        if (kl > 0) {
            for (int j0 = 0; j0 < n - 1; j0 += chunk) {
                const int jn = min(chunk, n - 1 - j0);
                // stage L columns j0..j0+jn-1, rows kd+1..kd+kl
                for (int e = tid; e < jn * kl; e += kSolveThreads) {
                    const int cj = e / kl, i = e - cj * kl;
                    stage[cj * kl + i] = ab[(size_t)(j0 + cj) * ldab + kd + 1 + i];
                }
                __syncthreads();
                for (int cj = 0; cj < jn; cj++) {
                    const int j = j0 + cj;
                    const int lm = min(kl, n - 1 - j);
                    const int l = (int)ipiv[j] - 1;
                    if (l != j && tid == 0) {
                        const double t = x[l];
                        x[l] = x[j];
                        x[j] = t;
                    }
                    __syncthreads();
                    const double temp = -x[j];
                    if (temp != 0.0)
                        for (int i = 1 + tid; i <= lm; i += kSolveThreads) x[j + i] += stage[cj * kl + i - 1] * temp;
                    __syncthreads();
                }
            }
        }
        // backward: dtbsv upper, no-transpose, non-unit
        for (int jhi = n - 1; jhi >= 0; jhi -= chunk) {
            const int jn = min(chunk, jhi + 1);
            // stage U columns jhi-jn+1 .. jhi, rows 0..kd
            for (int e = tid; e < jn * (k + 1); e += kSolveThreads) {
                const int cj = e / (k + 1), i = e - cj * (k + 1);
                stage[cj * (k + 1) + i] = ab[(size_t)(jhi - cj) * ldab + i];
            }
            __syncthreads();
            for (int cj = 0; cj < jn; cj++) {
                const int j = jhi - cj;
                const double xj = x[j];
                if (xj != 0.0) {
                    const double v = xj / stage[cj * (k + 1) + k];
                    __syncthreads();
                    if (tid == 0) x[j] = v;
                    const int lo = max(0, j - k);
                    for (int i = j - 1 - tid; i >= lo; i -= kSolveThreads)
                        x[i] = x[i] - v * stage[cj * (k + 1) + k + i - j];
                }
                __syncthreads();
            }
        }
    }
    for (int i = tid; i < n; i += kSolveThreads) rhs[i] = x[i];
}

namespace {

// rhs (energy-fast) -> psd slice (K-fast) for every node
__global__ void scatter_psd(const verb_grid g, const double* __restrict__ rhs, double* __restrict__ psd,
                            const int* __restrict__ k0s) {
    const int node = blockIdx.y;
    const int iR = node % g.nR;
    if (node_skipped(g, iR)) return;
    const int k0 = k0s ? k0s[node] : 0;
    const int nx = g.nV, ny = g.nK - k0, n = nx * ny, nfull = g.nV * g.nK;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int ix = p / ny, iy = p % ny;
    psd[(size_t)node * nfull + (size_t)ix * g.nK + k0 + iy] = rhs[(size_t)node * nfull + iy * nx + ix];
}

// psd slice (K-fast) -> energy-fast vector, the inverse of scatter_psd
__global__ void gather_psd(const verb_grid g, const double* __restrict__ psd, double* __restrict__ x) {
    const int node = blockIdx.y;
    const int iR = node % g.nR;
    if (node_skipped(g, iR)) return;
    const int nx = g.nV, ny = g.nK, n = nx * ny, nfull = n;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int ix = p / ny, iy = p % ny;
    x[(size_t)node * nfull + iy * nx + ix] = psd[(size_t)node * nfull + (size_t)ix * g.nK + iy];
}

// ---- mixed precision (FP32 factors, FP64 refinement) ----

constexpr int kMxThreads = 256;

// d32 = (float)diag, x = 0, done = skipped node, sweeps = 0, bmax = max|b|
__global__ __launch_bounds__(kMxThreads) void mixed_init(const verb_grid g, const double* __restrict__ diag,
                                                         const double* __restrict__ b, float* __restrict__ d32,
                                                         double* __restrict__ x, double* __restrict__ bmax,
                                                         int* __restrict__ done, int* __restrict__ sweeps) {
    const int node = blockIdx.x;
    const int iR = node % g.nR;
    const int n = g.nV * g.nK;
    const size_t base = (size_t)node * n;
    if (node_skipped(g, iR)) {
        if (threadIdx.x == 0) {
            done[node] = 1;
            sweeps[node] = 0;
            bmax[node] = 0.0;
        }
        return;
    }
    for (int i = threadIdx.x; i < 9 * n; i += kMxThreads) d32[9 * base + i] = (float)diag[9 * base + i];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += kMxThreads) {
        x[base + i] = 0.0;
        m = fmax(m, fabs(b[base + i]));
    }
    __shared__ double red[kMxThreads / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kMxThreads / 64; q++) t = fmax(t, red[q]);
        bmax[node] = t;
        done[node] = 0;
        sweeps[node] = 0;
    }
}

// r = b - A x in FP64 for the nodes still refining; a node whose
// max|r| <= tol * max|b| is done. count: whether a solve follows.
__global__ __launch_bounds__(kMxThreads) void mixed_residual(const verb_grid g, const double* __restrict__ diag,
                                                             const double* __restrict__ b,
                                                             const double* __restrict__ x, float* __restrict__ r32,
                                                             const double* __restrict__ bmax, double tol,
                                                             double* __restrict__ rres, int* __restrict__ done,
                                                             int* __restrict__ sweeps, int count) {
    const int node = blockIdx.x;
    if (done[node]) return;
    const int nx = g.nV, n = g.nV * g.nK;
    const size_t base = (size_t)node * n;
    const double* dg = diag + 9 * base;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += kMxThreads) {
        double acc = b[base + i];
#pragma unroll
        for (int s = 0; s < 9; s++) {
            const int j = i + (s / 3 - 1) * nx + (s % 3 - 1);
            if (j >= 0 && j < n) acc = fma(-dg[(size_t)s * n + i], x[base + j], acc);
        }
        r32[base + i] = (float)acc;
        m = (acc == acc) ? fmax(m, fabs(acc)) : (double)INFINITY;  // a NaN residual never converges
    }
    __shared__ double red[kMxThreads / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kMxThreads / 64; q++) t = fmax(t, red[q]);
        const double bm = bmax[node];
        rres[node] = bm > 0.0 ? t / bm : t;
        if (t <= tol * bm) done[node] = 1;
        else if (count) sweeps[node] += 1;
    }
}

__global__ void mixed_update(const verb_grid g, double* __restrict__ x, const float* __restrict__ c,
                             const int* __restrict__ done) {
    const int node = blockIdx.y;
    if (done[node]) return;
    const int n = g.nV * g.nK;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[(size_t)node * n + i] += (double)c[(size_t)node * n + i];
}

int solve_chunk(int n, int width) {
    int dev = 0, max_lds = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    const long avail = (long)max_lds - (long)n * 8;
    long ch = avail / ((long)width * 8);
    if (ch > 32) ch = 32;
    return (int)ch;
}

}  // namespace

int launch_local(const verb_grid& g, double* psd, const LocalWork& w, bool factor, bool cached, int* status,
                 hipStream_t stream, hipEvent_t* marks, const NodeSizes* sizes) {
    const int* k0s = sizes ? sizes->k0 : nullptr;
    const int* nns = sizes ? sizes->n : nullptr;
    const int* order = sizes ? sizes->order : nullptr;
    if (g.nV < 3 || g.nK < 3) return VERB_ERR_ARG;
    const int nx = g.nV, ny = g.nK, n = nx * ny;
    const int kl = nx + 1, ku = nx + 1, ldab = 2 * kl + ku + 1;
    const int nodes = g.nP * g.nR;
    const int num_substeps = max(1, (int)(g.dt / g.sub_dt));
    const double dt_sub = g.dt / num_substeps;
    const size_t ab_stride = (size_t)ldab * n;
    const int chunk = solve_chunk(n, kl + ku + 1);
    if (chunk < 1) return VERB_ERR_ARG;
    dim3 agrid((n + kAsmThreads - 1) / kAsmThreads, nodes);
    const bool window = local_lu_variant == 0 && kl <= 112 && w.diag && w.y && w.uext;
    if (window && factor && num_substeps == 1 && !cached) {
        // plain path: factor with the forward elimination fused, then U x = y
        if (marks) hipEventRecord(marks[0], stream);
        // (B0 and C0 are read only by the sub-stepped and cached solves)
        launch_assemble(g, psd, w.ab, nullptr, nullptr, w.rhs, w.diag, dt_sub, 0, 1, status, k0s, stream);
        if (marks) {
            hipEventRecord(marks[1], stream);
            hipEventRecord(marks[2], stream);
        }
        LuArgs la;
        la.diag = w.diag; la.rhs = w.rhs; la.y = w.y; la.ab = w.ab; la.ipiv = w.ipiv; la.uext = w.uext;
        la.n = n; la.nx = nx; la.write_L = 0; la.status = status; la.nR = g.nR;
        la.include_boundary = g.include_boundary;
        la.nn = nns; la.order = order;
        la.queue = w.queue; la.qnodes = nodes;
        // U goes out in 16-row blocks (the factors are not kept), which the
        // back-substitution streams
        if (!ublk_geometry(n, kl, ku, &la.ublk_cols, &la.ublk_stride)) la.ublk_cols = 0;
        // persistent launch with blocked U: the back-substitution runs in the
        // same kernel, on the workgroups that run out of nodes to factor
        const bool fused = w.queue && la.ublk_cols > 0;
        // ... and, when that launch is wide enough to hold it, writes x
        // straight back into the PSD (no scatter pass)
        const bool to_psd = fused && band_lu_fuses_backsolve(w.lu_variant, kl);
        if (fused) {
            la.bs_queue = w.queue + 1;
            la.bs_flags = w.queue + 2;
            la.bs_x = w.rhs;
        }
        if (to_psd) {
            la.bs_psd = psd;
            la.bs_k0 = k0s;
            la.bs_nK = g.nK;
        }
        int rc = launch_band_lu(la, nodes, stream, w.lu_variant);
        if (rc) return rc;
        if (marks) {
            hipEventRecord(marks[3], stream);
            hipEventRecord(marks[4], stream);
        }
        if (!fused) {
            rc = launch_backsolve(w.ab, w.y, w.uext, w.rhs, n, kl, ku, ldab, nodes, g.nR, g.include_boundary, stream,
                                  nns, order, nullptr, la.ublk_cols, la.ublk_stride);
            if (rc) return rc;
        }
        if (!to_psd) hipLaunchKernelGGL(scatter_psd, agrid, dim3(kAsmThreads), 0, stream, g, w.rhs, psd, k0s);
        if (marks) hipEventRecord(marks[5], stream);
        return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
    }
    if (marks) hipEventRecord(marks[0], stream);
    if (factor) {
        const size_t count = ab_stride * nodes;
        hipLaunchKernelGGL(fill_zero, dim3(4096), dim3(256), 0, stream, w.ab, count);
    }
    if (sizes && !window) return VERB_ERR_ARG;  // uneven extents need the register-window LU
    launch_assemble(g, psd, w.ab, w.b0, w.c0, w.rhs, (window && factor) ? w.diag : (double*)nullptr, dt_sub,
                    (factor && !window) ? 1 : 0, 0, status, k0s, stream);
    if (marks) {
        hipEventRecord(marks[1], stream);
        hipEventRecord(marks[2], stream);
    }
    if (factor && window) {
        // cached factors / several sub-steps: keep L (dgbtf2 order) as well
        LuArgs la;
        la.diag = w.diag; la.rhs = nullptr; la.y = nullptr; la.ab = w.ab; la.ipiv = w.ipiv; la.uext = w.uext;
        la.n = n; la.nx = nx; la.write_L = 1; la.status = status; la.nR = g.nR;
        la.include_boundary = g.include_boundary;
        la.nn = nns; la.order = order;
        la.queue = w.queue; la.qnodes = nodes;
        int rc = launch_band_lu(la, nodes, stream, w.lu_variant);
        if (rc) return rc;
    } else if (factor) {
        const size_t lds = sizeof(double) * (size_t)(kl + 1 + kl + ku + 1);
        hipLaunchKernelGGL(band_factor_kernel, dim3(nodes), dim3(kLuThreads), lds, stream, w.ab, w.ipiv, ab_stride, n,
                           kl, ku, ldab, status, g.nR, g.include_boundary);
    }
    if (marks) {
        hipEventRecord(marks[3], stream);
        hipEventRecord(marks[4], stream);
    }
    if (window && !sizes) {
        // per sub-step: y = L^-1 P (B0*x + C0), then x = U^-1 y
        for (int s = 0; s < num_substeps; s++) {
            int rc = launch_forward(w.ab, w.ipiv, n, kl, ku, ldab, w.b0, w.c0, w.rhs, w.y, nodes, g.nR,
                                    g.include_boundary, stream);
            if (rc) return rc;
            rc = launch_backsolve(w.ab, w.y, w.uext, w.rhs, n, kl, ku, ldab, nodes, g.nR, g.include_boundary, stream);
            if (rc) return rc;
        }
    } else {
        const size_t slds = sizeof(double) * ((size_t)n + (size_t)chunk * (kl + ku + 1));
        hipFuncSetAttribute((const void*)band_solve_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slds);
        hipLaunchKernelGGL(band_solve_kernel, dim3(nodes), dim3(kSolveThreads), slds, stream, w.ab, w.ipiv,
                           ab_stride, n, kl, ku, ldab, w.b0, w.c0, w.rhs, num_substeps, chunk, g.nR,
                           g.include_boundary, nns);
    }
    hipLaunchKernelGGL(scatter_psd, agrid, dim3(kAsmThreads), 0, stream, g, w.rhs, psd, k0s);
    if (marks) hipEventRecord(marks[5], stream);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_local_assembly(const verb_grid& g, const double* psd, double* diag, double* b0, double* c0, double* rhs,
                          int* status, hipStream_t stream) {
    if (g.nV < 3 || g.nK < 3) return VERB_ERR_ARG;
    const int num_substeps = max(1, (int)(g.dt / g.sub_dt));
    launch_assemble(g, psd, nullptr, b0, c0, rhs, diag, g.dt / num_substeps, 0, 1, status, nullptr, stream);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_dgbtrf(int n, int kl, int ku, double* ab, int ldab, int64_t* ipiv, int* status, hipStream_t stream) {
    if (n < 1 || kl < 0 || ku < 0 || ldab < 2 * kl + ku + 1) return VERB_ERR_ARG;
    const size_t lds = sizeof(double) * (size_t)(kl + 1 + kl + ku + 1);
    hipLaunchKernelGGL(band_factor_kernel, dim3(1), dim3(kLuThreads), lds, stream, ab, ipiv, (size_t)0, n, kl, ku,
                       ldab, status, 0, 1);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_dgbtrs(int n, int kl, int ku, const double* ab, int ldab, const int64_t* ipiv, double* b,
                  hipStream_t stream) {
    if (n < 1 || kl < 0 || ku < 0 || ldab < 2 * kl + ku + 1) return VERB_ERR_ARG;
    const int chunk = solve_chunk(n, kl + ku + 1);
    if (chunk < 1) return VERB_ERR_ARG;
    const size_t slds = sizeof(double) * ((size_t)n + (size_t)chunk * (kl + ku + 1));
    hipFuncSetAttribute((const void*)band_solve_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slds);
    hipLaunchKernelGGL(band_solve_kernel, dim3(1), dim3(kSolveThreads), slds, stream, ab, ipiv, (size_t)0, n, kl, ku,
                       ldab, (const double*)nullptr, (const double*)nullptr, b, 1, chunk, 0, 1, (const int*)nullptr);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}


int launch_local_gmres(const verb_grid& g, double* psd, const LocalWork& w, const GmresWork& gw, double tol,
                       int restart, int max_cycles, int max_iters, double inner_tol, const double* x0, int* status,
                       hipStream_t stream, hipEvent_t* marks) {
    if (g.nV < 3 || g.nK < 3) return VERB_ERR_ARG;
    const int nx = g.nV, ny = g.nK, n = nx * ny;
    const int kl = nx + 1, ku = nx + 1, ldab = 2 * kl + ku + 1;
    const int nodes = g.nP * g.nR;
    const int num_substeps = max(1, (int)(g.dt / g.sub_dt));
    if (num_substeps != 1) return VERB_ERR_ARG;
    if (kl > 112) return VERB_ERR_ARG;  // the fallback needs the register-window LU
    const double dt_sub = g.dt / num_substeps;
    dim3 agrid((n + kAsmThreads - 1) / kAsmThreads, nodes);
    if (marks) hipEventRecord(marks[0], stream);
    launch_assemble(g, psd, w.ab, w.b0, w.c0, w.rhs, w.diag, dt_sub, 0, 1, status, nullptr, stream);
    // initial iterate: the caller's x0 (same [P][R][V][K] layout as psd) or the incoming PSD
    hipLaunchKernelGGL(gather_psd, agrid, dim3(kAsmThreads), 0, stream, g, x0 ? x0 : psd, gw.x);
    if (marks) {
        hipEventRecord(marks[1], stream);
        hipEventRecord(marks[2], stream);
    }
    GmArgs ga;
    ga.diag = w.diag; ga.rhs = w.rhs; ga.x = gw.x; ga.d32 = gw.d32; ga.pf = gw.pf; ga.V = gw.V; ga.z = gw.z;
    ga.conv = gw.conv; ga.iters = gw.iters; ga.resid = gw.resid;
    ga.nx = nx; ga.ny = ny; ga.nfull = n; ga.m = restart; ga.max_cycles = max_cycles; ga.max_iters = max_iters;
    ga.tol = tol; ga.inner_tol = (float)inner_tol; ga.nR = g.nR; ga.include_boundary = g.include_boundary;
    int rc = launch_gmres(ga, nodes, stream);
    if (rc) return rc;
    if (marks) {
        hipEventRecord(marks[3], stream);
        hipEventRecord(marks[4], stream);
    }
    // nodes that missed the tolerance: the plain band-LU path, into the same x
    LuArgs la;
    la.diag = w.diag; la.rhs = w.rhs; la.y = w.y; la.ab = w.ab; la.ipiv = w.ipiv; la.uext = w.uext;
    la.n = n; la.nx = nx; la.write_L = 0; la.status = status; la.nR = g.nR;
    la.include_boundary = g.include_boundary;
    la.nn = nullptr; la.order = nullptr; la.done = gw.conv;
    la.queue = w.queue; la.qnodes = nodes;
    if (!ublk_geometry(n, kl, ku, &la.ublk_cols, &la.ublk_stride)) la.ublk_cols = 0;
    rc = launch_band_lu(la, nodes, stream, w.lu_variant);
    if (rc) return rc;
    rc = launch_backsolve(w.ab, w.y, w.uext, gw.x, n, kl, ku, ldab, nodes, g.nR, g.include_boundary, stream, nullptr,
                          nullptr, gw.conv, la.ublk_cols, la.ublk_stride);
    if (rc) return rc;
    hipLaunchKernelGGL(scatter_psd, agrid, dim3(kAsmThreads), 0, stream, g, gw.x, psd, (const int*)nullptr);
    if (marks) hipEventRecord(marks[5], stream);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}


int launch_local_mixed(const verb_grid& g, double* psd, const LocalWork& w, const MixedWork& mw, double tol,
                       int max_sweeps, int* status, hipStream_t stream, hipEvent_t* marks) {
    if (g.nV < 3 || g.nK < 3 || max_sweeps < 1) return VERB_ERR_ARG;
    const int nx = g.nV, ny = g.nK, n = nx * ny;
    const int kl = nx + 1, ku = nx + 1, ldab = 2 * kl + ku + 1;
    const int nodes = g.nP * g.nR;
    const int num_substeps = max(1, (int)(g.dt / g.sub_dt));
    if (num_substeps != 1) return VERB_ERR_ARG;
    const double dt_sub = g.dt;
    dim3 agrid((n + kAsmThreads - 1) / kAsmThreads, nodes);
    if (marks) hipEventRecord(marks[0], stream);
    launch_assemble(g, psd, w.ab, w.b0, w.c0, w.rhs, w.diag, dt_sub, 0, 1, status, nullptr, stream);
    hipLaunchKernelGGL(mixed_init, dim3(nodes), dim3(kMxThreads), 0, stream, g, w.diag, w.rhs, mw.d32, mw.x,
                       mw.bmax, mw.done, mw.sweeps);
    if (marks) {
        hipEventRecord(marks[1], stream);
        hipEventRecord(marks[2], stream);
    }
    LuArgsT<float> la;
    la.diag = mw.d32; la.rhs = nullptr; la.y = nullptr; la.ab = mw.ab32; la.ipiv = mw.ipiv; la.uext = nullptr;
    la.n = n; la.nx = nx; la.write_L = 1; la.status = status; la.nR = g.nR;
    la.include_boundary = g.include_boundary;
    la.nn = nullptr; la.order = nullptr;
    int rc = launch_band_lu32(la, nodes, stream);
    if (rc) return rc;
    if (marks) {
        hipEventRecord(marks[3], stream);
        hipEventRecord(marks[4], stream);
    }
    dim3 ugrid((n + 255) / 256, nodes);
    for (int s = 0; s <= max_sweeps; s++) {
        const int solve = s < max_sweeps;
        hipLaunchKernelGGL(mixed_residual, dim3(nodes), dim3(kMxThreads), 0, stream, g, w.diag, w.rhs, mw.x, mw.r32,
                           mw.bmax, tol, mw.rres, mw.done, mw.sweeps, solve);
        if (!solve) break;
        rc = launch_forward32(mw.ab32, mw.ipiv, n, kl, ku, ldab, mw.r32, mw.y32, nodes, g.nR, g.include_boundary,
                              stream, mw.done);
        if (rc) return rc;
        rc = launch_backsolve32(mw.ab32, mw.y32, mw.c32, n, kl, ku, ldab, nodes, g.nR, g.include_boundary, stream,
                                mw.done);
        if (rc) return rc;
        hipLaunchKernelGGL(mixed_update, ugrid, dim3(256), 0, stream, g, mw.x, mw.c32, mw.done);
    }
    hipLaunchKernelGGL(scatter_psd, agrid, dim3(kAsmThreads), 0, stream, g, mw.x, psd, (const int*)nullptr);
    if (marks) hipEventRecord(marks[5], stream);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace verb
