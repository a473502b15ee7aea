// Matrix-free restarted GMRES for the local (E, alpha) systems of
// Diffusion_2D (src/Diffusion_2D.cpp:43-153), one workgroup per (P,R) node.
//
// The system is the one the band LU solves: A x = B0*psd + C0 with A the
// nine-diagonal operator assembled by local_assemble (energy-fast unknowns,
// in = iK*nV + iV). The solve runs in two precisions:
//   - outer: the residual r = b - A x and the update x += d in FP64, and the
//     convergence test max|r| <= tol * max|b| on that FP64 residual;
//   - inner: one GMRES(m) cycle in FP32 on A d = r, right-preconditioned by
//     the exact inverse of the pitch-angle tridiagonal blocks (for every
//     energy, the +-nV couplings of A), classical Gram-Schmidt with the
//     basis vectors kept unnormalised and their scales held in LDS.
// A node that has not met tol after max_cycles cycles, or within
// max_iters inner iterations in all, is flagged; the caller factors exactly
// those nodes with the band LU.
#include "common.h"
#include "gmres.h"

namespace verb {

namespace {

constexpr int kGmThreads = 256;
constexpr int kGmWaves = kGmThreads / 64;

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ inline double wave_max_d(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// Sum over the workgroup of v[0..k-1]; results in out[0..k-1] (LDS), visible
// to every thread on return.
template <int K>
__device__ inline void block_sum(float (&v)[K], int k, float* red, float* out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < K; i++) {
        if (i < k) {
            const float s = wave_sum(v[i]);
            if (lane == 0) red[i * kGmWaves + wv] = s;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < k) {
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < kGmWaves; q++) s += red[threadIdx.x * kGmWaves + q];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

__device__ inline double block_max(double v, double* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    v = wave_max_d(v);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double m = red[0];
#pragma unroll
    for (int q = 1; q < kGmWaves; q++) m = fmax(m, red[q]);
    __syncthreads();
    return m;
}

// z = M^-1 (scale * v): for every energy ix, the pitch-angle tridiagonal
// (sub ps, inverse pivots pd, super-multipliers pc from the Thomas
// factorisation) over iy = 0..ny-1, unknown iy*nx + ix.
__device__ inline void precond(const float* v, float scale, float* z, const float* ps, const float* pc,
                               const float* pd, int nx, int ny) {
    for (int ix = threadIdx.x; ix < nx; ix += kGmThreads) {
        float yp = 0.0f;
        for (int iy = 0; iy < ny; iy++) {
            const int e = iy * nx + ix;
            const float y = (scale * v[e] - ps[e] * yp) * pd[e];
            z[e] = y;
            yp = y;
        }
        float zn = 0.0f;
        for (int iy = ny - 1; iy >= 0; iy--) {
            const int e = iy * nx + ix;
            const float zv = z[e] - pc[e] * zn;
            z[e] = zv;
            zn = zv;
        }
    }
}

}  // namespace

__global__ __launch_bounds__(kGmThreads) void gmres_nodes(GmArgs a) {
    const int node = blockIdx.x;
    const int tid = threadIdx.x;
    if (a.nR > 0 && !a.include_boundary && (node % a.nR == 0 || node % a.nR == a.nR - 1)) {
        if (tid == 0) {
            a.conv[node] = 1;
            a.iters[node] = 0;
            if (a.resid) a.resid[node] = 0.0;
        }
        return;
    }
    const int nx = a.nx, ny = a.ny, n = nx * ny;
    const size_t nf = (size_t)a.nfull;
    const int m = a.m;
    const double* dg = a.diag + (size_t)node * 9 * nf;
    const double* b = a.rhs + (size_t)node * nf;
    double* x = a.x + (size_t)node * nf;
    float* d32 = a.d32 + (size_t)node * 9 * nf;
    float* ps = a.pf + (size_t)node * 3 * nf;
    float* pc = ps + nf;
    float* pd = pc + nf;
    float* V = a.V + (size_t)node * (kGmMaxM + 1) * nf;
    float* z = a.z + (size_t)node * 2 * nf;
    float* u = z + nf;

    __shared__ float red[(kGmMaxM + 1) * kGmWaves];
    __shared__ float hs[kGmMaxM + 1];
    __shared__ float H[(kGmMaxM + 1) * kGmMaxM];
    __shared__ float scl[kGmMaxM + 1];
    __shared__ float cs[kGmMaxM], sn[kGmMaxM], gv[kGmMaxM + 1], yv[kGmMaxM];
    __shared__ double dred[kGmWaves];
    __shared__ int done, cols;

    int offs[9];
#pragma unroll
    for (int s = 0; s < 9; s++) offs[s] = (s / 3 - 1) * nx + (s % 3 - 1);

    // FP32 operator, and the Thomas factors of the pitch-angle blocks (in FP64)
    for (int i = tid; i < 9 * n; i += kGmThreads) d32[i] = (float)dg[i];
    for (int ix = tid; ix < nx; ix += kGmThreads) {
        double cprev = 0.0;
        for (int iy = 0; iy < ny; iy++) {
            const int e = iy * nx + ix;
            const double sub = (iy > 0) ? dg[1 * n + e] : 0.0;
            const double sup = (iy < ny - 1) ? dg[7 * n + e] : 0.0;
            const double den = dg[4 * n + e] - sub * cprev;
            const double inv = 1.0 / den;
            cprev = sup * inv;
            ps[e] = (float)sub;
            pc[e] = (float)cprev;
            pd[e] = (float)inv;
        }
    }
    double bm = 0.0;
    for (int i = tid; i < n; i += kGmThreads) bm = fmax(bm, fabs(b[i]));
    const double bmax = block_max(bm, dred);  // also orders the factor writes above
    const double thresh = a.tol * bmax;

    int total = 0;
    bool ok = false;
    double last = 0.0;  // relative FP64 residual at the last check
    for (int cyc = 0; cyc <= a.max_cycles; cyc++) {
        // r = b - A x in FP64; the FP32 copy starts the Krylov basis
        double rm = 0.0;
        float rs[1] = {0.0f};
        for (int i = tid; i < n; i += kGmThreads) {
            double acc = b[i];
#pragma unroll
            for (int s = 0; s < 9; s++) {
                const int j = i + offs[s];
                if (j >= 0 && j < n) acc = fma(-dg[(size_t)s * n + i], x[j], acc);
            }
            rm = fmax(rm, fabs(acc));
            const float r32 = (float)acc;
            V[i] = r32;
            rs[0] += r32 * r32;
        }
        const double rmax = block_max(rm, dred);
        last = (bmax > 0.0) ? rmax / bmax : rmax;
        if (rmax <= thresh) {
            ok = true;
            break;
        }
        if (cyc == a.max_cycles) break;
        // the iteration budget: a node that has used it is left to the caller
        const int mc = (a.max_iters > 0) ? min(m, a.max_iters - total) : m;
        if (mc < 1) break;
        block_sum(rs, 1, red, hs);
        const float beta = sqrtf(hs[0]);
        if (!(beta > 0.0f)) break;
        if (tid == 0) {
            scl[0] = 1.0f / beta;
            gv[0] = beta;
            done = 0;
            cols = 0;
        }
        __syncthreads();

        for (int j = 0; j < mc; j++) {
            // z = M^-1 v_j
            precond(V + (size_t)j * nf, scl[j], z, ps, pc, pd, nx, ny);
            __syncthreads();
            // w = A z into slot j+1, and h_i = <w, v_i>
            float hp[kGmMaxM + 1];
#pragma unroll
            for (int i = 0; i <= kGmMaxM; i++) hp[i] = 0.0f;
            for (int e = tid; e < n; e += kGmThreads) {
                float w = 0.0f;
#pragma unroll
                for (int s = 0; s < 9; s++) {
                    const int q = e + offs[s];
                    if (q >= 0 && q < n) w = fmaf(d32[(size_t)s * n + e], z[q], w);
                }
                V[(size_t)(j + 1) * nf + e] = w;
#pragma unroll
                for (int i = 0; i <= kGmMaxM; i++)
                    if (i <= j) hp[i] = fmaf(w, V[(size_t)i * nf + e], hp[i]);
            }
            block_sum(hp, j + 1, red, hs);
            // w -= sum h_i v_i (with the basis scales), and ||w||
            float h[kGmMaxM + 1];
#pragma unroll
            for (int i = 0; i <= kGmMaxM; i++) h[i] = (i <= j) ? hs[i] * scl[i] : 0.0f;
            float nr[1] = {0.0f};
            for (int e = tid; e < n; e += kGmThreads) {
                float w = V[(size_t)(j + 1) * nf + e];
#pragma unroll
                for (int i = 0; i <= kGmMaxM; i++)
                    if (i <= j) w = fmaf(-h[i] * scl[i], V[(size_t)i * nf + e], w);
                V[(size_t)(j + 1) * nf + e] = w;
                nr[0] = fmaf(w, w, nr[0]);
            }
            block_sum(nr, 1, red, hs + kGmMaxM);
            if (tid == 0) {
                const float hn = sqrtf(hs[kGmMaxM]);
                // new column of the Hessenberg matrix, previous rotations applied
                float col[kGmMaxM + 1];
                for (int i = 0; i <= j; i++) col[i] = h[i];
                for (int i = 0; i < j; i++) {
                    const float t = cs[i] * col[i] + sn[i] * col[i + 1];
                    col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
                    col[i] = t;
                }
                const float den = sqrtf(col[j] * col[j] + hn * hn);
                const float c = (den > 0.0f) ? col[j] / den : 1.0f;
                const float sv = (den > 0.0f) ? hn / den : 0.0f;
                cs[j] = c;
                sn[j] = sv;
                col[j] = den;
                for (int i = 0; i <= j; i++) H[i * kGmMaxM + j] = col[i];
                gv[j + 1] = -sv * gv[j];
                gv[j] = c * gv[j];
                scl[j + 1] = (hn > 0.0f) ? 1.0f / hn : 0.0f;
                cols = j + 1;
                if (!(hn > 0.0f) || fabsf(gv[j + 1]) <= a.inner_tol * beta) done = 1;
            }
            __syncthreads();
            if (done) break;
        }
        const int kc = cols;
        total += kc;
        // y = H^-1 g (upper triangular), then d = M^-1 (V y), x += d
        if (tid == 0) {
            for (int i = kc - 1; i >= 0; i--) {
                float s = gv[i];
                for (int q = i + 1; q < kc; q++) s -= H[i * kGmMaxM + q] * yv[q];
                const float d = H[i * kGmMaxM + i];
                yv[i] = (d != 0.0f) ? s / d : 0.0f;
            }
        }
        __syncthreads();
        float yy[kGmMaxM];
#pragma unroll
        for (int i = 0; i < kGmMaxM; i++) yy[i] = (i < kc) ? yv[i] * scl[i] : 0.0f;
        for (int e = tid; e < n; e += kGmThreads) {
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < kGmMaxM; i++)
                if (i < kc) s = fmaf(yy[i], V[(size_t)i * nf + e], s);
            u[e] = s;
        }
        __syncthreads();
        precond(u, 1.0f, z, ps, pc, pd, nx, ny);
        __syncthreads();
        for (int e = tid; e < n; e += kGmThreads) x[e] += (double)z[e];
        __syncthreads();
    }
    if (tid == 0) {
        a.conv[node] = ok ? 1 : 0;
        a.iters[node] = total;
        if (a.resid) a.resid[node] = last;
    }
}

int launch_gmres(const GmArgs& a, int nodes, hipStream_t stream) {
    if (a.m < 1 || a.m > kGmMaxM || a.max_cycles < 0 || a.max_iters < 0) return VERB_ERR_ARG;
    hipLaunchKernelGGL(gmres_nodes, dim3(nodes), dim3(kGmThreads), 0, stream, a);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace verb
