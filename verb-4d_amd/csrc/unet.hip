// The learned warm start for the local (E, alpha) solve: a 4-level U-Net
// (width 32 -> 64 -> 96 -> 128, ~0.51M parameters) that maps five per-node
// channels - log Daa, log DEE, DaE / sqrt(Daa DEE), log PSD_in and the
// normalised dt - to delta = (PSD_out - PSD_in) / s on the node's grid (s
// the node's max |PSD_in|), so that x0 = PSD_in + s delta starts the GMRES
// solve of the implicit system. No counterpart in the reference (src/Diffusion_2D.cpp:144-148
// factors every node with dgbtrf_).
//
// Every layer is a hand-written gfx950 kernel:
//   conv_gemm   3x3 convolution as an implicit GEMM on
//               v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation):
//               rows are pixels, columns output channels, k = (tap,
//               channel). The same kernel is the data gradient (taps
//               mirrored, weights transposed). Input concatenation and the
//               nearest 2x upsampling of the decoder are folded into its
//               operand loads.
//   conv_wgrad  weight gradient dW[t][c][o] = sum_p X(p + off_t)[c] dY[p][o]
//               on the same MFMA, pixels as the k dimension, FP32 atomics.
//   relu_bwd, pool (2x2 average) forward / backward, concat split, the
//   1x1 head with the loss, Adam, weight packing, feature and x0 kernels.
// Activations are NHWC BF16 (channels contiguous), gradients FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "unet.h"

namespace verb {
namespace unet {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even
__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kT = 256;  // threads per workgroup

// ---------------------------------------------------------------------------
// 3x3 convolution (forward, or data gradient with flip = -1)
// ---------------------------------------------------------------------------
struct Src {
    const uint16_t* s0;
    const uint16_t* s1;
    int c0, c1, up0;
};

// 4 BF16 channels [c, c+4) of the concatenated input at pixel (n, hs, ws)
// of an H x W level (zeros outside it); c is a multiple of 4.
__device__ inline bf16x4 load_in(const Src& s, int n, int hs, int ws, int H, int W, int c) {
    bf16x4 z = {0, 0, 0, 0};
    if (hs < 0 || hs >= H || ws < 0 || ws >= W) return z;
    const uint16_t* p;
    if (c < s.c0) {
        if (s.up0) {
            const int H0 = H >> 1, W0 = W >> 1;
            p = s.s0 + (((size_t)n * H0 + (hs >> 1)) * W0 + (ws >> 1)) * s.c0 + c;
        } else {
            p = s.s0 + (((size_t)n * H + hs) * W + ws) * s.c0 + c;
        }
    } else {
        p = s.s1 + (((size_t)n * H + hs) * W + ws) * s.c1 + (c - s.c0);
    }
    return *reinterpret_cast<const bf16x4*>(p);
}

// 64 output pixels per workgroup, 16 per wavefront, all 16*NB output
// channels per wavefront. wp: BF16 weights packed [9][cin/4][cout][4]
// (lane l of an MFMA B operand takes the 4 k-values of its column in one
// 8-byte load).
template <int NB>
__global__ __launch_bounds__(kT) void conv_gemm(Src s, const uint16_t* __restrict__ wp, int N, int H, int W, int flip,
                                                const float* __restrict__ bias, int relu,
                                                uint16_t* __restrict__ out_bf, float* __restrict__ out_f) {
    constexpr int COUT = NB * 16;
    const int cin = s.c0 + s.c1;
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const size_t M = (size_t)N * H * W;
    const size_t p0 = ((size_t)blockIdx.x * 4 + wv) * 16;
    if (p0 >= M) return;
    // the pixel whose A row this lane loads
    const size_t pa = p0 + (l & 15);
    const bool va = pa < M;
    const int na = va ? (int)(pa / ((size_t)H * W)) : 0;
    const int ra = va ? (int)(pa % ((size_t)H * W)) : 0;
    const int ha = ra / W, wa = ra % W;
    const int kg = l >> 4;  // 4-channel group of the 16-wide k chunk
    f32x4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nchunk = cin >> 4;
    for (int t = 0; t < 9; t++) {
        const int dh = flip * (t / 3 - 1), dw = flip * (t % 3 - 1);
        const int hs = ha + dh, ws = wa + dw;
        const bool vin = va && hs >= 0 && hs < H && ws >= 0 && ws < W;
        for (int cc = 0; cc < nchunk; cc++) {
            const int c = cc * 16 + 4 * kg;
            bf16x4 a = {0, 0, 0, 0};
            if (vin) a = load_in(s, na, hs, ws, H, W, c);
            const uint16_t* wb = wp + (((size_t)t * (cin >> 2) + (c >> 2)) * COUT + (l & 15)) * 4;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const bf16x4 bv = *reinterpret_cast<const bf16x4*>(wb + (size_t)b * 16 * 4);
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[b], 0, 0, 0);
            }
        }
    }
    // C layout: lane l, element i -> row 4*(l/16) + i, column l % 16
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int o = b * 16 + (l & 15);
        const float bo = bias ? bias[o] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const size_t p = p0 + 4 * kg + i;
            if (p >= M) continue;
            float v = acc[b][i] + bo;
            if (relu) v = fmaxf(v, 0.f);
            if (out_bf) out_bf[p * COUT + o] = f2bf(v);
            if (out_f) out_f[p * COUT + o] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[t][c][o] += sum over this workgroup's pixels of
// X(p + off_t)[c] * dY[p][o]; MFMA rows = input channels (16 per
// wavefront, 64 per workgroup), columns = output channels, k = pixels.
// Grid (9 taps, ceil(cin/64), pixel groups).
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(kT) void conv_wgrad(Src s, const uint16_t* __restrict__ dy, int N, int H, int W,
                                                 float* __restrict__ dw) {
    constexpr int COUT = NB * 16;
    const int cin = s.c0 + s.c1;
    const int t = blockIdx.x;
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int cb = blockIdx.y * 64 + wv * 16;
    if (cb >= cin) return;  // whole wavefront: channel blocks are 16-aligned
    const int dh = t / 3 - 1, dww = t % 3 - 1;
    const size_t M = (size_t)N * H * W;
    const size_t per = (M + gridDim.z - 1) / gridDim.z;
    const size_t pbeg = (size_t)blockIdx.z * per;
    const size_t pend = pbeg + per < M ? pbeg + per : M;
    const int kg = l >> 4;
    const int ca = cb + (l & 15);  // the A row (input channel) of this lane
    f32x4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (size_t pk = pbeg; pk < pend; pk += 16) {
        bf16x4 a;
        bf16x4 bv[NB];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const size_t p = pk + 4 * kg + i;
            short av = 0;
            if (p < pend) {
                const int n = (int)(p / ((size_t)H * W));
                const int r = (int)(p % ((size_t)H * W));
                const int hs = r / W + dh, ws = r % W + dww;
                if (hs >= 0 && hs < H && ws >= 0 && ws < W) {
                    const uint16_t* src;
                    if (ca < s.c0) {
                        if (s.up0) {
                            const int H0 = H >> 1, W0 = W >> 1;
                            src = s.s0 + (((size_t)n * H0 + (hs >> 1)) * W0 + (ws >> 1)) * s.c0 + ca;
                        } else {
                            src = s.s0 + (((size_t)n * H + hs) * W + ws) * s.c0 + ca;
                        }
                    } else {
                        src = s.s1 + (((size_t)n * H + hs) * W + ws) * s.c1 + (ca - s.c0);
                    }
                    av = (short)*src;
                }
            }
            a[i] = av;
#pragma unroll
            for (int b = 0; b < NB; b++) bv[b][i] = (p < pend) ? (short)dy[p * COUT + b * 16 + (l & 15)] : (short)0;
        }
#pragma unroll
        for (int b = 0; b < NB; b++) acc[b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv[b], acc[b], 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int o = b * 16 + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int c = cb + 4 * kg + i;
            if (c < cin) atomicAdd(&dw[((size_t)t * cin + c) * COUT + o], acc[b][i]);
        }
    }
}

// dpre = dOut where the ReLU output is positive (BF16), and the bias
// gradient sum_p dpre[p][o] (LDS partial sums, one atomic per channel and
// workgroup).
__global__ __launch_bounds__(kT) void relu_bwd(const float* __restrict__ g, const uint16_t* __restrict__ y,
                                               size_t count, int C, uint16_t* __restrict__ dpre,
                                               float* __restrict__ dbias) {
    __shared__ float part[128];
    for (int i = threadIdx.x; i < C; i += kT) part[i] = 0.f;
    __syncthreads();
    for (size_t e = (size_t)blockIdx.x * kT + threadIdx.x; e < count; e += (size_t)gridDim.x * kT) {
        const uint16_t d = (bf2f(y[e]) > 0.f) ? f2bf(g[e]) : (uint16_t)0;
        dpre[e] = d;
        if (d) atomicAdd(&part[e % C], bf2f(d));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < C; i += kT)
        if (part[i] != 0.f) atomicAdd(&dbias[i], part[i]);
}

// 2x2 average pooling of an H x W level (BF16 in and out, FP32 sum).
__global__ __launch_bounds__(kT) void pool_fwd(const uint16_t* __restrict__ x, int N, int H, int W, int C,
                                               uint16_t* __restrict__ y) {
    const int Ho = H >> 1, Wo = W >> 1;
    const size_t total = (size_t)N * Ho * Wo * C;
    for (size_t e = (size_t)blockIdx.x * kT + threadIdx.x; e < total; e += (size_t)gridDim.x * kT) {
        const int c = (int)(e % C);
        const size_t pix = e / C;
        const int wo = (int)(pix % Wo), ho = (int)((pix / Wo) % Ho);
        const size_t n = pix / ((size_t)Wo * Ho);
        const size_t b = ((n * H + 2 * ho) * W + 2 * wo) * C + c;
        const float s = bf2f(x[b]) + bf2f(x[b + C]) + bf2f(x[b + (size_t)W * C]) + bf2f(x[b + (size_t)W * C + C]);
        y[e] = f2bf(0.25f * s);
    }
}

// g_in (H x W) += g_out (H/2 x W/2) / 4
__global__ __launch_bounds__(kT) void pool_bwd_add(const float* __restrict__ gout, int N, int H, int W, int C,
                                                   float* __restrict__ gin) {
    const size_t total = (size_t)N * H * W * C;
    const int Ho = H >> 1, Wo = W >> 1;
    for (size_t e = (size_t)blockIdx.x * kT + threadIdx.x; e < total; e += (size_t)gridDim.x * kT) {
        const int c = (int)(e % C);
        const size_t pix = e / C;
        const int w = (int)(pix % W), h = (int)((pix / W) % H);
        const size_t n = pix / ((size_t)W * H);
        gin[e] += 0.25f * gout[((n * Ho + (h >> 1)) * Wo + (w >> 1)) * C + c];
    }
}

// data gradient of a concatenated input [c0 upsampled | c1]: g0 (half
// resolution) = 2x2 sum of the first c0 channels, g1 = the other c1
// (both assigned). One thread per half-resolution element of g0 and per
// full-resolution element of g1.
__global__ __launch_bounds__(kT) void split_grad(const float* __restrict__ dx, int N, int H, int W, int c0, int c1,
                                                 float* __restrict__ g0, float* __restrict__ g1) {
    const int cin = c0 + c1;
    const int Ho = H >> 1, Wo = W >> 1;
    const size_t n0 = (size_t)N * Ho * Wo * c0;
    const size_t n1 = (size_t)N * H * W * c1;
    for (size_t e = (size_t)blockIdx.x * kT + threadIdx.x; e < n0 + n1; e += (size_t)gridDim.x * kT) {
        if (e < n0) {
            const int c = (int)(e % c0);
            const size_t pix = e / c0;
            const int wo = (int)(pix % Wo), ho = (int)((pix / Wo) % Ho);
            const size_t n = pix / ((size_t)Wo * Ho);
            const size_t b = ((n * H + 2 * ho) * W + 2 * wo) * cin + c;
            g0[e] = dx[b] + dx[b + cin] + dx[b + (size_t)W * cin] + dx[b + (size_t)W * cin + cin];
        } else {
            const size_t f = e - n0;
            const int c = (int)(f % c1);
            const size_t pix = f / c1;
            g1[f] = dx[pix * cin + c0 + c];
        }
    }
}

// 1x1 head: delta[p] = sum_c y[p][c] w[c] + b. With targets: the loss
// sum (pin + delta - pout)^2 / norm, its gradient into gy [p][32] and the
// head's parameter gradients (LDS partial sums + atomics).
__global__ __launch_bounds__(kT) void head(const uint16_t* __restrict__ y, size_t M, const float* __restrict__ wb,
                                           float* __restrict__ delta, const float* __restrict__ pin,
                                           const float* __restrict__ pout, float inv_norm, float* __restrict__ gy,
                                           float* __restrict__ gw, float* __restrict__ loss) {
    __shared__ float pw[kHeadCh + 2];
    for (int i = threadIdx.x; i < kHeadCh + 2; i += kT) pw[i] = 0.f;
    __syncthreads();
    // this thread's sums of the head's parameter gradients and of the loss
    float acc[kHeadCh + 2];
#pragma unroll
    for (int c = 0; c < kHeadCh + 2; c++) acc[c] = 0.f;
    for (size_t p = (size_t)blockIdx.x * kT + threadIdx.x; p < M; p += (size_t)gridDim.x * kT) {
        float d = wb[kHeadCh];
        float yc[kHeadCh];
#pragma unroll
        for (int c = 0; c < kHeadCh; c++) {
            yc[c] = bf2f(y[p * kHeadCh + c]);
            d = fmaf(yc[c], wb[c], d);
        }
        if (delta) delta[p] = d;
        if (!pout) continue;
        const float diff = (pin[p] + d) - pout[p];
        acc[kHeadCh + 1] = fmaf(diff * diff, inv_norm, acc[kHeadCh + 1]);
        const float gd = 2.f * diff * inv_norm;
#pragma unroll
        for (int c = 0; c < kHeadCh; c++) {
            gy[p * kHeadCh + c] = gd * wb[c];
            acc[c] = fmaf(gd, yc[c], acc[c]);
        }
        acc[kHeadCh] += gd;
    }
    if (!pout) return;  // (no barrier follows on this path)
#pragma unroll
    for (int c = 0; c < kHeadCh + 2; c++)
        if (acc[c] != 0.f) atomicAdd(&pw[c], acc[c]);
    __syncthreads();
    for (int i = threadIdx.x; i < kHeadCh + 1; i += kT)
        if (pw[i] != 0.f) atomicAdd(&gw[i], pw[i]);
    if (threadIdx.x == 0 && loss) atomicAdd(loss, pw[kHeadCh + 1]);
}

// fp32 weights [9][cin][cout] -> BF16 packed for the forward B operand
// [9][cin/4][cout][4] and for the data gradient [9][cout/4][cin][4]
__global__ __launch_bounds__(kT) void pack_weights(const float* __restrict__ w, int cin, int cout,
                                                   uint16_t* __restrict__ fwd, uint16_t* __restrict__ bwd) {
    const size_t total = (size_t)9 * cin * cout;
    for (size_t e = (size_t)blockIdx.x * kT + threadIdx.x; e < total; e += (size_t)gridDim.x * kT) {
        const int o = (int)(e % cout);
        const int c = (int)((e / cout) % cin);
        const int t = (int)(e / ((size_t)cout * cin));
        const uint16_t h = f2bf(w[e]);
        fwd[(((size_t)t * (cin >> 2) + (c >> 2)) * cout + o) * 4 + (c & 3)] = h;
        if (bwd) bwd[(((size_t)t * (cout >> 2) + (o >> 2)) * cin + c) * 4 + (o & 3)] = h;
    }
}

__global__ __launch_bounds__(kT) void fill_f32(float* __restrict__ p, size_t n, float v) {
    for (size_t e = (size_t)blockIdx.x * kT + threadIdx.x; e < n; e += (size_t)gridDim.x * kT) p[e] = v;
}

__global__ __launch_bounds__(kT) void adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                           float* __restrict__ v, size_t n, float lr, float b1, float b2, float c1,
                                           float c2, float eps) {
    for (size_t e = (size_t)blockIdx.x * kT + threadIdx.x; e < n; e += (size_t)gridDim.x * kT) {
        const float ge = g[e];
        const float me = b1 * m[e] + (1.f - b1) * ge;
        const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
        m[e] = me;
        v[e] = ve;
        p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
    }
}

// max |psd| of each node (the node's normalisation)
__global__ __launch_bounds__(kT) void node_scale(const double* __restrict__ psd, int n, int node0,
                                                 double* __restrict__ scale) {
    __shared__ double red[kT];
    const double* x = psd + (size_t)(node0 + blockIdx.x) * n;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += kT) m = fmax(m, fabs(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[blockIdx.x] = red[0] > 0.0 ? red[0] : 1.0;
}

// feature channels of node node0 + blockIdx.y on its padded Hp x Wp grid
// (zeros outside nV x nK and in channels 5..15), and pin = psd / scale
__global__ __launch_bounds__(kT) void features(const verb_grid g, const double* __restrict__ psd, int node0,
                                               const double* __restrict__ scale, int Hp, int Wp,
                                               uint16_t* __restrict__ feat, float* __restrict__ pin) {
    const int j = blockIdx.y;
    const int node = node0 + j;
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p >= Hp * Wp) return;
    const int h = p / Wp, w = p % Wp;
    uint16_t* f = feat + ((size_t)j * Hp * Wp + p) * kInCh;
    float ch[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    float pv = 0.f;
    if (h < g.nV && w < g.nK) {
        const size_t i = ((size_t)node * g.nV + h) * g.nK + w;
        const double daa = g.DKK[i], dvv = g.DVV[i], dvk = g.DVK[i], e = g.Vg[i];
        const double dee = dvv / (e * e);
        const double s = scale[j];
        const double pn = psd[i] / s;
        ch[0] = (float)((log10(fmax(daa, 1e-30)) + 5.0) / 3.0);
        ch[1] = (float)((log10(fmax(dee, 1e-30)) + 5.0) / 3.0);
        ch[2] = (float)(dvk / sqrt(fmax(dvv * daa, 1e-300)));
        ch[3] = (float)(log10(fmax(pn, 1e-30)) / 15.0 + 1.0);
        ch[4] = (float)(g.dt / 600.0);
        pv = (float)pn;
    }
#pragma unroll
    for (int c = 0; c < kInCh; c++) f[c] = (c < 5) ? f2bf(ch[c]) : (uint16_t)0;
    pin[(size_t)j * Hp * Wp + p] = pv;
}

// pout = psd / scale on the padded grid (targets)
__global__ __launch_bounds__(kT) void normalize(const verb_grid g, const double* __restrict__ psd, int node0,
                                                const double* __restrict__ scale, int Hp, int Wp,
                                                float* __restrict__ out) {
    const int j = blockIdx.y;
    const int node = node0 + j;
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p >= Hp * Wp) return;
    const int h = p / Wp, w = p % Wp;
    float v = 0.f;
    if (h < g.nV && w < g.nK) v = (float)(psd[((size_t)node * g.nV + h) * g.nK + w] / scale[j]);
    out[(size_t)j * Hp * Wp + p] = v;
}

// x0 = psd + scale * delta for nodes node0.. of the grid ([P][R][V][K] like psd)
__global__ __launch_bounds__(kT) void make_x0(const verb_grid g, const double* __restrict__ psd, int node0,
                                              const float* __restrict__ delta, const double* __restrict__ scale,
                                              int Hp, int Wp, double* __restrict__ x0) {
    const int j = blockIdx.y;
    const int node = node0 + j;
    const int p = blockIdx.x * kT + threadIdx.x;
    const int n = g.nV * g.nK;
    if (p >= n) return;
    const int h = p / g.nK, w = p % g.nK;
    const double d = (double)delta[(size_t)j * Hp * Wp + h * Wp + w];
    const size_t i = (size_t)node * n + p;
    x0[i] = psd[i] + scale[j] * d;
}

int grid1(size_t work) {
    size_t b = (work + kT - 1) / kT;
    if (b > 8192) b = 8192;
    return (int)(b ? b : 1);
}

template <int NB>
void launch_conv(const Src& s, const uint16_t* wp, int N, int H, int W, int flip, const float* bias, int relu,
                 uint16_t* ob, float* of, hipStream_t st) {
    const size_t M = (size_t)N * H * W;
    const int blocks = (int)((M + 63) / 64);
    hipLaunchKernelGGL(conv_gemm<NB>, dim3(blocks), dim3(kT), 0, st, s, wp, N, H, W, flip, bias, relu, ob, of);
}

int conv_dispatch(int cout, const Src& s, const uint16_t* wp, int N, int H, int W, int flip, const float* bias,
                  int relu, uint16_t* ob, float* of, hipStream_t st) {
    switch (cout) {
        case 16: launch_conv<1>(s, wp, N, H, W, flip, bias, relu, ob, of, st); break;
        case 32: launch_conv<2>(s, wp, N, H, W, flip, bias, relu, ob, of, st); break;
        case 64: launch_conv<4>(s, wp, N, H, W, flip, bias, relu, ob, of, st); break;
        case 96: launch_conv<6>(s, wp, N, H, W, flip, bias, relu, ob, of, st); break;
        case 128: launch_conv<8>(s, wp, N, H, W, flip, bias, relu, ob, of, st); break;
        case 160: launch_conv<10>(s, wp, N, H, W, flip, bias, relu, ob, of, st); break;
        case 224: launch_conv<14>(s, wp, N, H, W, flip, bias, relu, ob, of, st); break;
        default: return VERB_ERR_ARG;
    }
    return VERB_OK;
}

template <int NB>
void launch_wgrad(const Src& s, const uint16_t* dy, int N, int H, int W, float* dw, hipStream_t st) {
    const int cin = s.c0 + s.c1;
    const size_t M = (size_t)N * H * W;
    int groups = (int)((M + 4095) / 4096);
    if (groups > 64) groups = 64;
    if (groups < 1) groups = 1;
    hipLaunchKernelGGL(conv_wgrad<NB>, dim3(9, (cin + 63) / 64, groups), dim3(kT), 0, st, s, dy, N, H, W, dw);
}

int wgrad_dispatch(int cout, const Src& s, const uint16_t* dy, int N, int H, int W, float* dw, hipStream_t st) {
    switch (cout) {
        case 32: launch_wgrad<2>(s, dy, N, H, W, dw, st); break;
        case 64: launch_wgrad<4>(s, dy, N, H, W, dw, st); break;
        case 96: launch_wgrad<6>(s, dy, N, H, W, dw, st); break;
        case 128: launch_wgrad<8>(s, dy, N, H, W, dw, st); break;
        default: return VERB_ERR_ARG;
    }
    return VERB_OK;
}

}  // namespace
}  // namespace unet
}  // namespace verb

using namespace verb::unet;

struct verb_unet {
    verb_ctx* ctx = nullptr;
    int max_batch = 0, Hp = 0, Wp = 0, train = 0;
    long step = 0;
    std::vector<void*> owned;
    float* params = nullptr;  // fp32 master
    float* grads = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* loss = nullptr;    // device scalar
    uint16_t* wfwd[kConvs] = {};
    uint16_t* wbwd[kConvs] = {};
    // activations (BF16): input x and a[l] the output of convolution l;
    // pooled p[1..3] (p[k] = pool of the level k-1 skip)
    uint16_t* act[kConvs] = {};
    uint16_t* pooled[4] = {};
    float* delta = nullptr;
    // gradients (FP32) of a[l] and p[k], the concatenated-input gradient,
    // and the BF16 pre-activation gradient
    float* gact[kConvs] = {};
    float* gpool[4] = {};
    float* dx = nullptr;
    uint16_t* dpre = nullptr;
};

namespace {

void* dalloc(verb_unet* u, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess) return nullptr;
    u->owned.push_back(p);
    return p;
}

size_t level_px(const verb_unet* u, int level, int batch) {
    return (size_t)batch * (u->Hp >> level) * (u->Wp >> level);
}

void release(verb_unet* u) {
    for (void* p : u->owned) hipFree(p);
    u->owned.clear();
}

// inputs of convolution l at batch/level geometry
Src conv_src(verb_unet* u, const uint16_t* x, int l) {
    const ConvSpec& c = kSpec[l];
    Src s{nullptr, nullptr, c.c0, c.c1, c.up0};
    switch (l) {
        case 0: s.s0 = x; break;
        case 1: s.s0 = u->act[0]; break;
        case 2: s.s0 = u->pooled[1]; break;
        case 3: s.s0 = u->pooled[2]; break;
        case 4: s.s0 = u->pooled[3]; break;
        case 5: s.s0 = u->act[4]; s.s1 = u->act[3]; break;
        case 6: s.s0 = u->act[5]; s.s1 = u->act[2]; break;
        case 7: s.s0 = u->act[6]; s.s1 = u->act[1]; break;
    }
    return s;
}

int pack(verb_unet* u) {
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    for (int l = 0; l < kConvs; l++) {
        const int cin = kSpec[l].c0 + kSpec[l].c1, cout = kSpec[l].cout;
        hipLaunchKernelGGL(pack_weights, dim3(grid1((size_t)9 * cin * cout)), dim3(kT), 0, st,
                           u->params + param_offset(l), cin, cout, u->wfwd[l], u->wbwd[l]);
    }
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int check_geometry(const verb_unet* u, int batch, int H, int W) {
    if (batch < 1 || batch > u->max_batch || H < 8 || W < 8 || (H & 7) || (W & 7)) return VERB_ERR_ARG;
    if ((size_t)H * W > (size_t)u->Hp * u->Wp) return VERB_ERR_ARG;
    return VERB_OK;
}

int forward(verb_unet* u, const uint16_t* x, int batch, int H, int W, float* delta) {
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    for (int l = 0; l < kConvs; l++) {
        const ConvSpec& c = kSpec[l];
        const int Hl = H >> c.level, Wl = W >> c.level;
        const Src s = conv_src(u, x, l);
        int rc = conv_dispatch(c.cout, s, u->wfwd[l], batch, Hl, Wl, 1,
                               u->params + param_offset(l) + (size_t)9 * (c.c0 + c.c1) * c.cout, 1, u->act[l],
                               nullptr, st);
        if (rc) return rc;
        // the encoder skips a1, a2, a3 (levels 0, 1, 2) are pooled into the
        // inputs p1, p2, p3 of the next level
        if (l >= 1 && l <= 3) {
            const size_t n = (size_t)batch * (Hl >> 1) * (Wl >> 1) * c.cout;
            hipLaunchKernelGGL(pool_fwd, dim3(grid1(n)), dim3(kT), 0, st, u->act[l], batch, Hl, Wl, c.cout,
                               u->pooled[l]);
        }
    }
    const size_t M = (size_t)batch * H * W;
    hipLaunchKernelGGL(head, dim3(grid1(M)), dim3(kT), 0, st, u->act[7], M, u->params + param_offset(kConvs), delta,
                       (const float*)nullptr, (const float*)nullptr, 0.f, (float*)nullptr, (float*)nullptr,
                       (float*)nullptr);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

// forward, loss and every parameter gradient into u->grads; the loss
// (sum over pixels of (pin + delta - pout)^2 / (batch H W)) into u->loss
int backward(verb_unet* u, const uint16_t* x, const float* pin, const float* pout, int batch, int H, int W) {
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    const size_t np = param_count();
    hipLaunchKernelGGL(fill_f32, dim3(grid1(np)), dim3(kT), 0, st, u->grads, np, 0.f);
    hipLaunchKernelGGL(fill_f32, dim3(1), dim3(kT), 0, st, u->loss, (size_t)1, 0.f);
    int rc = forward(u, x, batch, H, W, u->delta);
    if (rc) return rc;
    const size_t M = (size_t)batch * H * W;
    const float inv = 1.f / (float)M;
    hipLaunchKernelGGL(head, dim3(grid1(M)), dim3(kT), 0, st, u->act[7], M, u->params + param_offset(kConvs),
                       (float*)nullptr, pin, pout, inv, u->gact[7], u->grads + param_offset(kConvs), u->loss);
    for (int l = kConvs - 1; l >= 0; l--) {
        const ConvSpec& c = kSpec[l];
        const int cin = c.c0 + c.c1;
        const int Hl = H >> c.level, Wl = W >> c.level;
        const size_t ml = (size_t)batch * Hl * Wl;
        float* gw = u->grads + param_offset(l);
        float* gb = gw + (size_t)9 * cin * c.cout;
        hipLaunchKernelGGL(relu_bwd, dim3(grid1(ml * c.cout)), dim3(kT), 0, st, u->gact[l], u->act[l],
                           ml * c.cout, c.cout, u->dpre, gb);
        const Src s = conv_src(u, x, l);
        if ((rc = wgrad_dispatch(c.cout, s, u->dpre, batch, Hl, Wl, gw, st))) return rc;
        if (l == 0) break;
        // data gradient with respect to the (concatenated) input
        const Src sd{u->dpre, nullptr, c.cout, 0, 0};
        if (c.c1 == 0) {
            // single input: its gradient directly (a[l-1] for l = 1; the
            // pooled p for l = 2..4, then back through the pooling)
            float* gin = (l == 1) ? u->gact[0] : u->gpool[l - 1];
            if ((rc = conv_dispatch(cin, sd, u->wbwd[l], batch, Hl, Wl, -1, nullptr, 0, nullptr, gin, st))) return rc;
            if (l >= 2) {
                // p[l-1] is the pool of a[l-1] (l = 2: a1, l = 3: a2, l = 4: a3)
                const int src = l - 1;
                const int Hs = Hl * 2, Ws = Wl * 2;
                const size_t n = (size_t)batch * Hs * Ws * cin;
                hipLaunchKernelGGL(pool_bwd_add, dim3(grid1(n)), dim3(kT), 0, st, gin, batch, Hs, Ws, cin,
                                   u->gact[src]);
            }
        } else {
            if ((rc = conv_dispatch(cin, sd, u->wbwd[l], batch, Hl, Wl, -1, nullptr, 0, nullptr, u->dx, st)))
                return rc;
            // decoder l: src0 = a[l-1] (upsampled), src1 = the skip
            const int skip = (l == 5) ? 3 : (l == 6) ? 2 : 1;
            const size_t n = (size_t)batch * (Hl >> 1) * (Wl >> 1) * c.c0 + ml * c.c1;
            hipLaunchKernelGGL(split_grad, dim3(grid1(n)), dim3(kT), 0, st, u->dx, batch, Hl, Wl, c.c0, c.c1,
                               u->gact[l - 1], u->gact[skip]);
        }
    }
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // namespace

extern "C" {

size_t verb_unet_param_count(void) { return param_count(); }

int verb_unet_create(verb_ctx* ctx, int max_batch, int height, int width, int train, verb_unet** out) {
    if (!ctx || !out || max_batch < 1 || height < 8 || width < 8 || (height & 7) || (width & 7)) return VERB_ERR_ARG;
    *out = nullptr;
    verb_unet* u = new verb_unet();
    u->ctx = ctx;
    u->max_batch = max_batch;
    u->Hp = height;
    u->Wp = width;
    u->train = train ? 1 : 0;
    const size_t np = param_count();
    bool ok = true;
    auto take = [&](size_t bytes) {
        void* p = dalloc(u, bytes);
        ok = ok && p;
        return p;
    };
    u->params = (float*)take(np * 4);
    u->delta = (float*)take(level_px(u, 0, max_batch) * 4);
    for (int l = 0; l < kConvs; l++) {
        const int cin = kSpec[l].c0 + kSpec[l].c1, cout = kSpec[l].cout;
        u->wfwd[l] = (uint16_t*)take((size_t)9 * cin * cout * 2);
        u->wbwd[l] = (uint16_t*)take((size_t)9 * cin * cout * 2);
        u->act[l] = (uint16_t*)take(level_px(u, kSpec[l].level, max_batch) * cout * 2);
    }
    // pooled inputs of levels 1..3: channels of the skip they pool
    const int pc[4] = {0, 32, 64, 96};
    for (int k = 1; k <= 3; k++) u->pooled[k] = (uint16_t*)take(level_px(u, k, max_batch) * pc[k] * 2);
    if (u->train) {
        u->grads = (float*)take(np * 4);
        u->m = (float*)take(np * 4);
        u->v = (float*)take(np * 4);
        u->loss = (float*)take(4);
        size_t dxmax = 0, dpmax = 0;
        for (int l = 0; l < kConvs; l++) {
            const size_t px = level_px(u, kSpec[l].level, max_batch);
            u->gact[l] = (float*)take(px * kSpec[l].cout * 4);
            if (kSpec[l].c1) dxmax = std::max(dxmax, px * (kSpec[l].c0 + kSpec[l].c1));
            dpmax = std::max(dpmax, px * kSpec[l].cout);
        }
        for (int k = 1; k <= 3; k++) u->gpool[k] = (float*)take(level_px(u, k, max_batch) * pc[k] * 4);
        u->dx = (float*)take(dxmax * 4);
        u->dpre = (uint16_t*)take(dpmax * 2);
        hipStream_t st = (hipStream_t)verb_ctx_stream(ctx);
        if (ok) {
            hipLaunchKernelGGL(fill_f32, dim3(grid1(np)), dim3(kT), 0, st, u->m, np, 0.f);
            hipLaunchKernelGGL(fill_f32, dim3(grid1(np)), dim3(kT), 0, st, u->v, np, 0.f);
        }
    }
    if (!ok) {
        release(u);
        delete u;
        return VERB_ERR_NOMEM;
    }
    *out = u;
    return VERB_OK;
}

int verb_unet_destroy(verb_unet* u) {
    if (!u) return VERB_ERR_ARG;
    hipStreamSynchronize((hipStream_t)verb_ctx_stream(u->ctx));
    release(u);
    delete u;
    return VERB_OK;
}

int verb_unet_set_params(verb_unet* u, const float* params) {
    if (!u || !params) return VERB_ERR_ARG;
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    if (hipMemcpyAsync(u->params, params, param_count() * 4, hipMemcpyHostToDevice, st) != hipSuccess)
        return VERB_ERR_HIP;
    int rc = pack(u);
    if (hipStreamSynchronize(st) != hipSuccess) return VERB_ERR_HIP;
    return rc;
}

int verb_unet_get_params(verb_unet* u, float* params) {
    if (!u || !params) return VERB_ERR_ARG;
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    if (hipMemcpyAsync(params, u->params, param_count() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return VERB_ERR_HIP;
    return VERB_OK;
}

int verb_unet_features(verb_unet* u, const verb_grid* g, const double* psd, int node0, int count, uint16_t* feat,
                       float* pin, double* scale) {
    if (!u || !g || !psd || !feat || !pin || !scale || count < 1 || node0 < 0 || node0 + count > g->nP * g->nR)
        return VERB_ERR_ARG;
    const int Hp = (g->nV + 7) / 8 * 8, Wp = (g->nK + 7) / 8 * 8;
    if (!g->DKK || !g->DVV || !g->DVK || !g->Vg) return VERB_ERR_ARG;
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    hipLaunchKernelGGL(node_scale, dim3(count), dim3(kT), 0, st, psd, g->nV * g->nK, node0, scale);
    hipLaunchKernelGGL(features, dim3((Hp * Wp + kT - 1) / kT, count), dim3(kT), 0, st, *g, psd, node0, scale, Hp,
                       Wp, feat, pin);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int verb_unet_normalize(verb_unet* u, const verb_grid* g, const double* psd, int node0, int count,
                        const double* scale, float* out) {
    if (!u || !g || !psd || !scale || !out || count < 1 || node0 < 0 || node0 + count > g->nP * g->nR)
        return VERB_ERR_ARG;
    const int Hp = (g->nV + 7) / 8 * 8, Wp = (g->nK + 7) / 8 * 8;
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    hipLaunchKernelGGL(normalize, dim3((Hp * Wp + kT - 1) / kT, count), dim3(kT), 0, st, *g, psd, node0, scale, Hp,
                       Wp, out);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int verb_unet_forward(verb_unet* u, const uint16_t* feat, int batch, int height, int width, float* delta) {
    if (!u || !feat || !delta) return VERB_ERR_ARG;
    int rc = check_geometry(u, batch, height, width);
    if (rc) return rc;
    return forward(u, feat, batch, height, width, delta);
}

int verb_unet_gradients(verb_unet* u, const uint16_t* feat, const float* pin, const float* pout, int batch,
                        int height, int width, float* grad, float* loss) {
    if (!u || !u->train || !feat || !pin || !pout) return VERB_ERR_ARG;
    int rc = check_geometry(u, batch, height, width);
    if (rc) return rc;
    if ((rc = backward(u, feat, pin, pout, batch, height, width))) return rc;
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    if (grad && hipMemcpyAsync(grad, u->grads, param_count() * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return VERB_ERR_HIP;
    if (loss) {
        if (hipMemcpyAsync(loss, u->loss, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return VERB_ERR_HIP;
    }
    return VERB_OK;
}

int verb_unet_train_step(verb_unet* u, const uint16_t* feat, const float* pin, const float* pout, int batch,
                         int height, int width, float lr, float* loss) {
    if (!u || !u->train || !feat || !pin || !pout || !(lr > 0)) return VERB_ERR_ARG;
    int rc = check_geometry(u, batch, height, width);
    if (rc) return rc;
    if ((rc = backward(u, feat, pin, pout, batch, height, width))) return rc;
    hipStream_t st = (hipStream_t)verb_ctx_stream(u->ctx);
    u->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
    const size_t np = param_count();
    hipLaunchKernelGGL(adam, dim3(grid1(np)), dim3(kT), 0, st, u->params, u->grads, u->m, u->v, np, lr, b1, b2, c1,
                       c2, 1e-8f);
    if ((rc = pack(u))) return rc;
    if (loss) {
        if (hipMemcpyAsync(loss, u->loss, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return VERB_ERR_HIP;
    }
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

}  // extern "C"

namespace verb {

// delta for nodes [node0, node0+count) of g from psd, then x0 = psd + scale delta
// into x0 ([P][R][V][K]); feat/pin/scale are caller workspace sized for
// count nodes of the padded grid, delta [count][Hp][Wp].
int unet_warm_start(verb_unet* u, const verb_grid& g, const double* psd, int node0, int count, uint16_t* feat,
                    float* pin, double* scale, float* delta, double* x0, hipStream_t st) {
    const int Hp = (g.nV + 7) / 8 * 8, Wp = (g.nK + 7) / 8 * 8;
    int rc = verb_unet_features(u, &g, psd, node0, count, feat, pin, scale);
    if (rc) return rc;
    if ((rc = verb_unet_forward(u, feat, count, Hp, Wp, delta))) return rc;
    const int n = g.nV * g.nK;
    hipLaunchKernelGGL(unet::make_x0, dim3((n + unet::kT - 1) / unet::kT, count), dim3(unet::kT), 0, st, g, psd,
                       node0, delta, scale, Hp, Wp, x0);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int unet_max_batch(const verb_unet* u) { return u ? u->max_batch : 0; }
verb_ctx* unet_ctx(const verb_unet* u) { return u ? u->ctx : nullptr; }

}  // namespace verb
