// Batched banded LU with LAPACK's partial pivoting, one workgroup per
// (P,R) node, window resident in MFMA accumulators.
//
// Semantics: the factorisation of dgbtrf/dgbtf2 as the reference calls it
// (src/MatrixSolver.cpp:272/302): pivot = first largest |a| among rows
// j..j+kl of column j; ipiv 1-based; U stored in rows 0..kl+ku of the LAPACK
// band column, the multipliers in rows kl+ku+1.. with the dgbtf2 convention
// (each column's multipliers in the row order of its own step). The
// right-hand side B0*psd + C0 is forward-eliminated panel by panel, in
// dgbtrs order (swap, then subtract the column's multipliers unless b(j) is
// zero), so the plain path never stores L.
//
// Blocking. The matrix is energy-fast (kl = ku = nV+1). Panels are NB = 16
// columns wide. At panel p the live window is RB row blocks x CB column
// blocks of 16x16 tiles in MFMA accumulator layout (RB*16 >= kl+16,
// CB*16 >= kl+ku+16); row block b lives in wavefront b mod RB, and slot s of
// every wavefront holds column block p+s (the registers rotate by one slot per
// panel, so every tile offset is a compile-time constant). Tiles that enter
// the window come straight from the assembled 9-diagonal operator (entries
// that enter the window have not yet been touched by any elimination step).
//
// Look-ahead schedule, four barriers per panel:
//   stage p: one wavefront (the one that held the panel rows of p-1) factors
//      panel p in registers with full-row interchanges, while the others
//      finish the trailing update of panel p-1 on MFMA, stream its U rows
//      to HBM and fetch the factoring wavefront's new row block into an LDS
//      stash; the one that holds the next panel rows also forward-eliminates
//      panel p-1's right-hand-side window;
//   3a: the rows that the panel's interchanges move are staged to LDS, the
//      factoring wavefront forms L11^-1, and the panel-row wavefront writes
//      ipiv and the U fill extents;
//   3b/3c: U12 = L11^-1 P A12 on MFMA; moved rows below the panel take their
//      new contents;
//   4a: column block p+1 is updated first and written to the other panel
//      buffer: it is panel p+1.
// U goes out in the LAPACK band layout (and, for the cached path, with the
// multipliers re-ordered to the dgbtf2 convention) or, when the factors are
// not kept, in contiguous 16-row blocks that the back-substitution streams.
#include <type_traits>
#include <utility>

#include "common.h"
#include "band_lu.h"

namespace verb {


int local_lu_variant = 0;

namespace {

constexpr int NB = 16;

// The assembled operator of one node (nine diagonals of n entries) is read
// through a raw buffer resource: an offset past the end of the buffer reads
// as 0, so entries outside the diagonals need no select on the loaded value
// (which would wait for every load in turn) and no 64-bit address math.
constexpr unsigned kOutside = 0x7ffffff8u;  // past the end of any operator buffer

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t operator_rsrc(const T* dg, int n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(dg), (short)0, 9 * n * (int)sizeof(T), 0x00020000);
}

template <typename T>
__device__ inline T load_entry(__amdgpu_buffer_rsrc_t rs, unsigned off);

template <>
__device__ inline double load_entry<double>(__amdgpu_buffer_rsrc_t rs, unsigned off) {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    const u2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
    return __hiloint2double((int)v.y, (int)v.x);
}

template <>
__device__ inline float load_entry<float>(__amdgpu_buffer_rsrc_t rs, unsigned off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}

__device__ inline double load_entry(__amdgpu_buffer_rsrc_t rs, unsigned off) { return load_entry<double>(rs, off); }

// Byte offset of entry (r, c) in the diagonal storage, or kOutside.
template <typename T>
__device__ inline unsigned a_off(int n, int nx, int r, int c) {
    const int off = c - r;
    const int dy = (off >= nx - 1) ? 1 : ((off <= 1 - nx) ? -1 : 0);
    const int dx = off - dy * nx;
    const bool ok = (r < n) & (c < n) & (dx >= -1) & (dx <= 1);
    return ok ? (unsigned)(((dy + 1) * 3 + dx + 1) * n + r) * (unsigned)sizeof(T) : kOutside;
}

__device__ inline double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// DPP move within 16-lane rows (ctrl: a row-local DPP pattern).
template <int CTRL>
__device__ inline double dpp_row(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ inline float dpp_row(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Wavefront maximum, returned as a wave-uniform value: quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror and row_mirror leave every lane of a
// 16-lane row with the row maximum; the four row maxima are read with
// readlane and combined.
template <typename T, class RL>
__device__ inline T wave_max_t(T v, RL readl) {
    v = fmax(v, dpp_row<0xB1>(v));
    v = fmax(v, dpp_row<0x4E>(v));
    v = fmax(v, dpp_row<0x141>(v));
    v = fmax(v, dpp_row<0x140>(v));
    const T r0 = readl(v, 0), r1 = readl(v, 16), r2 = readl(v, 32), r3 = readl(v, 48);
    return fmax(fmax(r0, r1), fmax(r2, r3));
}

// Arithmetic of the register window in precision T. The accumulator layout of
// the 16x16x4 MFMA (lane j + 16*g, element r, column j) puts row 4*r + g there
// for f64 (rows interleave across the four 16-lane groups) and row 4*g + r
// for f32.
template <typename T>
struct Prec;

template <>
struct Prec<double> {
    typedef double acc4 __attribute__((ext_vector_type(4)));
    static constexpr double maxfinite = 1.7976931348623157e308;
    // high word of |v|: orders magnitudes (ties need the low word)
    __device__ static unsigned key(double v) { return (unsigned)(__double_as_longlong(v) >> 32) & 0x7fffffffu; }
    __device__ static int crow(int g, int r) { return 4 * r + g; }
    __device__ static double readlane(double v, int lane) { return readlane_d(v, lane); }
    __device__ static double wmax(double v) { return wave_max_t(v, [](double x, int q) { return readlane_d(x, q); }); }
    __device__ static acc4 mfma(double a, double b, acc4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <>
struct Prec<float> {
    typedef float acc4 __attribute__((ext_vector_type(4)));
    static constexpr float maxfinite = 3.40282346638528859812e+38f;
    // bits of |v|: orders magnitudes exactly
    __device__ static unsigned key(float v) { return (unsigned)__float_as_int(v) & 0x7fffffffu; }
    __device__ static int crow(int g, int r) { return 4 * g + r; }
    __device__ static float readlane(float v, int lane) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
    }
    __device__ static float wmax(float v) { return wave_max_t(v, [](float x, int q) { return readlane(x, q); }); }
    __device__ static acc4 mfma(float a, float b, acc4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

// 16x16 tile (row block rb, column block cb) of the operator in accumulator
// layout. A tile that meets exactly one diagonal band (always the case for
// nx >= 2*NB) has a wave-uniform row offset dy, so each element costs one
// add, one compare and one select before its load.
template <typename T>
__device__ inline typename Prec<T>::acc4 load_tile(__amdgpu_buffer_rsrc_t rs, int n, int nx, int rb, int cb, int jj,
                                                   int g) {
    using P = Prec<T>;
    const int d = NB * (cb - rb);
    const int lo = d - (NB - 1), hi = d + (NB - 1);
    const bool b0 = (lo <= -nx + 1) & (hi >= -nx - 1);
    const bool b1 = (lo <= 1) & (hi >= -1);
    const bool b2 = (lo <= nx + 1) & (hi >= nx - 1);
    typename P::acc4 t = {T(0), T(0), T(0), T(0)};
    if ((int)b0 + (int)b1 + (int)b2 == 1) {
        const int dy = b0 ? -1 : (b1 ? 0 : 1);
        const bool inside = (NB * rb + NB <= n) & (NB * cb + NB <= n);
        // dx = d - dy*nx + jj - row; offset = ((dy+1)*3 + dx + 1)*n + row
        const int kd = d - dy * nx;
        const int row0 = NB * rb;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int rr = P::crow(g, r);
            const int dx = kd + jj - rr;
            bool ok = (unsigned)(dx + 1) <= 2u;
            if (!inside) ok = ok & (row0 + rr < n) & (NB * cb + jj < n);
            const unsigned off = (unsigned)(((dy + 1) * 3 + dx + 1) * n + row0 + rr) * (unsigned)sizeof(T);
            t[r] = load_entry<T>(rs, ok ? off : kOutside);
        }
    } else if (b0 | b1 | b2) {
#pragma unroll
        for (int r = 0; r < 4; r++) t[r] = load_entry<T>(rs, a_off<T>(n, nx, NB * rb + P::crow(g, r), NB * cb + jj));
    }
    return t;
}

__device__ inline int wave_min_int(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off));
    return __builtin_amdgcn_readfirstlane(v);
}

// Largest unsigned value over the wavefront: four row shifts (lanes that
// shift in from outside the row read 0), then lane 15 of each row into the
// next and lane 31 into rows 2 and 3; lane 63 ends with the maximum.
__device__ inline unsigned wave_max_u32(unsigned v) {
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true));  // row_shr:1
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true));  // row_shr:2
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true));  // row_shr:4
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true));  // row_shr:8
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Calls f(std::integral_constant<int, I>()) for I = B..E-1: the loop index
// is a compile-time constant from the start, so arrays indexed by it are
// split into registers before any loop unrolling takes place.
template <int B, class F, int... I>
__device__ inline void static_for_seq(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, B + I>()), ...);
}

template <int B, int E, class F>
__device__ inline void static_for(F&& f) {
    if constexpr (E > B) static_for_seq<B>(f, std::make_integer_sequence<int, E - B>());
}

// Orders this wavefront's LDS accesses before the call against those after
// it, across lanes (a wavefront executes its LDS instructions in order).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane index read by an opaque instruction sequence. The panel loop takes its
// lane-dependent indices from this at the start of each phase: computed once
// outside the loop, the many LDS addresses derived from them stay live across
// all phases and are spilled to scratch, whose reloads wait on every
// outstanding vector-memory load.
__device__ inline int lane_fresh() {
    int v;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(v));
    return v;
}

__device__ inline int lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
}

template <typename T, int RB, int CB>
struct Lds {
    static constexpr int W = NB * (CB - 1);  // window columns right of the panel
    static constexpr int PR = NB * RB;       // window rows
    // column stride of the panel buffers: 6 elements of padding spread both
    // the 16 columns that one 16x16 tile writes and the 4 columns that an
    // MFMA operand reads over the LDS banks (FP64 column starts stay 16-byte
    // aligned for the L11 reads)
    static constexpr int PS = PR + 6;
    static constexpr int LS = NB + 2;        // row stride of linv (bank spread)
    T pnl[2][NB * PS];                       // panels p and p+1, column-major (full-row interchanges)
    // U12 rows, stride WU = W + 1: the U-row stores read it a column at a
    // time (16 lanes, 16 rows), which an even stride puts on one bank pair
    static constexpr int WU = W + 1;
    T u12[NB * WU];
    T sb[2 * NB * W];                        // pre-interchange contents of the rows that move
    T stash[(CB - 1) * 4 * 64];              // slots 1..CB-1 of the factoring wavefront's tiles
    T rw[PR];                                // right-hand-side window (index mod PR)
    T linv[NB * LS];                         // L11^-1 of the current panel, row-major
    int slot[PR];                            // sb row holding window row i's old contents, -1: none
    int dslot[PR];                           // sb row whose contents window row i takes, -1: unchanged
    int perm[2][PR];                         // per panel parity: window row that the interchanges put at each position
    int jp[NB];
    int ju, info;
    unsigned jz;                             // columns of the current panel with a zero pivot (bit c)
    int lfin;                                // the panel's multipliers are all finite
    int qnode;                               // persistent mode: current dispatch slot
    // The wavefront that shares a SIMD with the factoring one hands the
    // trailing update of NV of its tiles to wavefronts on other SIMDs (an
    // FP64 MFMA holds the SIMD's FP64 pipeline for its 64 cycles, during
    // which the factorisation issues nothing): the tiles go out and come back
    // through vx ([tile][4][lane]), vflag = the panel whose tiles are there
    static constexpr int NV = (sizeof(T) == 8 && RB == 8 && CB == 14) ? 8 : 0;
    T vx[NV > 0 ? NV * 256 : 1];
    int vflag, vdone;
    int u12done;                             // wavefronts done with their deferred U12 tiles in this stage
};
static_assert(sizeof(Lds<double, 8, 14>) <= 160 * 1024, "LDS of the 8x14 window");

}  // namespace

__global__ void zero_counter(int* c) { *c = 0; }
// the counters of the fused launch: LU slots, back-substitution slots, and
// the per-node "factors written" flags
__global__ void zero_lu_counters(int* q, int* bq, int* flags, int nflags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *q = 0;
        *bq = 0;
    }
    if (i < nflags) flags[i] = 0;
}

constexpr int kBsThreads = 256;
constexpr int kXRing = 512;  // x entries kept in LDS: >= kl+ku+32, a power of two
// row stride of the staged diagonal block and of the partial sums: 17 puts
// the 16 rows that the triangle's lanes read on different LDS banks
constexpr int kBsPad = NB + 1;
constexpr size_t kBsLdsDoubles = kXRing + 2 * NB * kBsPad;
// waits of a fused back-substitution for its node's factors: s_sleep(32)
// rounds (about 1 us each); past this the node is reported and skipped
constexpr int kBsMaxSpins = 1 << 18;
// U blocks in flight per fused back-substitution: one node per workgroup, so
// its latency is hidden by prefetch depth (the LU's register budget allows it)
constexpr int kBsFusedSets = 3;
// Where a back-substitution stores x besides its LDS ring: x_all in
// energy-fast order, or, with psd given, the node's PSD slice (row r = iy*nx
// + ix at psd[ix*nK + k0 + iy]: the scatter back to [V][K])
template <typename T>
struct XOut {
    T* psd = nullptr;
    int nx = 1, nK = 0, k0 = 0;
};
template <typename T, bool BLK, int NI, int D>
__device__ __forceinline__ void backsolve_node(const T* __restrict__ ab_all, const T* __restrict__ y_all,
                                               T* __restrict__ x_all, int n, int stride, int kl, int ku, int ldab,
                                               int node, int ublk_cols, size_t ublk_stride, T* sh, int tid,
                                               XOut<T> xout = XOut<T>());

// (FP32 windows up to 8x14 fit 128 registers: two workgroups per CU)
template <typename T, int RB, int CB>
__global__ __launch_bounds__(RB * 64) __attribute__((amdgpu_waves_per_eu((sizeof(T) == 4 && RB <= 8 && CB <= 14) ? 4 : 1))) void band_lu_window(LuArgsT<T> a) {
    using L = Lds<T, RB, CB>;
    using P = Prec<T>;
    using A4 = typename P::acc4;
    constexpr int W = L::W;
    constexpr int WU = L::WU;
    constexpr int PR = L::PR;
    constexpr int PS = L::PS;
    static_assert(RB >= 3, "the trailing update's U rows are written by row blocks 2..RB-1");
    extern __shared__ double lds_raw[];  // (8-byte aligned)
    L& S = *reinterpret_cast<L*>(lds_raw);

    const int tid = threadIdx.x;
    for (int it = 0;; it++) {
    int node;
    if (a.queue) {
        // persistent mode: the next dispatch slot from the node queue
        if (tid == 0) S.qnode = atomicAdd(a.queue, 1);
        __syncthreads();
        const int k = S.qnode;
        __syncthreads();
        if (k >= a.qnodes) break;
        node = a.order ? a.order[k] : k;
    } else {
        if (it > 0) break;
        node = a.order ? a.order[blockIdx.x] : blockIdx.x;
    }
    if (a.nR > 0 && !a.include_boundary && (node % a.nR == 0 || node % a.nR == a.nR - 1)) continue;
    if (a.done && a.done[node]) continue;
    const int stride = a.n;
    const int n = a.nn ? a.nn[node] : a.n, nx = a.nx;
    const int kl = nx + 1, ku = nx + 1, kv = kl + ku, ldab = 2 * kl + ku + 1;
    const T* dg = a.diag + (size_t)node * 9 * stride;
    const T* rhs = a.rhs ? a.rhs + (size_t)node * stride : nullptr;
    T* yv = a.y ? a.y + (size_t)node * stride : nullptr;
    T* ab = a.ab + (size_t)node * ldab * stride;
    int64_t* ipiv = a.ipiv + (size_t)node * stride;
    int* uext = a.uext ? a.uext + (size_t)node * stride : nullptr;

    // wavefront index as a uniform value: per-wavefront branches below are
    // scalar, so the register allocator sees them as exclusive
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l = lane_fresh(), g = l >> 4, jj = l & 15;

    const __amdgpu_buffer_rsrc_t rs = operator_rsrc<T>(dg, n);
    A4 acc[CB];
    // initial window: row block w, column blocks 0..CB-1 in slots 0..CB-1
#pragma unroll
    for (int s = 0; s < CB; s++) acc[s] = load_tile<T>(rs, n, nx, w, s, jj, g);
    for (int i = tid; i < PR; i += RB * 64) S.rw[i] = (rhs && i < n) ? rhs[i] : T(0);
    if (tid == 0) {
        S.ju = 0;
        S.info = 0;
        S.vflag = -1;
        S.vdone = 0;
        S.u12done = 0;
    }
    __syncthreads();

    const int npan = (n + NB - 1) / NB;
    // Look-ahead schedule. Stage p factors panel p on one wavefront while the
    // others finish the trailing update of panel p-1; then come panel p's row
    // interchanges, U12, and the update of column block p+1, which is written
    // to the other panel buffer. The factoring wavefront is the one that held
    // the panel rows of p-1: it has no update work left for p-1.
    {
        // 1. column block 0 -> panel buffer 0; the wavefront that factors
        //    panel 0 keeps its other tiles in the stash meanwhile
        const int l = lane_fresh(), g = l >> 4, jj = l & 15;
#pragma unroll
        for (int r = 0; r < 4; r++) S.pnl[0][jj * PS + NB * w + P::crow(g, r)] = acc[0][r];
        if (w == RB - 1)
#pragma unroll
            for (int s = 1; s < CB; s++)
#pragma unroll
                for (int r = 0; r < 4; r++) S.stash[((s - 1) * 4 + r) * 64 + l] = acc[s][r];
    }
    __syncthreads();
    int ju_prev = 0;
    // sb rows whose contents this wavefront's rows r take in slots 2..CB-1,
    // applied in the next stage's trailing update (-1: none)
    int dfx[4] = {-1, -1, -1, -1};
    // U12 of the previous panel beyond its first column block, formed at the
    // start of the next stage by the wavefronts of row blocks 1..3 and 5..7
    // (finite multipliers, L not kept): the sb rows its moved rows took
    int dslc[4] = {-1, -1, -1, -1};
    bool u12_later = false;
    for (int p = 0;; p++) {
        const int j0 = NB * p;
        const int nc = min(NB, n - j0);
        T* const pn = S.pnl[p & 1];
        const bool fac = (w == (p + RB - 1) % RB);
        constexpr int NV = L::NV;
        bool vpartner = false;  // this wavefront's tiles 2..NV+1 are updated by others in this stage

        // 2. factorisation of panel p in registers. Lane l keeps
        //    window rows l + 64q (q < NQ) as 16-wide vectors; an interchange
        //    only swaps the rows' logical positions, the pivot row is
        //    broadcast with readlane, the rank-1 update runs over the full row
        //    with the pivot-row entries left of the step zeroed, and at the
        //    end every row is stored at its final position (the result of
        //    full-row interchanges).
        if (fac && p < npan) {
            // The factorisation is the critical path of the stage and runs
            // on one wavefront; it shares its SIMD with a wavefront of the
            // trailing update, which has time to spare: issue priority goes
            // to the factorisation.
            __builtin_amdgcn_s_setprio(3);
            constexpr int NQ = (PR + 63) / 64;
            static_assert(NQ >= 1 && NQ <= 3, "window rows per lane");
            const int l = lane_fresh();
            // this wavefront's tiles are dead here: slot 0 is the panel, the
            // others come from the stash after the stage barrier
#pragma unroll
            for (int s = 0; s < CB; s++) acc[s] = __builtin_nondeterministic_value(acc[s]);
            int info = S.info;
            unsigned zmask = 0;  // (uniform) columns with a zero pivot
            constexpr int kNone = 1 << 20;
            // (all indices into these arrays are compile-time constants, so
            // they live in registers)
            T x[NQ][NB];
            int pos[NQ];
            // (loads from a clamped row, so that they need no branch; a
            // non-finite entry turns the sum of the entries times zero into
            // NaN)
            T chk[2] = {T(0), T(0)};
            static_for<0, NQ>([&](auto qc) {
                constexpr int q = qc;
                const int row = l + 64 * q;
                const bool h = 64 * (q + 1) <= PR || row < PR;
                const int rc = h ? row : PR - 1;
                static_for<0, NB>([&](auto kc) {
                    constexpr int k = kc;
                    const T v = pn[k * PS + rc];
                    x[q][k] = h ? v : T(0);
                    chk[k & 1] = fma(x[q][k], T(0), chk[k & 1]);
                });
                pos[q] = h ? row : kNone;
            });
            const bool bad = !isfinite(chk[0] + chk[1]);
            // Columns are taken by an unrolled sequence with the column index
            // known at compile time (only entries right of the step are
            // touched) while the panel holds only finite values, the largest
            // candidate is unique in its key and the pivot is nonzero with a
            // finite reciprocal. Any other column takes the general step
            // below, and the sequence resumes after it. Per column c, lane c
            // records the interchange; columns with a zero pivot are marked.
            const bool fast = __ballot(bad) == 0;
            bool nonfin = !fast;  // (multipliers may be non-finite)
            int myjp = 0;
            int c = 0;
            while (c < nc) {
                static_for<0, NB>([&](auto Cc) {
                    constexpr int C = Cc;
                    if (!(fast && C == c && C < nc)) return;
                    const int j = j0 + C;
                    T xc[NQ];
                    bool cd[NQ];
                    unsigned kk[NQ];
                    unsigned mloc = 0;
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        xc[q] = x[q][C];
                        // candidates: logical rows C..C+km. The rows past
                        // C+km have not moved and hold exact zeros in column
                        // C (below the band, or past the last row), so they
                        // neither reach a nonzero maximum key nor change a
                        // value when they take a zero multiplier.
                        cd[q] = pos[q] >= C;
                        kk[q] = cd[q] ? P::key(xc[q]) : 0u;
                        mloc = max(mloc, kk[q]);
                    });
                    const unsigned m = wave_max_u32(mloc);
                    unsigned long long b[NQ];
                    int cnt = 0;
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        // (non-candidates have key 0, and a zero maximum
                        // matches every lane)
                        b[q] = __ballot(kk[q] == m);
                        cnt += __popcll(b[q]);
                    });
                    if (cnt != 1) return;
                    int qp = 0;
                    unsigned long long bp = b[0];
                    T xcp = xc[0];
                    int posp = pos[0];
                    static_for<1, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        const bool here = b[q] != 0;
                        qp = here ? q : qp;
                        bp = here ? b[q] : bp;
                        xcp = here ? xc[q] : xcp;
                        posp = here ? pos[q] : posp;
                    });
                    const int pl = __ffsll(bp) - 1;
                    const T piv = P::readlane(xcp, pl);
                    const T rinv = T(1) / piv;
                    if (!(piv != T(0) && fabs(rinv) <= P::maxfinite)) return;
                    const int jpos = __builtin_amdgcn_readlane(posp, pl);
                    const int jp = jpos - C;
                    // (jp is uniform: written into lane C)
                    asm("v_writelane_b32 %0, %1, %2" : "+v"(myjp) : "s"(jp), "n"(C));
                    // pivot row right of the step (the branches end
                    // differently so that the compiler keeps them apart
                    // instead of merging them through a copy of a row set)
                    T pr[NB];
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        if (qp == q) {
                            static_for<C + 1, NB>([&](auto kc) {
                                constexpr int k = kc;
                                pr[k] = P::readlane(x[q][k], pl);
                            });
                            asm volatile("; pivot row in set %0" ::"n"(q));
                        }
                    });
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        // interchange: pivot row -> position C, the row at C
                        // -> jpos; the candidates other than the pivot row
                        // take multipliers
                        pos[q] = (pos[q] == jpos) ? C : ((pos[q] == C) ? jpos : pos[q]);
                        const bool e = cd[q] & (pos[q] != C);
                        const T Lq = e ? xc[q] * rinv : T(0);
                        x[q][C] = e ? Lq : xc[q];
                        // finite values: fma(0, t, v) == v for the rows the
                        // step does not touch
                        static_for<C + 1, NB>([&](auto kc) {
                            constexpr int k = kc;
                            x[q][k] = fma(Lq, -pr[k], x[q][k]);
                        });
                    });
                    c = C + 1;
                });
                if (c >= nc) break;

                // general step for column c, on the rows staged in the panel
                // buffer at their physical index
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    const int row = l + 64 * q;
                    if (row < PR)
                        static_for<0, NB>([&](auto kc) {
                            constexpr int k = kc;
                            pn[k * PS + row] = x[q][k];
                        });
                });
                wave_sync();
                const int j = j0 + c;
                const int km = min(kl, n - 1 - j);
                T xc[NQ], v[NQ];
                bool cand[NQ];
                T mloc = T(-1);
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    const int row = l + 64 * q;
                    xc[q] = (row < PR) ? pn[c * PS + row] : T(0);
                    // candidates: logical rows c..c+km
                    cand[q] = (pos[q] >= c) & (pos[q] <= c + km);
                    v[q] = cand[q] ? fabs(xc[q]) : T(-1);
                    mloc = fmax(mloc, v[q]);
                });
                const T m = P::wmax(mloc);
                unsigned long long tb[NQ];
                int ntie = 0;
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    tb[q] = __ballot(cand[q] & (v[q] == m));
                    ntie += __popcll(tb[q]);
                });
                int jpos = c;
                if (ntie == 1) {
                    // a unique largest entry
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        if (tb[q]) jpos = __builtin_amdgcn_readlane(pos[q], __ffsll(tb[q]) - 1);
                    });
                } else if (ntie > 1) {
                    // ties: LAPACK takes the first logical row
                    int mn = kNone;
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        mn = min(mn, (cand[q] & (v[q] == m)) ? pos[q] : kNone);
                    });
                    jpos = wave_min_int(mn);
                }
                // the lane and register set holding logical row jpos
                int qp = 0, pl = 0;
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    const unsigned long long pb = __ballot(pos[q] == jpos);
                    if (pb) {
                        qp = q;
                        pl = __ffsll(pb) - 1;
                    }
                });
                const int jp = jpos - c;
                myjp = (l == c) ? jp : myjp;
                T piv = T(0);
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    if (q == qp) piv = P::readlane(xc[q], pl);
                });
                if (piv != T(0)) {
                    const int prow = pl + 64 * qp;
                    const T rinv = T(1) / piv;
                    const bool finite = fabs(rinv) <= P::maxfinite;
                    nonfin |= !finite;
                    T Lv[NQ];
                    bool ev[NQ];
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        // interchange: pivot row -> position c, the row at c -> jpos
                        pos[q] = (pos[q] == jpos) ? c : ((pos[q] == c) ? jpos : pos[q]);
                        const bool e = (pos[q] > c) & (pos[q] <= c + km);
                        // multipliers; 0 for rows the step does not touch
                        const T Lq = e ? xc[q] * rinv : T(0);
                        const int row = l + 64 * q;
                        if (row < PR) pn[c * PS + row] = e ? Lq : xc[q];
                        Lv[q] = Lq;
                        ev[q] = e;
                    });
                    // rank-1 update right of the step; the pivot row's own
                    // entries stay as they are (its multiplier is 0)
                    for (int k = c + 1; k < NB; k++) {
                        const T t = -pn[k * PS + prow];
                        static_for<0, NQ>([&](auto qc) {
                            constexpr int q = qc;
                            const int row = l + 64 * q;
                            if (row < PR) {
                                const T xv = pn[k * PS + row];
                                // finite multipliers: fma(0, t, x) == x for
                                // untouched rows and fma(L, 0, x) == x for a
                                // zero pivot-row entry
                                pn[k * PS + row] = (finite || (t != T(0) && ev[q])) ? fma(Lv[q], t, xv) : xv;
                            }
                        });
                    }
                } else {
                    if (info == 0) info = j + 1;
                    zmask |= 1u << c;
                }
                wave_sync();
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    const int row = l + 64 * q;
                    if (row < PR)
                        static_for<0, NB>([&](auto kc) {
                            constexpr int k = kc;
                            x[q][k] = pn[k * PS + row];
                        });
                });
                wave_sync();
                c++;
            }
            // rows to their final positions, and the sb maps: panel rows keep
            // their index; a row below the panel that moved was pulled up as a
            // pivot (at most one per column) and takes sb row 16 + its final
            // position. dslot gives each window row the sb row it takes.
            static_for<0, NQ>([&](auto qc) {
                constexpr int q = qc;
                const int row = l + 64 * q;
                if (64 * (q + 1) <= PR || row < PR) {
                    static_for<0, NB>([&](auto kc) {
                        constexpr int k = kc;
                        pn[k * PS + pos[q]] = x[q][k];
                    });
                    S.perm[p & 1][pos[q]] = row;
                    const int sl = (row < NB) ? row : ((pos[q] != row) ? NB + pos[q] : -1);
                    S.slot[row] = sl;
                    S.dslot[pos[q]] = (pos[q] < NB || pos[q] != row) ? sl : -1;
                }
            });
            // (ipiv and the fill extents are written from these in step 3a)
            if (l < NB) S.jp[l] = (l < nc) ? myjp : 0;
            if (l == 0) {
                S.jz = zmask;
                S.info = info;
                S.lfin = !nonfin;
            }
            __builtin_amdgcn_s_setprio(0);
        }
        else if (!fac && p > 0) {
            // 4b. trailing update of panel q = p-1 on the remaining column
            //     blocks; the finished U rows stream out
            const int q = p - 1;
            const int j0 = NB * q, lb = (w - q % RB + RB) % RB, nc = min(NB, n - j0);
            const T* const pn = S.pnl[q & 1];
            const int l = lane_fresh(), g = l >> 4, jj = l & 15;
            const int rowbase = NB * lb;
            // column block q+CB enters the window in slot 0, whose panel tile
            // is done with (requested first: the wavefront that holds the next
            // panel rows stores this tile right after the barrier)
            acc[0] = load_tile<T>(rs, n, nx, q + lb, q + CB, jj, g);
            // The factoring wavefront's row block (p+RB-1, which entered the
            // window at panel p) in slots 1..CB-1: these wavefronts request
            // its tiles and leave them in the stash.
            constexpr int NT = (CB - 1 + RB - 2) / (RB - 1);
            A4 ft[NT];
            static_for<0, NT>([&](auto tc) {
                constexpr int t = tc;
                const int sft = 1 + (lb - 1) + t * (RB - 1);
                if (p < npan && sft < CB) ft[t] = load_tile<T>(rs, n, nx, p + RB - 1, p + sft, jj, g);
            });
            T an[4];
#pragma unroll
            for (int kk = 0; kk < 4; kk++) an[kk] = -pn[(4 * kk + g) * PS + rowbase + jj];
            // U12 tiles 1..CB-2 of panel q (tile 0 was formed in step 3b for
            // step 4a), dealt out to the wavefronts of row blocks 1..3 and 5..7 (not on
            // the factoring wavefront's SIMD), as step 3b forms them
            if (u12_later) {
                const int hh = (lb < 4) ? lb - 1 : lb - 2;  // 0..5 (row block 4: 3, unused)
                if (lb != 4) {
                    T la[4];
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) la[kk] = S.linv[jj * L::LS + 4 * kk + g];
                    static_for<0, (CB - 2 + 5) / 6>([&](auto tc) {
                        constexpr int t = tc;
                        const int cb = 1 + hh + 6 * t;
                        if (cb > CB - 2) return;
                        A4 d = A4{T(0), T(0), T(0), T(0)};
#pragma unroll
                        for (int kk = 0; kk < 4; kk++) {
                            const T bv = (dslc[kk] >= 0) ? S.sb[dslc[kk] * W + NB * cb + jj] : T(0);
                            d = P::mfma(la[kk], bv, d);
                        }
#pragma unroll
                        for (int r = 0; r < 4; r++) S.u12[P::crow(g, r) * WU + NB * cb + jj] = d[r];
                    });
                    if (l == 0) __hip_atomic_fetch_add(&S.u12done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            // the contents that rows moved by panel q's interchanges take in
            // slots 2..CB-1 (slot 1 took them in step 3c)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (dfx[r] >= 0) {
#pragma unroll
                    for (int s = 2; s < CB; s++) acc[s][r] = S.sb[dfx[r] * W + NB * (s - 1) + jj];
                }
            }
            if (u12_later) {
                // (bounded wait for the six)
                int spins = 0;
                while (__hip_atomic_load(&S.u12done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != RB - 2 &&
                       spins < (1 << 16)) {
                    __builtin_amdgcn_s_sleep(1);
                    spins++;
                }
                if (spins >= (1 << 16) && l == 0) raise_error(a.status, VERB_ERR_HIP);
            }
            // slot 1 was updated in step 4a. The wavefront on the factoring
            // one's SIMD (local row block 4 of 8: wavefront w runs on SIMD
            // w mod 4) hands tiles 2..NV+1 to the wavefronts of row blocks
            // 2, 3, 5, 6 and 7, which take them after their own tiles.
            if constexpr (NV > 0) {
                vpartner = lb == 4;
                if (vpartner) {
                    static_for<0, NV>([&](auto tc) {
                        constexpr int t = tc;
                        if (NB * (q + 2 + t) > ju_prev) return;
#pragma unroll
                        for (int r = 0; r < 4; r++) S.vx[(t * 4 + r) * 64 + l] = acc[2 + t][r];
                    });
                    if (l == 0) {
                        S.vdone = 0;
                        __hip_atomic_store(&S.vflag, p, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
            }
#pragma unroll
            for (int s = 2; s < CB; s++) {
                if (NB * (q + s) > ju_prev) continue;
                if (NV > 0 && vpartner && s < 2 + NV) continue;
                A4 c4 = acc[s];
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const T bv = S.u12[(4 * kk + g) * WU + NB * (s - 1) + jj];
                    c4 = P::mfma(an[kk], bv, c4);
                }
                acc[s] = c4;
            }
            // slot s holds column block p + s again
            {
                const A4 t0 = acc[0];
#pragma unroll
                for (int s = 0; s + 1 < CB; s++) acc[s] = acc[s + 1];
                acc[CB - 1] = t0;
            }
            static_for<0, NT>([&](auto tc) {
                constexpr int t = tc;
                const int sft = 1 + (lb - 1) + t * (RB - 1);
                if (p < npan && sft < CB)
#pragma unroll
                    for (int r = 0; r < 4; r++) S.stash[((sft - 1) * 4 + r) * 64 + l] = ft[t][r];
            });
            if constexpr (NV > 0) {
                // the row-block-4 tiles this wavefront takes (t = h, h + 5, ... for
                // row blocks 2, 3, 5, 6, 7), with the same MFMA sequence, so
                // the same values, as their owner's
                const int h = (lb == 2) ? 0 : (lb == 3) ? 1 : (lb == 5) ? 2 : (lb == 6) ? 3 : (lb == 7) ? 4 : -1;
                if (h >= 0) {
                    // (the owner publishes its tiles at the start of its
                    // trailing update; bounded wait)
                    int spins = 0;
                    while (__hip_atomic_load(&S.vflag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != p &&
                           spins < (1 << 16)) {
                        __builtin_amdgcn_s_sleep(1);
                        spins++;
                    }
                    if (spins >= (1 << 16) && l == 0) raise_error(a.status, VERB_ERR_HIP);
                    T ap[4];
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) ap[kk] = -pn[(4 * kk + g) * PS + NB * 4 + jj];
#pragma unroll 1
                    for (int t = h; t < NV; t += 5) {
                        if (NB * (q + 2 + t) > ju_prev) continue;
                        A4 c4;
#pragma unroll
                        for (int r = 0; r < 4; r++) c4[r] = S.vx[(t * 4 + r) * 64 + l];
#pragma unroll
                        for (int kk = 0; kk < 4; kk++) {
                            const T bv = S.u12[(4 * kk + g) * WU + NB * (1 + t) + jj];
                            c4 = P::mfma(ap[kk], bv, c4);
                        }
#pragma unroll
                        for (int r = 0; r < 4; r++) S.vx[(t * 4 + r) * 64 + l] = c4[r];
                    }
                    if (l == 0) __hip_atomic_fetch_add(&S.vdone, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            // Right-hand-side window of panel q, on the wavefront that holds
            // the next panel rows (it has the least trailing-update work):
            // the panel's interchanges, then per column c the multipliers
            // times b(c), skipped when b(c) is zero (dgbtrs). Window
            // position i holds the row that the interchanges put there; its
            // multipliers in column c sit at position i of the panel buffer
            // (rows past the band hold zeros). The panel rows' slots then
            // take the rows that enter the window of panel q+1.
            if (lb == 1 && rhs) {
                constexpr int NQ = (PR + 63) / 64;
                const int* const perm = S.perm[q & 1];
                T rq[NQ];
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    const int i = l + 64 * q;
                    const bool h = 64 * (q + 1) <= PR || i < PR;
                    const int src = perm[h ? i : 0];
                    const T v = S.rw[(unsigned)(j0 + src) % PR];
                    rq[q] = h ? v : T(0);
                });
                // (the multipliers of all columns are requested first)
                T lm[NB][NQ];
                static_for<0, NB>([&](auto cc) {
                    constexpr int c = cc;
                    static_for<0, NQ>([&](auto qc) {
                        constexpr int q = qc;
                        const int i = l + 64 * q;
                        const bool h = 64 * (q + 1) <= PR || i < PR;
                        lm[c][q] = pn[c * PS + (h ? i : 0)];
                    });
                });
                static_for<0, NB>([&](auto cc) {
                    constexpr int c = cc;
                    const T bc = P::readlane(rq[0], c);  // position c < 16: lane c
                    if (c < nc && bc != T(0)) {
                        static_for<0, NQ>([&](auto qc) {
                            constexpr int q = qc;
                            const int i = l + 64 * q;
                            const bool h = 64 * (q + 1) <= PR || i < PR;
                            rq[q] = (h && i > c) ? fma(lm[c][q], -bc, rq[q]) : rq[q];
                        });
                    }
                });
                static_for<0, NQ>([&](auto qc) {
                    constexpr int q = qc;
                    const int i = l + 64 * q;
                    if (64 * (q + 1) <= PR || i < PR) S.rw[(unsigned)(j0 + i) % PR] = rq[q];
                });
                // the panel rows' entries are final
                if (yv && l < nc) yv[j0 + l] = rq[0];
                const int ie = j0 + PR + l;
                if (l < NB) S.rw[(unsigned)ie % PR] = (ie < n) ? rhs[ie] : T(0);
            }
            // U rows j0..j0+nc-1: entry e is row e%16 at column j0 + e/16, up
            // to kl+ku past the diagonal (the entries past the row's fill
            // extent are exact zeros in the window). 16 consecutive lanes
            // write 16 consecutive band rows of one column. The wavefront
            // that holds the next panel rows writes none: it waits for its
            // entering tile right after the barrier, and that wait would
            // cover these stores.
            if (lb >= 2 && a.ublk_cols) {
                // blocked U: the whole 16 x ublk_cols block, zeros outside
                // the band, in one contiguous stretch
                const int c = l & 15;
                const int r = j0 + c;
                const int ue = (c < nc) ? min(r + kv, n - 1) : -1;
                const int wc = a.ublk_cols;
                T* dst = a.ab + (size_t)node * a.ublk_stride + (size_t)q * NB * wc;
                for (int e = 64 * (lb - 2) + l; e < NB * wc; e += 64 * (RB - 2)) {
                    const int col_off = e >> 4;
                    const int col = j0 + col_off;
                    T v = T(0);
                    if (col >= r && col <= ue) v = (col_off < NB) ? pn[col_off * PS + c] : S.u12[c * WU + (col_off - NB)];
                    dst[e] = v;
                }
            } else if (lb >= 2) {
                const int c = l & 15;
                const int r = j0 + c;
                const int ue = (c < nc) ? min(r + kv, n - 1) : -1;
                for (int e = 64 * (lb - 2) + l; e < NB * (NB + W); e += 64 * (RB - 2)) {
                    const int col_off = e >> 4;
                    const int col = j0 + col_off;
                    if (col >= r && col <= ue) {
                        const T v = (col_off < NB) ? pn[col_off * PS + c] : S.u12[c * WU + (col_off - NB)];
                        ab[(size_t)col * ldab + kv + r - col] = v;
                    }
                }
            }
            if constexpr (NV > 0) {
                // tiles 2..NV+1, now in slots 1..NV, back once the five
                // helpers are done (bounded wait)
                if (vpartner) {
                    int spins = 0;
                    while (__hip_atomic_load(&S.vdone, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 5 &&
                           spins < (1 << 16)) {
                        __builtin_amdgcn_s_sleep(1);
                        spins++;
                    }
                    if (spins >= (1 << 16) && l == 0) raise_error(a.status, VERB_ERR_HIP);
                    static_for<0, NV>([&](auto tc) {
                        constexpr int t = tc;
                        if (NB * (q + 2 + t) > ju_prev) return;
#pragma unroll
                        for (int r = 0; r < 4; r++) acc[1 + t][r] = S.vx[(t * 4 + r) * 64 + l];
                    });
                }
            }
        }
        __syncthreads();
        if (p == npan) break;
        const int lb = (w - p % RB + RB) % RB;  // local row block of this wavefront
        const bool lfin = S.lfin != 0;
        if (fac) {
            const int l = lane_fresh();
            // L11^-1 for U12 below (finite multipliers): lane m solves
            // L11 y = e_m column-wise; the other wavefronts are busier here.
            // (Before this wavefront's tiles come back from the stash, so
            // that the L11 entries can all be requested at once.)
            if (lfin && l < NB) {
                T y[NB];
                static_for<0, NB>([&](auto cc) {
                    constexpr int c = cc;
                    y[c] = (c == l) ? T(1) : T(0);
                });
                // the L11 entries of columns [I0, I1) are requested together
                // (one wait per group instead of one per column)
                T lc[NB * NB];
                // (16-byte reads at immediate offsets from the panel base:
                // column i starts on a 16-byte boundary)
                constexpr int E = 16 / sizeof(T);
                using V16 = T __attribute__((ext_vector_type(E)));
                // (FP32 keeps element reads: its register budget is that of
                // two workgroups per CU)
                auto fetch = [&](auto i0c, auto i1c) {
                    static_for<decltype(i0c)::value, decltype(i1c)::value>([&](auto ic) {
                        constexpr int i = ic;
                        if constexpr (sizeof(T) == 8) {
                            static_for<(i + 1) / E, NB / E>([&](auto vc) {
                                constexpr int cv = E * vc;
                                const V16 t = *reinterpret_cast<const V16*>(pn + i * PS + cv);
                                static_for<0, E>([&](auto ec) {
                                    constexpr int e = ec;
                                    constexpr int c = cv + e;
                                    if constexpr (c > i) lc[i * NB + c] = t[e];
                                });
                            });
                        } else {
                            static_for<i + 1, NB>([&](auto cc) {
                                constexpr int c = cc;
                                lc[i * NB + c] = pn[i * PS + c];
                            });
                        }
                    });
                };
                auto eliminate = [&](auto i0c, auto i1c) {
                    static_for<decltype(i0c)::value, decltype(i1c)::value>([&](auto ic) {
                        constexpr int i = ic;
                        const T yi = y[i];
                        static_for<i + 1, NB>([&](auto cc) {
                            constexpr int c = cc;
                            y[c] = fma(-lc[i * NB + c], yi, y[c]);
                        });
                    });
                };
                using I0 = std::integral_constant<int, 0>;
                using I3 = std::integral_constant<int, 3>;
                using I7 = std::integral_constant<int, 7>;
                using IE = std::integral_constant<int, NB - 1>;
                fetch(I0(), I3());
                fetch(I3(), I7());
                __builtin_amdgcn_sched_barrier(0);
                eliminate(I0(), I3());
                fetch(I7(), IE());
                __builtin_amdgcn_sched_barrier(0);
                eliminate(I3(), I7());
                eliminate(I7(), IE());
                static_for<0, NB>([&](auto cc) {
                    constexpr int c = cc;
                    S.linv[c * L::LS + l] = y[c];
                });
            }
#pragma unroll
            for (int s = 1; s < CB; s++)
#pragma unroll
                for (int r = 0; r < 4; r++) acc[s][r] = S.stash[((s - 1) * 4 + r) * 64 + l];
        }

        // the panel's ipiv and U fill extents (dgbtf2: ju = max(ju, min(j +
        // ku + jp, n-1)) after each column with a nonzero pivot), on the
        // wavefront that holds the panel rows: lane c takes column c, the
        // running maximum is a prefix maximum over the 16 lanes
        if (lb == 0) {
            const int l = lane_fresh();
            const int c = l & (NB - 1);
            const int jpc = S.jp[c];
            const unsigned zm = S.jz;
            const int jprev = S.ju;
            int v = (c < nc && !((zm >> c) & 1u)) ? min(j0 + c + ku + jpc, n - 1) : 0;
            v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));  // row_shr:1
            v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));  // row_shr:2
            v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));  // row_shr:4
            v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));  // row_shr:8
            const int uj = max(jprev, v);
            if (l < nc) {
                ipiv[j0 + l] = (int64_t)(j0 + l + jpc + 1);
                if (uext) uext[j0 + l] = uj;
            }
            const int jnew = __builtin_amdgcn_readlane(uj, nc - 1);
            if (l == 0) S.ju = jnew;
        }
        // 3a. pre-interchange contents of the rows that move -> sb (columns
        //     right of the panel)
        {
            const int l = lane_fresh(), g = l >> 4, jj = l & 15;
            int sl[4];
#pragma unroll
            for (int r = 0; r < 4; r++) sl[r] = S.slot[NB * lb + P::crow(g, r)];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (sl[r] >= 0) {
#pragma unroll
                    for (int s = 1; s < CB; s++) S.sb[sl[r] * W + NB * (s - 1) + jj] = acc[s][r];
                }
            }
        }
        __syncthreads();
        if (lb == 0) {
            // The panel rows' registers are free: the row block that enters
            // the window at the bottom for panel p+1 belongs to this
            // wavefront. Its tile in column block p+1 is requested while the
            // other wavefronts form U12 (nothing waits for it before 4a).
            const int l = lane_fresh(), g = l >> 4, jj = l & 15;
            const int rb = p + RB;
            // slot 0 of the layout of panel p+1 (column block p+1); the
            // other slots come through the stash while panel p+1 is factored
            acc[0] = load_tile<T>(rs, n, nx, rb, p + 1, jj, g);
        }

        // 3b. U12 = L11^-1 P A12 on the wavefronts that do not hold the panel
        //     rows: with finite multipliers as L11^-1 times the moved rows,
        //     one 16x16 tile per MFMA chain (column blocks dealt out to the
        //     wavefronts); otherwise by substitution, one lane per column
        // (with RB = 8 and L not kept, only tile 0, for step 4a; the others
        // at the start of the next stage)
        u12_later = RB == 8 && lfin && !a.write_L && p + 1 <= npan;
        if (tid == 0) S.u12done = 0;  // (nobody waits on it until the next stage)
        if (lb != 0 && lfin && u12_later) {
            const int l = lane_fresh(), g = l >> 4, jj = l & 15;
#pragma unroll
            for (int kk = 0; kk < 4; kk++) dslc[kk] = (4 * kk + g < nc) ? S.dslot[4 * kk + g] : -1;
            if (lb == 1) {
                A4 d = A4{T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const T la = S.linv[jj * L::LS + 4 * kk + g];
                    const T bv = (dslc[kk] >= 0) ? S.sb[dslc[kk] * W + jj] : T(0);
                    d = P::mfma(la, bv, d);
                }
#pragma unroll
                for (int r = 0; r < 4; r++) S.u12[P::crow(g, r) * WU + jj] = d[r];
            }
        } else if (lb != 0 && lfin) {
            const int l = lane_fresh(), g = l >> 4, jj = l & 15;
            T la[4];
            int dsl[4];
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                la[kk] = S.linv[jj * L::LS + 4 * kk + g];
                dsl[kk] = (4 * kk + g < nc) ? S.dslot[4 * kk + g] : -1;
            }
            // this wavefront's column blocks lb-1, lb-1+(RB-1), ...: their
            // MFMA chains are interleaved (a chain is four dependent MFMAs)
            constexpr int NBW = (CB - 1 + RB - 2) / (RB - 1);
            A4 d[NBW];
            T bv[NBW][4];
            static_for<0, NBW>([&](auto tc) {
                constexpr int t = tc;
                const int cb = lb - 1 + t * (RB - 1);
                const int cbr = (cb < CB - 1) ? cb : lb - 1;
#pragma unroll
                for (int kk = 0; kk < 4; kk++)
                    bv[t][kk] = (dsl[kk] >= 0) ? S.sb[dsl[kk] * W + NB * cbr + jj] : T(0);
                d[t] = A4{T(0), T(0), T(0), T(0)};
            });
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
                static_for<0, NBW>([&](auto tc) {
                    constexpr int t = tc;
                    d[t] = P::mfma(la[kk], bv[t][kk], d[t]);
                });
            static_for<0, NBW>([&](auto tc) {
                constexpr int t = tc;
                const int cb = lb - 1 + t * (RB - 1);
                if (cb < CB - 1)
#pragma unroll
                    for (int r = 0; r < 4; r++) S.u12[P::crow(g, r) * WU + NB * cb + jj] = d[t][r];
            });
        } else if (lb != 0) {
            const int l = lane_fresh();
            for (int t = NB * 4 * (lb - 1) + l; t < W; t += 64 * (RB - 1)) {
                T u[NB];
                static_for<0, NB>([&](auto cc) {
                    constexpr int c = cc;
                    u[c] = (c < nc) ? S.sb[S.dslot[c] * W + t] : T(0);
                });
                // column-oriented, as the reference TRSM: once u[m] is final
                // it updates every later row (a zero u[m] is skipped), so the
                // rows' updates are independent of each other
                static_for<0, NB - 1>([&](auto mc) {
                    constexpr int m = mc;
                    const T temp = -u[m];
                    if (temp != T(0))
                        static_for<m + 1, NB>([&](auto cc) {
                            constexpr int c = cc;
                            u[c] = fma(pn[m * PS + c], temp, u[c]);
                        });
                });
                static_for<0, NB>([&](auto cc) {
                    constexpr int c = cc;
                    S.u12[c * WU + t] = u[c];
                });
            }
        }
        // 3c. rows below the panel whose contents changed take them from sb:
        //     slot 1 now (step 4a updates it); slots 2..CB-1 at the start of
        //     the next stage's trailing update, which is the next to use them
        //     (sb is rewritten in the next step 3a, or, when L is kept, as
        //     scratch in step 4a: then all slots now)
        if (lb != 0) {
            const int l = lane_fresh(), g = l >> 4, jj = l & 15;
#pragma unroll
            for (int r = 0; r < 4; r++) dfx[r] = S.dslot[NB * lb + P::crow(g, r)];
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (dfx[r] >= 0) acc[1][r] = S.sb[dfx[r] * W + jj];
            if (a.write_L) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (dfx[r] >= 0) {
#pragma unroll
                        for (int s = 2; s < CB; s++) acc[s][r] = S.sb[dfx[r] * W + NB * (s - 1) + jj];
                    }
                    dfx[r] = -1;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++) dfx[r] = -1;
        }
        __syncthreads();

        // 4a. update of column block p+1 (slot 1), which becomes panel p+1:
        //     its tiles go to the other panel buffer at their rows in the
        //     window of p+1 (the wavefront that held the panel rows brings the
        //     row block that entered the window, in slot 0 of its new layout)
        T* const pnx = S.pnl[(p + 1) & 1];
        const int ju = S.ju;  // (formed in step 3a)
        if (lb != 0) {
            if (p + 1 < npan) {
                const int l = lane_fresh(), g = l >> 4, jj = l & 15;
                if (NB * (p + 1) <= ju) {
                    T an[4];
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) an[kk] = -pn[(4 * kk + g) * PS + NB * lb + jj];
                    A4 c4 = acc[1];
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) c4 = P::mfma(an[kk], S.u12[(4 * kk + g) * WU + jj], c4);
                    acc[1] = c4;
                }
#pragma unroll
                for (int r = 0; r < 4; r++) pnx[jj * PS + NB * (lb - 1) + P::crow(g, r)] = acc[1][r];
            }
        } else {
            const int l = lane_fresh(), g = l >> 4, jj = l & 15;
            if (a.write_L) {
                // multipliers in dgbtf2 order: undo the later interchanges of
                // the panel on the columns left of each pivot step, in a copy
                // of the panel (sb is free until the next panel's row moves)
                T* const cp = S.sb;
                for (int i = l; i < PS * NB; i += 64) cp[i] = pn[i];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int c = nc - 1; c >= 0; c--) {
                    const int jp = S.jp[c];
                    if (jp != 0 && l < c) {
                        const T t = cp[l * PS + c];
                        cp[l * PS + c] = cp[l * PS + c + jp];
                        cp[l * PS + c + jp] = t;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                for (int c = 0; c < nc; c++) {
                    const int km = min(kl, n - 1 - (j0 + c));
                    for (int i = l; i < kl; i += 64)
                        ab[(size_t)(j0 + c) * ldab + kv + 1 + i] = (i < km) ? cp[c * PS + c + 1 + i] : T(0);
                }
            }
            if (p + 1 < npan) {
#pragma unroll
                for (int r = 0; r < 4; r++) pnx[jj * PS + NB * (RB - 1) + P::crow(g, r)] = acc[0][r];
            }
            // This wavefront factors the next panel: its global stores are
            // waited for here, where it has time to spare, rather than at the
            // start of the factorisation.
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        __syncthreads();
        ju_prev = ju;
    }
    if (tid == 0 && S.info != 0) raise_pivot(a.status, S.info, node);
    // fused back-substitution: this node's U blocks and y are published
    // (every thread's stores made visible device-wide, then the flag)
    if (a.bs_flags) __threadfence();
    __syncthreads();  // the LDS state is reused by the next node
    if (a.bs_flags && tid == 0) __hip_atomic_store(&a.bs_flags[node], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if constexpr (sizeof(T) == 8) {
        // Fused back-substitution: once the node queue is drained, the
        // workgroup takes x = U^-1 y tasks in the same node order from a
        // second counter, each after its node's factors are published.
        // Workgroups that run out of LU nodes early (the last, partial round
        // of nodes over the CUs) thus solve the nodes finished before.
        if (a.bs_queue) {
            static_assert(sizeof(L) >= kBsLdsDoubles * sizeof(double), "back-substitution LDS");
            __shared__ int bs_slot, bs_skip;
            T* sh = reinterpret_cast<T*>(lds_raw);
            const int nx = a.nx, kl = nx + 1, ldab = 3 * kl + 1;
            for (;;) {
                __syncthreads();
                if (tid == 0) bs_slot = atomicAdd(a.bs_queue, 1);
                __syncthreads();
                const int k = bs_slot;
                if (k >= a.qnodes) break;
                const int node = a.order ? a.order[k] : k;
                if (a.nR > 0 && !a.include_boundary && (node % a.nR == 0 || node % a.nR == a.nR - 1)) continue;
                if (a.done && a.done[node]) continue;
                if (tid == 0) {
                    int spins = 0;
                    while (__hip_atomic_load(&a.bs_flags[node], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                           spins <= kBsMaxSpins) {
                        __builtin_amdgcn_s_sleep(32);
                        spins++;
                    }
                    bs_skip = spins > kBsMaxSpins;
                    if (bs_skip) raise_error(a.status, VERB_ERR_HIP);
                }
                __syncthreads();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (bs_skip) continue;
                XOut<T> xout;
                if (a.bs_psd) {
                    xout.psd = a.bs_psd + (size_t)node * a.n;
                    xout.nx = nx;
                    xout.nK = a.bs_nK;
                    xout.k0 = a.bs_k0 ? a.bs_k0[node] : 0;
                }
                backsolve_node<T, true, CB, kBsFusedSets>(a.ab, a.y, a.bs_x, a.nn ? a.nn[node] : a.n, a.n, kl, kl, ldab, node,
                                               a.ublk_cols, a.ublk_stride, sh, tid, xout);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// x = U^-1 y, one workgroup per node. U is read in blocks of 16 rows x
// (kl+ku+16) columns (the factorisation stores every U row up to kl+ku past
// its diagonal, zeros included) into three register sets used in turn, with
// buffer loads issued two blocks ahead; x is kept in LDS only over the band
// ahead of the current block (a ring of kXRing entries; the solution streams
// to HBM). Within a block: each of the 16 threads of a row reduces its
// columns right of the block straight from registers, then one wavefront
// solves the 16x16 triangle (staged in LDS) last row first, broadcasting
// each x with readlane.
// ---------------------------------------------------------------------------
constexpr int kBsMaxI = 20;  // staged 16-column groups per thread: (kl+ku+16)/16 <= 20

// One node's x = U^-1 y on the first kBsThreads threads of the workgroup
// (tid: this thread's index; threads past kBsThreads only take part in the
// barriers); sh: LDS of bs_lds_doubles doubles. n: the node's unknowns,
// stride: the per-node vector stride.
template <typename T, bool BLK, int NI, int D>
__device__ __forceinline__ void backsolve_node(const T* __restrict__ ab_all, const T* __restrict__ y_all,
                                               T* __restrict__ x_all, int n, int stride, int kl, int ku, int ldab,
                                               int node, int ublk_cols, size_t ublk_stride, T* sh, int tid,
                                               XOut<T> xout) {
    const bool act = tid < kBsThreads;
    const int kv = kl + ku;
    const int WS = BLK ? ublk_cols : kv + NB;  // columns per block (relative to the block start)
    const int nI = (WS + NB - 1) / NB;
    const T* ab = ab_all + (BLK ? (size_t)node * ublk_stride : (size_t)node * ldab * stride);
    const T* yv = y_all + (size_t)node * stride;
    T* xo = x_all + (size_t)node * stride;
    T* x = sh;               // ring of kXRing: x[i] at i % kXRing
    T* sd = sh + kXRing;     // [NB][NB] diagonal block of the current block row
    T* part = sd + NB * kBsPad;  // [NB][16] partial sums (rows padded to kBsPad)
    const int ri = tid & 15, t = tid >> 4;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(ab), (short)0, ldab * n * (int)sizeof(T), 0x00020000);
    const int nblk = (n + NB - 1) / NB;
    (void)rs;

    // this thread's entries of block blk: row jb+ri, columns jb + t + 16i
    // (and, for the 16 threads of the triangle, y of its row)
    // (every load is unconditional, from a clamped address, and its value
    // is used as loaded: a load under a branch, or a select on its value,
    // makes the compiler wait for all loads in flight at that point, which
    // would undo the prefetch. Entries of groups i >= nI and y past n are
    // never used: the partial sums take columns up to hi only, the triangle
    // rows below n only.)
    auto issue = [&](int blk, T (&v)[NI], T& yr) {
        const int jb = NB * blk, r = jb + ri;
        yr = yv[min(r, n - 1)];
        const int hi = min(r + kv, n - 1);
        if constexpr (BLK) {
            // the block is contiguous and clean: entry (ri, t + 16i) at
            // 256i + tid
            const T* src = ab + (size_t)blk * NB * WS;
            static_for<0, NI>([&](auto ic) {
                constexpr int i = ic;
                v[i] = src[NB * NB * (i < nI ? i : 0) + tid];
            });
        } else {
            static_for<0, NI>([&](auto ic) {
                constexpr int i = ic;
                const int co = t + NB * i;
                const int col = jb + co;
                const bool ok = (i < nI) & (co < WS) & (r < n) & (col >= r) & (col <= hi);
                const unsigned off = (unsigned)(col * (ldab - 1) + kv + r) * (unsigned)sizeof(T);
                v[i] = load_entry<T>(rs, ok ? off : kOutside);
            });
        }
    };
    // one block: partial sums over the band right of it straight from the
    // registers the block was loaded into, then the triangle
    auto solve_block = [&](int blk, const T (&v)[NI], T yr) {
        const int jb = NB * blk;
        // (branch-free: every x of the band is read and the terms outside it
        // are fma(0, 0, s), so the LDS reads are issued together instead of
        // one per branch)
        {
            const int r = jb + ri;
            sd[ri * kBsPad + t] = v[0];  // column jb + t
            const int hi = (r < n) ? min(r + kv, n - 1) - jb : -1;
            T sacc = T(0);
            static_for<1, NI>([&](auto ic) {
                constexpr int i = ic;
                const int co = t + NB * i;
                const T xv = x[(jb + co) & (kXRing - 1)];
                const bool in = co <= hi;
                sacc = fma(in ? v[i] : T(0), in ? xv : T(0), sacc);
            });
            part[ri * kBsPad + t] = sacc;
        }
        __syncthreads();
        if (tid < 64) {
            // (the triangle is the serial part of the block: issue priority
            // over the wavefronts of the other workgroups on the CU)
            __builtin_amdgcn_s_setprio(3);
            const int i = tid & 15;
            const int r = jb + i;
            const bool live = (tid < NB) & (r < n);
            T sp = T(0);
#pragma unroll
            for (int q = 0; q < 16; q++) sp += part[i * kBsPad + q];
            T urow[NB];
#pragma unroll
            for (int k = 0; k < NB; k++) urow[k] = sd[i * kBsPad + k];
            const T dv = sd[i * kBsPad + i];
            T b = live ? yr - sp : T(0);
            const T rd = live ? T(1) / dv : T(0);
            // rows past n (last block only) solve to 0 and are not stored;
            // lane i's b is final once step i is past, so its x is stored
            // after the loop (the value step i broadcast)
#pragma unroll
            for (int k = NB - 1; k >= 0; k--) {
                const T xk = Prec<T>::readlane(b * rd, k);
                b = (i < k) ? fma(-urow[k], xk, b) : b;
            }
            const T xi = b * rd;
            if (tid < NB) {
                x[(jb + i) & (kXRing - 1)] = xi;
                if (r < n) {
                    if (xout.psd) {
                        const int iy = r / xout.nx, ix = r - iy * xout.nx;
                        xout.psd[(size_t)ix * xout.nK + xout.k0 + iy] = xi;
                    } else {
                        xo[r] = xi;
                    }
                }
            }
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
    };

    const int top = nblk - 1;
    if (!act) {
        // threads past kBsThreads: the barriers of every block only
        for (int b = top; b >= 0; b--) {
            __syncthreads();
            __syncthreads();
        }
        return;
    }
    // U blocks are requested D-1 blocks ahead of their solve into D
    // register sets used in turn (past block 0, block 0 again: loads that
    // are never used, so that the sequence has no branch)
    T v[D][NI];
    T yq[D];
    static_for<0, D - 1>([&](auto sc) {
        constexpr int sset = sc;
        issue(max(top - sset, 0), v[sset], yq[sset]);
    });
    for (int blk = top; blk >= 0; blk -= D) {
        static_for<0, D>([&](auto kc) {
            constexpr int k = kc;
            constexpr int nxt = (k + D - 1) % D;  // the set freed by the previous solve
            const int b = blk - k;
            if (b < 0) return;
            issue(max(b - (D - 1), 0), v[nxt], yq[nxt]);
            solve_block(b, v[k], yq[k]);
        });
    }
}


template <typename T, bool BLK, int NI, int D>
__global__ __launch_bounds__(kBsThreads) void band_backsolve(const T* __restrict__ ab_all,
                                                             const T* __restrict__ y_all,
                                                             T* __restrict__ x_all, int n, int kl, int ku,
                                                             int ldab, int nR, int include_boundary,
                                                             const int* __restrict__ nn,
                                                             const int* __restrict__ order,
                                                             const int* __restrict__ done, int ublk_cols,
                                                             size_t ublk_stride) {
    const int node = order ? order[blockIdx.x] : blockIdx.x;
    if (nR > 0 && !include_boundary && (node % nR == 0 || node % nR == nR - 1)) return;
    if (done && done[node]) return;
    extern __shared__ double sh_raw[];  // 8-byte aligned
    backsolve_node<T, BLK, NI, D>(ab_all, y_all, x_all, nn ? nn[node] : n, n, kl, ku, ldab, node, ublk_cols,
                                  ublk_stride, reinterpret_cast<T*>(sh_raw), threadIdx.x);
}

// ---------------------------------------------------------------------------
// y = L^-1 P b with factors in LAPACK band layout and dgbtf2-order
// multipliers (the dgbtrs forward loop: for j < n-1, swap b(j) with
// b(ipiv(j)), then b(j+1..j+lm) -= L(:,j) * b(j), skipped when b(j) = 0).
// One wavefront per node. The 128 entries of b that the next 16 columns
// touch live in two registers per lane (positions l and l+64 of the window
// starting at column j0); interchanges and the pivot value go through
// readlane, and each lane's multipliers for all 16 columns of a block
// are loaded (column j's multipliers are contiguous, so the
// lanes of one column load consecutive words), two blocks ahead of use. When B0 is given, the
// right-hand side is formed first as rhs*B0 + C0 (src/MatrixSolver.cpp:318).
// ---------------------------------------------------------------------------
constexpr int kFwdBlock = 16;

template <typename T, int NW>
__global__ __launch_bounds__(64, 1) void band_forward(const T* __restrict__ ab_all,
                                                   const int64_t* __restrict__ ipiv_all, int n, int kl, int ku,
                                                   int ldab, const double* __restrict__ b0_all,
                                                   const double* __restrict__ c0_all,
                                                   const T* __restrict__ rhs_all, T* __restrict__ y_all,
                                                   int nR, int include_boundary, const int* __restrict__ done) {
    const int node = blockIdx.x;
    if (nR > 0 && !include_boundary && (node % nR == 0 || node % nR == nR - 1)) return;
    if (done && done[node]) return;
    const int kd = kl + ku;
    const T* ab = ab_all + (size_t)node * ldab * n;
    const int64_t* ipiv = ipiv_all + (size_t)node * n;
    const T* rhs = rhs_all + (size_t)node * n;
    const double* B0 = b0_all ? b0_all + (size_t)node * n : nullptr;
    const double* C0 = c0_all ? c0_all + (size_t)node * n : nullptr;
    T* y = y_all + (size_t)node * n;
    const int l = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(ab), (short)0, ldab * n * (int)sizeof(T), 0x00020000);
    auto b_at = [&](int i) -> T {
        if (i >= n) return T(0);
        if (!B0) return rhs[i];
        return (T)__dadd_rn(__dmul_rn((double)rhs[i], B0[i]), C0[i]);
    };
    // window positions 64q + l, q < NW, of the entries starting at column j0
    T w[NW];
#pragma unroll
    for (int q = 0; q < NW; q++) w[q] = b_at(64 * q + l);
    const int ncol = n - 1;  // columns with an elimination step
    const int nblk = (ncol + kFwdBlock - 1) / kFwdBlock;
    T L0[NW][kFwdBlock], L1[NW][kFwdBlock], L2[NW][kFwdBlock];
    int pv0 = 0, pv1 = 0, pv2 = 0;
    T f0 = T(0), f1 = T(0), f2 = T(0);
    // multipliers of block blk for this lane: column j0+c at the rows of
    // window positions 64q + l (i = position - c - 1), the pivots of the
    // block, and the entries that enter the window after the block
    auto fetch = [&](int blk, T (&La)[NW][kFwdBlock], int& pv, T& fr) {
        const int j0 = kFwdBlock * blk;
        fr = (l >= 64 - kFwdBlock) ? b_at(j0 + kFwdBlock + 64 * (NW - 1) + l) : T(0);
#pragma unroll
        for (int c = 0; c < kFwdBlock; c++) {
            const int j = j0 + c;
            const int lm = min(kl, n - 1 - j);
            const unsigned base = (unsigned)(j * ldab + kd + 1);
#pragma unroll
            for (int q = 0; q < NW; q++) {
                const int i = 64 * q + l - c - 1;
                const bool ok = (j < ncol) & (i >= 0) & (i < lm);
                La[q][c] = load_entry<T>(rs, ok ? (base + (unsigned)i) * (unsigned)sizeof(T) : kOutside);
            }
        }
        pv = (l < kFwdBlock && j0 + l < ncol) ? (int)ipiv[j0 + l] - 1 - j0 : l;
    };
    auto process = [&](int blk, const T (&La)[NW][kFwdBlock], int pv, T fresh) {
        const int j0 = kFwdBlock * blk;
#pragma unroll
        for (int c = 0; c < kFwdBlock; c++) {
            const int j = j0 + c;
            if (j < ncol) {
                const int lm = min(kl, n - 1 - j);
                const int p = __builtin_amdgcn_readlane(pv, c);  // window position of row ipiv(j)
                if (p != c) {
                    const T vc = Prec<T>::readlane(w[0], c);
                    T vp = T(0);
#pragma unroll
                    for (int q = 0; q < NW; q++)
                        if ((p >> 6) == q) vp = Prec<T>::readlane(w[q], p & 63);
                    if (l == c) w[0] = vp;
#pragma unroll
                    for (int q = 0; q < NW; q++)
                        if ((p >> 6) == q && l == (p & 63)) w[q] = vc;
                }
                const T bj = Prec<T>::readlane(w[0], c);
                if (bj != T(0)) {
                    const T t = -bj;
                    // only rows j+1..j+lm change
#pragma unroll
                    for (int q = 0; q < NW; q++) {
                        const int i = 64 * q + l - c - 1;
                        w[q] = (i >= 0 && i < lm) ? fma(La[q][c], t, w[q]) : w[q];
                    }
                }
            }
        }
        // columns j0..j0+15 are final: out, and the window moves by 16
        if (l < kFwdBlock && j0 + l < n) y[j0 + l] = w[0];
        T sh[NW];
#pragma unroll
        for (int q = 0; q < NW; q++) sh[q] = __shfl(w[q], (l + kFwdBlock) & 63);
#pragma unroll
        for (int q = 0; q < NW; q++) w[q] = (l < 64 - kFwdBlock) ? sh[q] : (q + 1 < NW ? sh[q + 1 < NW ? q + 1 : q] : fresh);
    };
    // multipliers are requested two blocks ahead of their use
    if (nblk > 0) fetch(0, L0, pv0, f0);
    if (nblk > 1) fetch(1, L1, pv1, f1);
    for (int blk = 0; blk < nblk; blk += 3) {
        if (blk + 2 < nblk) fetch(blk + 2, L2, pv2, f2);
        process(blk, L0, pv0, f0);
        if (blk + 1 >= nblk) break;
        if (blk + 3 < nblk) fetch(blk + 3, L0, pv0, f0);
        process(blk + 1, L1, pv1, f1);
        if (blk + 2 >= nblk) break;
        if (blk + 4 < nblk) fetch(blk + 4, L1, pv1, f1);
        process(blk + 2, L2, pv2, f2);
    }
    const int j0 = kFwdBlock * nblk;
#pragma unroll
    for (int q = 0; q < NW; q++)
        if (j0 + 64 * q + l < n) y[j0 + 64 * q + l] = w[q];
}

template <typename T>
int launch_forward_t(const T* ab, const int64_t* ipiv, int n, int kl, int ku, int ldab, const double* b0,
                     const double* c0, const T* rhs, T* y, int nodes, int nR, int include_boundary,
                     hipStream_t stream, const int* done = nullptr) {
    if ((size_t)ldab * n * sizeof(T) >= 0x7ffffff8u) return VERB_ERR_ARG;
    if (kl + kFwdBlock <= 128) {
        hipLaunchKernelGGL((band_forward<T, 2>), dim3(nodes), dim3(64), 0, stream, ab, ipiv, n, kl, ku, ldab, b0, c0,
                           rhs, y, nR, include_boundary, done);
    } else if (kl + kFwdBlock <= 192) {
        hipLaunchKernelGGL((band_forward<T, 3>), dim3(nodes), dim3(64), 0, stream, ab, ipiv, n, kl, ku, ldab, b0, c0,
                           rhs, y, nR, include_boundary, done);
    } else {
        return VERB_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_forward(const double* ab, const int64_t* ipiv, int n, int kl, int ku, int ldab, const double* b0,
                   const double* c0, const double* rhs, double* y, int nodes, int nR, int include_boundary,
                   hipStream_t stream) {
    return launch_forward_t<double>(ab, ipiv, n, kl, ku, ldab, b0, c0, rhs, y, nodes, nR, include_boundary, stream);
}

int launch_forward32(const float* ab, const int64_t* ipiv, int n, int kl, int ku, int ldab, const float* rhs,
                     float* y, int nodes, int nR, int include_boundary, hipStream_t stream, const int* done) {
    return launch_forward_t<float>(ab, ipiv, n, kl, ku, ldab, nullptr, nullptr, rhs, y, nodes, nR, include_boundary,
                                   stream, done);
}

template <typename T, bool BLK, int NI, int D>
int launch_backsolve_v(const T* ab, const T* y, T* x, int n, int kl, int ku, int ldab, int nodes, int nR,
                       int include_boundary, hipStream_t stream, const int* nn, const int* order, const int* done,
                       int ublk_cols, size_t ublk_stride, size_t lds) {
    hipFuncSetAttribute((const void*)band_backsolve<T, BLK, NI, D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds);
    hipLaunchKernelGGL((band_backsolve<T, BLK, NI, D>), dim3(nodes), dim3(kBsThreads), lds, stream, ab, y, x, n, kl,
                       ku, ldab, nR, include_boundary, nn, order, done, ublk_cols, ublk_stride);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

// Register sets per thread: NI 16-column groups of a block (the smallest
// instantiation that covers kl+ku+16 columns) times D blocks in flight.
template <typename T, bool BLK>
int launch_backsolve_b(const T* ab, const T* y, T* x, int n, int kl, int ku, int ldab, int nodes, int nR,
                       int include_boundary, hipStream_t stream, const int* nn, const int* order, const int* done,
                       int ublk_cols, size_t ublk_stride, size_t lds) {
    const int cols = BLK ? ublk_cols : kl + ku + NB;
    if (cols <= NB * 8)
        return launch_backsolve_v<T, BLK, 8, 6>(ab, y, x, n, kl, ku, ldab, nodes, nR, include_boundary, stream, nn,
                                                order, done, ublk_cols, ublk_stride, lds);
    if (cols <= NB * 14)
        return launch_backsolve_v<T, BLK, 14, 3>(ab, y, x, n, kl, ku, ldab, nodes, nR, include_boundary, stream, nn,
                                                 order, done, ublk_cols, ublk_stride, lds);
    return launch_backsolve_v<T, BLK, kBsMaxI, 3>(ab, y, x, n, kl, ku, ldab, nodes, nR, include_boundary, stream, nn,
                                                  order, done, ublk_cols, ublk_stride, lds);
}

template <typename T>
int launch_backsolve_t(const T* ab, const T* y, T* x, int n, int kl, int ku, int ldab, int nodes, int nR,
                       int include_boundary, hipStream_t stream, const int* nn, const int* order, const int* done,
                       int ublk_cols = 0, size_t ublk_stride = 0) {
    if (kl + ku + NB > NB * kBsMaxI || kl + ku + 2 * NB > kXRing) return VERB_ERR_ARG;
    if (ublk_cols && (ublk_cols > NB * kBsMaxI || ublk_cols < kl + ku + NB)) return VERB_ERR_ARG;
    const size_t lds = sizeof(T) * ((size_t)kXRing + 2 * NB * kBsPad);
    int dev = 0, max_lds = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (lds + 256 > (size_t)max_lds) return VERB_ERR_ARG;
    if (ublk_cols)
        return launch_backsolve_b<T, true>(ab, y, x, n, kl, ku, ldab, nodes, nR, include_boundary, stream, nn, order,
                                           done, ublk_cols, ublk_stride, lds);
    return launch_backsolve_b<T, false>(ab, y, x, n, kl, ku, ldab, nodes, nR, include_boundary, stream, nn, order,
                                        done, 0, 0, lds);
}

int launch_backsolve(const double* ab, const double* y, const int* uext, double* x, int n, int kl, int ku, int ldab,
                     int nodes, int nR, int include_boundary, hipStream_t stream, const int* nn, const int* order,
                     const int* done, int ublk_cols, size_t ublk_stride) {
    (void)uext;  // U rows are stored up to kl+ku past the diagonal, so extents are not needed
    return launch_backsolve_t<double>(ab, y, x, n, kl, ku, ldab, nodes, nR, include_boundary, stream, nn, order,
                                      done, ublk_cols, ublk_stride);
}

int launch_backsolve32(const float* ab, const float* y, float* x, int n, int kl, int ku, int ldab, int nodes, int nR,
                       int include_boundary, hipStream_t stream, const int* done) {
    return launch_backsolve_t<float>(ab, y, x, n, kl, ku, ldab, nodes, nR, include_boundary, stream, nullptr, nullptr,
                                     done);
}

template <typename T, int RB, int CB>
int launch_band_lu_t(const LuArgsT<T>& a, int nodes, hipStream_t stream) {
    const size_t lds = sizeof(Lds<T, RB, CB>);
    int dev = 0, max_lds = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (lds > (size_t)max_lds) return VERB_ERR_ARG;
    hipFuncSetAttribute((const void*)band_lu_window<T, RB, CB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int grid = nodes;
    if (a.queue) {
        int cus = 0;
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        grid = min(nodes, max(1, cus));
        if (a.bs_queue && RB * 64 < kBsThreads) {
            // too few threads for the back-substitution in the same kernel:
            // the LU, then the back-substitution kernel (which writes x in
            // energy-fast order only)
            if (a.bs_psd) return VERB_ERR_ARG;
            LuArgsT<T> b = a;
            b.bs_queue = nullptr;
            b.bs_flags = nullptr;
            b.bs_x = nullptr;
            int rc = launch_band_lu_t<T, RB, CB>(b, nodes, stream);
            if (rc) return rc;
            if constexpr (sizeof(T) == 8)
                return launch_backsolve(a.ab, a.y, nullptr, a.bs_x, a.n, a.nx + 1, a.nx + 1, 3 * (a.nx + 1) + 1,
                                        nodes, a.nR, a.include_boundary, stream, a.nn, a.order, a.done, a.ublk_cols,
                                        a.ublk_stride);
            return VERB_ERR_ARG;
        }
        if (a.bs_queue) {
            if (sizeof(T) != 8 || !a.bs_flags || (!a.bs_x && !a.bs_psd) || !a.y || a.write_L || a.ublk_cols <= 0 ||
                a.ublk_cols > CB * NB || (a.bs_psd && a.bs_nK < 1))
                return VERB_ERR_ARG;
            hipLaunchKernelGGL(zero_lu_counters, dim3((nodes + 255) / 256), dim3(256), 0, stream, a.queue, a.bs_queue,
                               a.bs_flags, nodes);
        } else {
            hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, stream, a.queue);
        }
    } else if (a.bs_queue) {
        return VERB_ERR_ARG;
    }
    hipLaunchKernelGGL((band_lu_window<T, RB, CB>), dim3(grid), dim3(RB * 64), lds, stream, a);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

namespace {

template <typename T>
struct LuVariant {
    int rb, cb;
    int (*launch)(const LuArgsT<T>&, int, hipStream_t);
};

// Register-window instantiations: RB*16 >= kl+16 rows, CB*16 >= 2kl+16
// columns. FP64 keeps CB*8 accumulator registers per lane and stops at
// kl = 112; FP32 needs half as many and reaches kl = 152 with 11 wavefronts.
const LuVariant<double> kLuVariants[] = {
    {3, 5, launch_band_lu_t<double, 3, 5>},   {4, 7, launch_band_lu_t<double, 4, 7>},
    {5, 9, launch_band_lu_t<double, 5, 9>},   {6, 11, launch_band_lu_t<double, 6, 11>},
    {7, 13, launch_band_lu_t<double, 7, 13>}, {8, 14, launch_band_lu_t<double, 8, 14>},
    {8, 15, launch_band_lu_t<double, 8, 15>},
};
const LuVariant<float> kLuVariants32[] = {
    {5, 9, launch_band_lu_t<float, 5, 9>},
    {7, 13, launch_band_lu_t<float, 7, 13>},  {8, 14, launch_band_lu_t<float, 8, 14>},
    {8, 15, launch_band_lu_t<float, 8, 15>},  {9, 17, launch_band_lu_t<float, 9, 17>},
    {10, 19, launch_band_lu_t<float, 10, 19>}, {11, 20, launch_band_lu_t<float, 11, 20>},
};

template <typename T>
bool lu_variant_fits(const LuVariant<T>& v, int kl) {
    return NB * v.rb >= kl + NB && NB * v.cb >= 2 * kl + NB;
}

}  // namespace

int lu_variant_count() { return (int)(sizeof(kLuVariants) / sizeof(kLuVariants[0])); }

int lu_variant_shape(int i, int kl, int* rb, int* cb) {
    if (i < 0 || i >= lu_variant_count()) return 0;
    if (rb) *rb = kLuVariants[i].rb;
    if (cb) *cb = kLuVariants[i].cb;
    return lu_variant_fits(kLuVariants[i], kl) ? 1 : 0;
}

// variant < 0: the smallest instantiation that covers kl. Returns
// VERB_ERR_ARG when no register-window instantiation covers kl.
int launch_band_lu(const LuArgs& a, int nodes, hipStream_t stream, int variant) {
    const int kl = a.nx + 1;
    if (variant >= 0) {
        if (variant >= lu_variant_count() || !lu_variant_fits(kLuVariants[variant], kl)) return VERB_ERR_ARG;
        return kLuVariants[variant].launch(a, nodes, stream);
    }
    for (const auto& v : kLuVariants)
        if (lu_variant_fits(v, kl)) return v.launch(a, nodes, stream);
    return VERB_ERR_ARG;
}

bool band_lu_fuses_backsolve(int variant, int kl) {
    if (variant >= 0) return variant < lu_variant_count() && kLuVariants[variant].rb * 64 >= kBsThreads;
    for (const auto& v : kLuVariants)
        if (lu_variant_fits(v, kl)) return v.rb * 64 >= kBsThreads;
    return false;
}

int launch_band_lu32(const LuArgsT<float>& a, int nodes, hipStream_t stream) {
    const int kl = a.nx + 1;
    for (const auto& v : kLuVariants32)
        if (lu_variant_fits(v, kl)) return v.launch(a, nodes, stream);
    return VERB_ERR_ARG;
}

}  // namespace verb
