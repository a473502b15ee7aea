// Convection: Convection_2D over every (V,K) plane.
//
// Restates src/Convection_2D.cpp:36-169 (outer Courant sub-steps, MLT sweep
// over iR < R-1, G_conv weighting, R sweep, exp losses, sources) and the
// ULTIMATE-QUICKEST face scheme of src/Convection_1D_ULTIMATE_QUICKEST6.cpp
// (inner sub-steps, ghost cells, order-9 unlimited faces, discriminator and
// ULTIMATE limiter, conservative update, boundary fix-ups) in the
// reference's operation order. Compiled with -ffp-contract=off: the
// limiter's branch choices and the integer sub-step counts are the
// reference's, bit for bit.
//
// Mapping: a workgroup owns a tile of TQ consecutive (V,K) planes and keeps
// it in LDS for the whole call, so HBM sees one read and one write of PSD
// per point (plus VP, VR, G_conv, losses, sources read once or from L2).
// The reference's sequential face loop carries only bookkeeping: every face
// depends on point values alone, so each QUICKEST step runs as four
// data-parallel passes with one lane per grid point:
//   (a) unlimited faces, (b) unlimited values, (c) limited faces,
//   (d) conservative update, with the boundary fix-ups folded into (b)/(d).
// Results ping-pong between two LDS buffers, so no pass reads a value
// another lane of the same pass writes.
#include <utility>

#include "common.h"

namespace verb {
namespace {

constexpr int kThreads = 256;
constexpr int GST = 5;
// Largest sub-step count a launch accepts (outer Courant-0.25 steps of
// Convection_2D, inner Courant-1 steps of QUICKEST6).
constexpr double kMaxSubsteps = 1 << 20;

struct LineBC {
    int lt, ut;        // boundary types
    double lbc, ubc;   // boundary values
    double sl, su;     // lbc*dx_lower, ubc*dx_upper (ConstantDerivative ghosts)
};

// Value of point j of a line whose points sit at X[base + i*stride], i in
// [0, n), with the ghost rules of QUICKEST6.cpp:89-143 for j outside.
__device__ inline double orig_at(const double* X, int base, int stride, int n, int j, const LineBC& bc) {
    if (j >= 0 && j < n) return X[base + j * stride];
    if (j < 0) {
        if (bc.lt == VERB_BC_PERIODIC) return X[base + (j + n - 1) * stride];
        if (bc.lt == VERB_BC_CONSTANT_VALUE) return bc.lbc;
        double v = X[base] - bc.sl;
        for (int ig = 2; ig <= -j; ig++) v = v - bc.sl;
        return v;
    }
    const int m = j - (n - 1);
    if (bc.ut == VERB_BC_PERIODIC) return X[base + m * stride];
    if (bc.ut == VERB_BC_CONSTANT_VALUE) return bc.ubc;
    double v = X[base + (n - 1) * stride] + bc.su;
    for (int ig = 2; ig <= m; ig++) v = v + bc.su;
    return v;
}

// Ghost-extended unlimited solution (QUICKEST6.cpp:276-330): lower ghosts
// come from the pre-step values X, upper ghosts from the unlimited values Y.
__device__ inline double unl_at(const double* X, const double* Y, int base, int stride, int n, int j,
                                const LineBC& bc) {
    if (j >= 0 && j < n) return Y[base + j * stride];
    if (j < 0) {
        if (bc.lt == VERB_BC_PERIODIC) return X[base + (j + n - 1) * stride];
        if (bc.lt == VERB_BC_CONSTANT_VALUE) return bc.lbc;
        double v = X[base] - bc.sl;
        for (int ig = 2; ig <= -j; ig++) v = v - bc.sl;
        return v;
    }
    const int m = j - (n - 1);
    if (bc.ut == VERB_BC_PERIODIC) return Y[base + m * stride];
    if (bc.ut == VERB_BC_CONSTANT_VALUE) return bc.ubc;
    double v = Y[base + (n - 1) * stride] + bc.su;
    for (int ig = 2; ig <= m; ig++) v = v + bc.su;
    return v;
}

// Points j0 .. j0+N-1 of a line with the ghost rules of orig_at (Y == X)
// or unl_at (lower ghosts from X, the line and upper ghosts from Y). The
// two boundary kinds of the common grids (periodic at both ends, constant
// value at both ends) gather without per-point branches; any other pair
// goes through the point rules above. (n >= 6, so a periodic ghost within
// five points of an end is a point of the line.)
template <int N>
__device__ inline void gather(const double* X, const double* Y, int base, int stride, int n, int j0,
                              const LineBC& bc, double (&out)[N]) {
    if (bc.lt == VERB_BC_PERIODIC && bc.ut == VERB_BC_PERIODIC) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int j = j0 + i;
            const int jj = (j < 0) ? j + n - 1 : ((j >= n) ? j - (n - 1) : j);
            const double* src = (j < 0) ? X : Y;
            out[i] = src[base + jj * stride];
        }
    } else if (bc.lt == VERB_BC_CONSTANT_VALUE && bc.ut == VERB_BC_CONSTANT_VALUE) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int j = j0 + i;
            const int jc = min(max(j, 0), n - 1);
            const double v = Y[base + jc * stride];
            out[i] = (j < 0) ? bc.lbc : ((j >= n) ? bc.ubc : v);
        }
    } else if (X == Y) {
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = orig_at(X, base, stride, n, j0 + i, bc);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = unl_at(X, Y, base, stride, n, j0 + i, bc);
    }
}

// Order-9 unlimited face value (QUICKEST6.cpp:192-239).
__device__ inline double unlimited_face(const double W[10], double courant) {
    const double cs = courant * courant;
    const double ca = fabs(courant);
    double U1, U2, U3, U4, C, D1, D2, D3, D4, D5;
    if (courant >= 0) {
        C = W[4];
        U1 = W[3]; U2 = W[2]; U3 = W[1]; U4 = W[0];
        D1 = W[5]; D2 = W[6]; D3 = W[7]; D4 = W[8]; D5 = W[9];
    } else {
        C = W[5];
        D1 = W[4]; D2 = W[3]; D3 = W[2]; D4 = W[1]; D5 = W[0];
        U1 = W[6]; U2 = W[7]; U3 = W[8]; U4 = W[9];
    }
    double f = 0.5 * (C + D1);
    const double d1 = C - D1;
    f += 0.5 * ca * d1;
    const double Dd2 = (U1 - C - D1 + D2) / 2;
    const double d3 = U1 - 3 * C + 3 * D1 - D2;
    f += (cs - 1) / 6.0 * (Dd2 + ca / 4 * d3);
    const double Dd4 = (U2 - 3 * U1 + 2 * C + 2 * D1 - 3 * D2 + D3) / 2;
    const double d5 = U2 - 5 * U1 + 10 * C - 10 * D1 + 5 * D2 - D3;
    f += (cs - 1) * (cs - 4) / 120.0 * (Dd4 + ca / 6 * d5);
    const double Dd6 = (U3 - 5 * U2 + 9 * U1 - 5 * C - 5 * D1 + 9 * D2 - 5 * D3 + D4) / 2;
    const double d7 = U3 - 7 * U2 + 21 * U1 - 35 * C + 35 * D1 - 21 * D2 + 7 * D3 - D4;
    f += (cs - 1) * (cs - 4) * (cs - 9) / 5040.0 * (Dd6 + ca / 8 * d7);
    const double Dd8 =
        (U4 - 7 * U3 + 20 * U2 - 28 * U1 + 14 * C + 14 * D1 - 28 * D2 + 20 * D3 - 7 * D4 + D5) / 2;
    const double d9 = U4 - 9 * U3 + 36 * U2 - 84 * U1 + 126 * C - 126 * D1 + 84 * D2 - 36 * D3 + 9 * D4 - D5;
    f += (cs - 1) * (cs - 4) * (cs - 9) * (cs - 16) / 362880.0 * (Dd8 + d9 / 2);
    return f;
}

// Discriminator + ULTIMATE limiter (QUICKEST6.cpp:394-447).
__device__ inline double limited_face(const double UW[10], double ou, double oc, double od, double face,
                                      double courant) {
    const double ca = fabs(courant);
    double UU1, UU2, UU3, UC, UD1, UD2, UD3;
    if (courant >= 0) {
        UC = UW[4];
        UU1 = UW[3]; UU2 = UW[2]; UU3 = UW[1];
        UD1 = UW[5]; UD2 = UW[6]; UD3 = UW[7];
    } else {
        UC = UW[5];
        UD1 = UW[4]; UD2 = UW[3]; UD3 = UW[2];
        UU1 = UW[6]; UU2 = UW[7]; UU3 = UW[8];
    }
    const double DEL = od - ou;
    const double CURV = od - 2 * oc + ou;
    const double ADEL = fabs(DEL);
    const double ACURV = fabs(CURV);
    const double dif1 = UD2 - UD3;
    const double dif2 = UD1 - UD2;
    const double dif3 = UC - UD1;
    const double dif4 = UU1 - UC;
    const double dif5 = UU2 - UU1;
    const double dif6 = UU3 - UU2;
    const double th = 0;
    const bool peak_d = dif1 > th && dif2 > th && dif3 < -th && dif4 < -th && dif2 < dif1 - th && dif3 > dif4 + th;
    const bool peak_c = dif2 > th && dif3 > th && dif4 < -th && dif5 < -th && dif3 < dif2 - th && dif4 > dif5 + th;
    const bool peak_u = dif3 > th && dif4 > th && dif5 < -th && dif6 < -th && dif4 < dif3 - th && dif5 > dif6 + th;
    double fr = face;
    if (!(peak_d || peak_c || peak_u)) {
        if (ACURV >= ADEL) {
            fr = oc;
        } else if (ca > 1e-31) {
            const double PSD_REF = ou + (oc - ou) / ca;
            if (DEL > 0) {
                fr = fmax(fr, oc);
                fr = fmin(fr, fmin(PSD_REF, od));
            } else if (DEL < 0) {
                fr = fmin(fr, oc);
                fr = fmax(fr, fmax(PSD_REF, od));
            }
        }
    }
    return fr;
}

struct Tile {
    int nP, nR, TQ, npt;
    int tq_shift;   // log2(TQ) (TQ divides the 256 threads, so it is a power of two)
    unsigned mR;    // ceil(2^32 / nR): pr / nR = umulhi(pr, mR) for pr * nR < 2^32
    long NQ;
    long q0;
    const double* bc;  // per-line boundary data in LDS: [dir][line][lbc, ubc, sl, su]
};

// Geometry of the line through point (p, r, tq) for sweep dir (0 = MLT
// along p, 1 = R along r).
struct LineGeom {
    int n, ix, base, stride, line;
};

__device__ inline LineGeom line_of(const Tile& T, int dir, int p, int r, int tq) {
    LineGeom L;
    if (dir == 0) {
        L.n = T.nP; L.ix = p; L.base = r * T.TQ + tq; L.stride = T.nR * T.TQ; L.line = r * T.TQ + tq;
    } else {
        L.n = T.nR; L.ix = r; L.base = p * T.nR * T.TQ + tq; L.stride = T.TQ; L.line = p * T.TQ + tq;
    }
    return L;
}

__device__ inline void decompose(const Tile& T, int t, int& p, int& r, int& tq) {
    tq = t & (T.TQ - 1);
    const int pr = t >> T.tq_shift;
    p = (T.nR == 1) ? pr : (int)__umulhi((unsigned)pr, T.mR);
    r = pr - p * T.nR;
}

// Boundary data of a line (values, and the ConstantDerivative steps).
__device__ inline LineBC line_bc(const verb_grid& g, const Tile& T, int dir, int p, int r, int tq,
                                 const double* Gc) {
    LineBC bc;
    const long q = T.q0 + tq;
    const long NQ = T.NQ;
    if (dir == 0) {
        bc.lt = g.Pl_type; bc.ut = g.Pu_type;
        // periodic lines never read their boundary values (QUICKEST6.cpp:96-100)
        bc.lbc = (bc.lt == VERB_BC_PERIODIC) ? 0.0 : g.Pl[(size_t)r * NQ + q];
        bc.ubc = (bc.ut == VERB_BC_PERIODIC) ? 0.0 : g.Pu[(size_t)r * NQ + q];
        bc.sl = bc.su = 0.0;
        if (bc.lt == VERB_BC_CONSTANT_DERIVATIVE) {
            const double dx = g.Pg[((size_t)1 * T.nR + r) * NQ + q] - g.Pg[((size_t)0 * T.nR + r) * NQ + q];
            bc.sl = bc.lbc * dx;
        }
        if (bc.ut == VERB_BC_CONSTANT_DERIVATIVE) {
            const double dx = g.Pg[((size_t)(T.nP - 1) * T.nR + r) * NQ + q] -
                              g.Pg[((size_t)(T.nP - 2) * T.nR + r) * NQ + q];
            bc.su = bc.ubc * dx;
        }
    } else {
        bc.lt = g.Rl_type; bc.ut = g.Ru_type;
        // R_LBC[iP] * G_conv_R[0], R_UBC[iP] * G_conv_R[R_size-1] (Convection_2D.cpp:135-136)
        bc.lbc = (bc.lt == VERB_BC_PERIODIC) ? 0.0 : g.Rl[(size_t)p * NQ + q] * Gc[((size_t)p * T.nR + 0) * NQ + q];
        bc.ubc = (bc.ut == VERB_BC_PERIODIC) ? 0.0
                                             : g.Ru[(size_t)p * NQ + q] * Gc[((size_t)p * T.nR + (T.nR - 1)) * NQ + q];
        bc.sl = bc.su = 0.0;
        if (bc.lt == VERB_BC_CONSTANT_DERIVATIVE) {
            const double dx = g.Rg[((size_t)p * T.nR + 1) * NQ + q] - g.Rg[((size_t)p * T.nR + 0) * NQ + q];
            bc.sl = bc.lbc * dx;
        }
        if (bc.ut == VERB_BC_CONSTANT_DERIVATIVE) {
            const double dx = g.Rg[((size_t)p * T.nR + T.nR - 1) * NQ + q] -
                              g.Rg[((size_t)p * T.nR + T.nR - 2) * NQ + q];
            bc.su = bc.ubc * dx;
        }
    }
    return bc;
}

// Boundary data of line `line` of sweep dir, as line_bc computed it at the
// start of the call (the data do not change between sub-steps).
__device__ inline LineBC line_bc_lds(const verb_grid& g, const Tile& T, int dir, int line) {
    LineBC bc;
    bc.lt = dir == 0 ? g.Pl_type : g.Rl_type;
    bc.ut = dir == 0 ? g.Pu_type : g.Ru_type;
    const double* e = T.bc + 4 * (dir == 0 ? line : T.nR * T.TQ + line);
    bc.lbc = e[0];
    bc.ubc = e[1];
    bc.sl = e[2];
    bc.su = e[3];
    return bc;
}

__device__ inline double face_value_or_zero(const double* FC, const double* F, int idx, bool has) {
    return has ? FC[idx] * F[idx] : 0.0 * 0.0;
}

// One QUICKEST step (four passes) of sweep `dir` on every line of the tile
// whose plane is active and whose inner-step count exceeds k. X holds the
// current values; the result is left in Y (every point written).
__device__ void quickest_step(const verb_grid& g, const Tile& T, int dir, int k, const double* X, double* Y,
                              double* F, const double* FC, const int* ninner, const int* active,
                              const double* Gc) {
    const int tid = threadIdx.x;
    // (a) unlimited faces
    for (int t = tid; t < T.npt; t += kThreads) {
        int p, r, tq;
        decompose(T, t, p, r, tq);
        if (!active[tq]) continue;
        if (dir == 0 && r == T.nR - 1) continue;
        const LineGeom L = line_of(T, dir, p, r, tq);
        if (k >= ninner[L.line]) continue;
        const LineBC bc = line_bc_lds(g, T, dir, L.line);
        const bool has = (L.ix > 0) || (bc.lt == VERB_BC_PERIODIC);
        if (!has) continue;
        double W[10];
        gather<10>(X, X, L.base, L.stride, L.n, L.ix - 5, bc, W);
        F[t] = unlimited_face(W, FC[t]);
    }
    __syncthreads();
    // (b) unlimited values -> Y
    for (int t = tid; t < T.npt; t += kThreads) {
        int p, r, tq;
        decompose(T, t, p, r, tq);
        bool run = active[tq] && !(dir == 0 && r == T.nR - 1);
        LineGeom L = line_of(T, dir, p, r, tq);
        if (run && k >= ninner[L.line]) run = false;
        if (!run) { Y[t] = X[t]; continue; }
        const LineBC bc = line_bc_lds(g, T, dir, L.line);
        const int n = L.n;
        int i = L.ix;
        double fixed = 0.0;
        bool use_fixed = false;
        if (i == n - 1) {
            if (bc.ut == VERB_BC_CONSTANT_VALUE) { fixed = bc.ubc; use_fixed = true; }
            else if (bc.ut == VERB_BC_CONSTANT_DERIVATIVE) i = n - 2;
            else i = 0;  // lower periodic: U[n-1] = U[0]
        } else if (i == 0) {
            if (bc.lt == VERB_BC_CONSTANT_VALUE) { fixed = bc.lbc; use_fixed = true; }
            else if (bc.lt == VERB_BC_CONSTANT_DERIVATIVE) i = 1;
        }
        if (use_fixed) { Y[t] = fixed; continue; }
        const int ti = L.base + i * L.stride;        // point i
        const int tr = L.base + (i + 1) * L.stride;  // its right face
        const bool hasl = (i > 0) || (bc.lt == VERB_BC_PERIODIC);
        Y[t] = X[ti] - (FC[tr] * F[tr] - face_value_or_zero(FC, F, ti, hasl));
    }
    __syncthreads();
    // (c) limited faces, in place in F
    for (int t = tid; t < T.npt; t += kThreads) {
        int p, r, tq;
        decompose(T, t, p, r, tq);
        if (!active[tq]) continue;
        if (dir == 0 && r == T.nR - 1) continue;
        const LineGeom L = line_of(T, dir, p, r, tq);
        if (k >= ninner[L.line]) continue;
        const LineBC bc = line_bc_lds(g, T, dir, L.line);
        const bool has = (L.ix > 0) || (bc.lt == VERB_BC_PERIODIC);
        if (!has) continue;
        const double cr = FC[t];
        double UW[10];
        gather<10>(X, Y, L.base, L.stride, L.n, L.ix - 5, bc, UW);
        // original points ix-2 .. ix+1: (ou, oc, od) is (ix-2, ix-1, ix)
        // for a non-negative Courant number, else (ix+1, ix, ix-1)
        double o4[4];
        gather<4>(X, X, L.base, L.stride, L.n, L.ix - 2, bc, o4);
        const double ou = (cr >= 0) ? o4[0] : o4[3];
        const double oc = (cr >= 0) ? o4[1] : o4[2];
        const double od = (cr >= 0) ? o4[2] : o4[1];
        F[t] = limited_face(UW, ou, oc, od, F[t], cr);
    }
    __syncthreads();
    // (d) conservative update -> Y
    for (int t = tid; t < T.npt; t += kThreads) {
        int p, r, tq;
        decompose(T, t, p, r, tq);
        bool run = active[tq] && !(dir == 0 && r == T.nR - 1);
        LineGeom L = line_of(T, dir, p, r, tq);
        if (run && k >= ninner[L.line]) run = false;
        if (!run) { Y[t] = X[t]; continue; }
        const LineBC bc = line_bc_lds(g, T, dir, L.line);
        const int n = L.n;
        int i = L.ix;
        if (i == n - 1) {
            if (bc.ut == VERB_BC_CONSTANT_VALUE) { Y[t] = bc.ubc; continue; }
            else if (bc.ut == VERB_BC_CONSTANT_DERIVATIVE) i = n - 2;
            else i = 0;
        } else if (i == 0) {
            if (bc.lt == VERB_BC_CONSTANT_VALUE) { Y[t] = bc.lbc; continue; }
            else if (bc.lt == VERB_BC_CONSTANT_DERIVATIVE) i = 1;
        }
        const int ti = L.base + i * L.stride;
        const int tr = L.base + (i + 1) * L.stride;
        const bool hasl = (i > 0) || (bc.lt == VERB_BC_PERIODIC);
        Y[t] = X[ti] - (FC[tr] * F[tr] - face_value_or_zero(FC, F, ti, hasl));
    }
    __syncthreads();
}

// lost (nullable, [P][R][V][K]): SAVE_PSD_LOST_CONV, each outer sub-step
// adds what the exp losses took (Convection_2D.cpp:150-155); fast:
// FAST_CONVECTION, the smaller of the two outer sub-step counts (:80-84).
__global__ __launch_bounds__(kThreads) void convection_kernel(const verb_grid g, double* __restrict__ psd,
                                                              int* __restrict__ status, int TQ, int ntiles,
                                                              int per_xcd, int single_outer,
                                                              double* __restrict__ lost, int fast) {
    // XCD-aware tile order: consecutive workgroup ids land on different
    // XCDs (round-robin), so give each XCD a contiguous run of tiles to keep
    // neighbouring 64-byte plane segments in one L2.
    const int bid = blockIdx.x;
    const int tile = (bid % 8) * per_xcd + bid / 8;
    if (tile >= ntiles) return;

    extern __shared__ double lds[];
    Tile T;
    T.nP = g.nP; T.nR = g.nR; T.TQ = TQ; T.npt = g.nP * g.nR * TQ;
    T.tq_shift = __builtin_ctz((unsigned)TQ);
    T.mR = g.nR > 1 ? 0xffffffffu / (unsigned)g.nR + 1u : 0u;
    T.NQ = (long)g.nV * g.nK;
    T.q0 = (long)tile * TQ;
    const int npt = T.npt;
    double* bufA = lds;
    double* bufB = lds + npt;
    double* F = lds + 2 * npt;
    double* FCP = lds + 3 * npt;
    double* FCR = lds + 4 * npt;
    // the velocity reduction [2*kThreads] is done before F, FCP and FCR are
    // first written (a barrier lies between), so it shares their space
    double* red = F;
    const int red_extra = 2 * kThreads > 3 * npt ? 2 * kThreads - 3 * npt : 0;
    double* dtq = lds + 5 * npt + red_extra;  // [TQ]
    int* ninP = (int*)(dtq + TQ);           // [nR*TQ]
    int* ninR = ninP + g.nR * TQ;           // [nP*TQ]
    int* nsteps = ninR + g.nP * TQ;         // [TQ]
    int* active = nsteps + TQ;              // [TQ]
    const int nint = (g.nR + g.nP) * TQ + 2 * TQ;  // ints above
    double* bcl = (double*)(ninP + nint + (nint & 1));  // [(nR + nP) * TQ][4], 8-byte aligned
    T.bc = bcl;

    const int tid = threadIdx.x;
    const long NQ = T.NQ;
    const double dt_total = g.dt;

    // load the PSD tile and reduce max|VP|, max|VR| per plane
    double mvp = 0.0, mvr = 0.0;
    for (int t = tid; t < npt; t += kThreads) {
        int p, r, tq;
        decompose(T, t, p, r, tq);
        const long q = T.q0 + tq;
        double v = 0.0;
        if (q < NQ) {
            const size_t gi = ((size_t)p * g.nR + r) * NQ + q;
            v = psd[gi];
            mvp = fmax(mvp, fabs(g.VP[gi]));
            mvr = fmax(mvr, fabs(g.VR[gi]));
        }
        bufA[t] = v;
    }
    red[tid] = mvp;
    red[kThreads + tid] = mvr;
    __syncthreads();
    // (kThreads is a multiple of TQ, so thread t only saw plane t % TQ)
    if (tid < TQ) {
        double a = 0.0, b = 0.0;
        for (int i = tid; i < kThreads; i += TQ) {
            a = fmax(a, red[i]);
            b = fmax(b, red[kThreads + i]);
        }
        const long q = T.q0 + tid;
        double nsP = 1, nsR = 1;
        if (q < NQ) {
            if (g.nP >= 3) {
                const double dP = g.Pg[((size_t)1 * g.nR + 0) * NQ + q] - g.Pg[q];
                const double nd = (double)dt_total * a / dP / 0.25;
                nsP = nd <= 1 ? 1 : ceil(nd);
            }
            if (g.nR >= 3) {
                const double dR = g.Rg[((size_t)0 * g.nR + 1) * NQ + q] - g.Rg[q];
                const double nd = (double)dt_total * b / dR / 0.25;
                nsR = nd <= 1 ? 1 : ceil(nd);
            }
        }
        double ns = fast ? ((nsR < nsP) ? nsR : nsP)   // std::min
                         : ((nsP < nsR) ? nsR : nsP);  // std::max
        if (single_outer) ns = 1;  // a bare QUICKEST6 call: inner sub-steps only
        // A count beyond kMaxSubsteps (or an infinite one) could never finish:
        // report it and leave the plane unchanged; NaN counts run no step, as
        // the reference's loop `it < num_steps` does for NaN.
        int nst = 0;
        if (q < NQ && ns <= kMaxSubsteps) nst = (int)ns;
        else if (q < NQ && ns > kMaxSubsteps) raise_error(status, VERB_ERR_COURANT);
        nsteps[tid] = nst;
        dtq[tid] = dt_total / ns;
    }
    __syncthreads();

    // per-line inner QUICKEST sub-steps (QUICKEST6.cpp:37-52)
    const int nlinesP = g.nR * TQ;  // MLT lines (r, tq); r = nR-1 unused
    const int nlinesR = g.nP * TQ;  // R lines (p, tq)
    for (int l = tid; l < nlinesP + nlinesR; l += kThreads) {
        const bool isP = l < nlinesP;
        const int ll = isP ? l : l - nlinesP;
        const int tq = ll % TQ;
        const int a = ll / TQ;  // r for MLT lines, p for R lines
        const long q = T.q0 + tq;
        int nin = 0;
        if (q < NQ && nsteps[tq] > 0 && ((isP && g.nP >= 3 && a < g.nR - 1) || (!isP && g.nR >= 3))) {
            double umax = 0.0, dx;
            if (isP) {
                for (int p = 0; p < g.nP; p++) umax = fmax(umax, fabs(g.VP[((size_t)p * g.nR + a) * NQ + q]));
                dx = g.Pg[((size_t)1 * g.nR + a) * NQ + q] - g.Pg[((size_t)0 * g.nR + a) * NQ + q];
            } else {
                for (int r = 0; r < g.nR; r++) umax = fmax(umax, fabs(g.VR[((size_t)a * g.nR + r) * NQ + q]));
                dx = g.Rg[((size_t)a * g.nR + 1) * NQ + q] - g.Rg[((size_t)a * g.nR + 0) * NQ + q];
            }
            const double sd = (double)dtq[tq] * umax / dx / 1.0;
            if (sd <= 1) nin = 1;
            else if (sd <= kMaxSubsteps) nin = (int)ceil(sd);
            else if (sd > kMaxSubsteps) raise_error(status, VERB_ERR_COURANT);  // incl. infinite velocity
            if (nin > 0) {
                const double dti = dtq[tq] / nin;
                if (fabs(umax / dx * dti) > 1) raise_error(status, VERB_ERR_COURANT);
            }
        }
        (isP ? ninP : ninR)[ll] = nin;
        // the line's boundary data, read by every sub-step
        if (q < NQ) {
            const LineBC bc = isP ? line_bc(g, T, 0, 0, a, tq, g.G_conv) : line_bc(g, T, 1, a, 0, tq, g.G_conv);
            double* e = bcl + 4 * l;
            e[0] = bc.lbc;
            e[1] = bc.ubc;
            e[2] = bc.sl;
            e[3] = bc.su;
        }
    }
    __syncthreads();

    // face Courant numbers: FC[point ix] = (C[ix-1] + C[ix]) / 2, with the
    // periodic face 0 = (C[n-2] + C[0]) / 2 and C[i] = Ux[i] / dx * dt_inner
    for (int t = tid; t < npt; t += kThreads) {
        int p, r, tq;
        decompose(T, t, p, r, tq);
        const long q = T.q0 + tq;
        double fcp = 0.0, fcr = 0.0;
        if (q < NQ) {
            // MLT direction
            const int lp = r * TQ + tq;
            if (ninP[lp] > 0) {
                const double dx = g.Pg[((size_t)1 * g.nR + r) * NQ + q] - g.Pg[((size_t)0 * g.nR + r) * NQ + q];
                const double dti = dtq[tq] / ninP[lp];
                const int pa = (p == 0) ? g.nP - 2 : p - 1;
                const bool has = (p > 0) || (g.Pl_type == VERB_BC_PERIODIC);
                if (has) {
                    const double ca = g.VP[((size_t)pa * g.nR + r) * NQ + q] / dx * dti;
                    const double cb = g.VP[((size_t)p * g.nR + r) * NQ + q] / dx * dti;
                    fcp = (ca + cb) / 2;
                }
            }
            const int lr = p * TQ + tq;
            if (ninR[lr] > 0) {
                const double dx = g.Rg[((size_t)p * g.nR + 1) * NQ + q] - g.Rg[((size_t)p * g.nR + 0) * NQ + q];
                const double dti = dtq[tq] / ninR[lr];
                const int ra = (r == 0) ? g.nR - 2 : r - 1;
                const bool has = (r > 0) || (g.Rl_type == VERB_BC_PERIODIC);
                if (has) {
                    const double ca = g.VR[((size_t)p * g.nR + ra) * NQ + q] / dx * dti;
                    const double cb = g.VR[((size_t)p * g.nR + r) * NQ + q] / dx * dti;
                    fcr = (ca + cb) / 2;
                }
            }
        }
        FCP[t] = fcp;
        FCR[t] = fcr;
    }
    int maxsteps = 0, maxinP = 0, maxinR = 0;
    for (int i = 0; i < TQ; i++) maxsteps = max(maxsteps, nsteps[i]);
    __syncthreads();
    for (int i = 0; i < nlinesP; i++) maxinP = max(maxinP, ninP[i]);
    for (int i = 0; i < nlinesR; i++) maxinR = max(maxinR, ninR[i]);

    double* X = bufA;
    double* Y = bufB;
    for (int it = 0; it < maxsteps; it++) {
        if (tid < TQ) active[tid] = it < nsteps[tid];
        __syncthreads();
        if (g.nP >= 3) {
            for (int k = 0; k < maxinP; k++) {
                quickest_step(g, T, 0, k, X, Y, F, FCP, ninP, active, g.G_conv);
                double* s = X; X = Y; Y = s;
            }
        }
        if (g.nR >= 3) {
            // PSD_PR_G = PSD_PR.times(G_conv)
            for (int t = tid; t < npt; t += kThreads) {
                int p, r, tq;
                decompose(T, t, p, r, tq);
                if (!active[tq]) continue;
                X[t] = X[t] * g.G_conv[((size_t)p * g.nR + r) * NQ + T.q0 + tq];
            }
            __syncthreads();
            for (int k = 0; k < maxinR; k++) {
                quickest_step(g, T, 1, k, X, Y, F, FCR, ninR, active, g.G_conv);
                double* s = X; X = Y; Y = s;
            }
            // PSD_PR[iP][iR] = PSD_R[iR] / G_conv_R[iR]
            for (int t = tid; t < npt; t += kThreads) {
                int p, r, tq;
                decompose(T, t, p, r, tq);
                if (!active[tq]) continue;
                X[t] = X[t] / g.G_conv[((size_t)p * g.nR + r) * NQ + T.q0 + tq];
            }
            __syncthreads();
        }
        // losses (exp(L*dt)) on every point, sources on interior R
        for (int t = tid; t < npt; t += kThreads) {
            int p, r, tq;
            decompose(T, t, p, r, tq);
            if (!active[tq]) continue;
            const size_t gi = ((size_t)p * g.nR + r) * NQ + T.q0 + tq;
            const double d = dtq[tq];
            const double before = X[t];
            double v = before * exp(g.losses_conv[gi] * d);
            // the point's thread is the same on every sub-step: no other
            // lane touches lost[gi]
            if (lost) lost[gi] = lost[gi] + (before - v);
            if (r >= 1 && r < g.nR - 1) v = v + g.sources[gi] * d;
            X[t] = v;
        }
        __syncthreads();
    }
    // store the tile
    for (int t = tid; t < npt; t += kThreads) {
        int p, r, tq;
        decompose(T, t, p, r, tq);
        const long q = T.q0 + tq;
        if (q < NQ) psd[((size_t)p * g.nR + r) * NQ + q] = X[t];
    }
}


// ---------------------------------------------------------------------------
// Line kernel: one lane per line of one plane, the line in registers.
//
// A workgroup owns TQ consecutive (V,K) planes, kept in LDS ([p][r][tq]) for
// the whole call as above, and 32 lanes per plane. For the MLT sweep lane li
// takes the line r = li (points p = 0..NP-1); for the R sweep, the G_conv
// weighting and the losses and sources, the line p = li (points r =
// 0..NR-1). A lane loads its line from LDS once per sweep, runs every inner
// QUICKEST step on it in registers (the four passes become four unrolled
// loops over the line with compile-time indices, ghosts included), and
// stores it back: two barriers per outer sub-step, one between the sweeps
// and one after the second. Line lengths are template arguments.
// ---------------------------------------------------------------------------
template <int B, class Fn, int... I>
__device__ inline void sfor_seq(Fn&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, B + I>{}), ...);
}
template <int B, int E, class Fn>
__device__ inline void sfor(Fn&& f) {
    if constexpr (E > B) sfor_seq<B>(f, std::make_integer_sequence<int, E - B>{});
}

// Point J (compile time; a ghost outside 0..N-1) of the line xr with the
// ghost rules of QUICKEST6.cpp:89-143 (orig_at).
template <int N, int J>
__device__ inline double xo(const double (&xr)[N], const LineBC& bc) {
    if constexpr (J >= 0 && J < N) {
        return xr[J];
    } else if constexpr (J < 0) {
        if (bc.lt == VERB_BC_PERIODIC) return xr[J + N - 1];
        if (bc.lt == VERB_BC_CONSTANT_VALUE) return bc.lbc;
        double v = xr[0] - bc.sl;
#pragma unroll
        for (int ig = 2; ig <= -J; ig++) v = v - bc.sl;
        return v;
    } else {
        constexpr int m = J - (N - 1);
        if (bc.ut == VERB_BC_PERIODIC) return xr[m];
        if (bc.ut == VERB_BC_CONSTANT_VALUE) return bc.ubc;
        double v = xr[N - 1] + bc.su;
#pragma unroll
        for (int ig = 2; ig <= m; ig++) v = v + bc.su;
        return v;
    }
}

// Point J of the ghost-extended unlimited solution (unl_at): lower ghosts
// from the pre-step values xr, the line and its upper ghosts from yr.
template <int N, int J>
__device__ inline double xu(const double (&xr)[N], const double (&yr)[N], const LineBC& bc) {
    if constexpr (J < 0) {
        return xo<N, J>(xr, bc);
    } else if constexpr (J < N) {
        return yr[J];
    } else {
        constexpr int m = J - (N - 1);
        if (bc.ut == VERB_BC_PERIODIC) return yr[m];
        if (bc.ut == VERB_BC_CONSTANT_VALUE) return bc.ubc;
        double v = yr[N - 1] + bc.su;
#pragma unroll
        for (int ig = 2; ig <= m; ig++) v = v + bc.su;
        return v;
    }
}

// Updated value of point K from face values A (face i left of point i) and
// face Courant numbers fc: QUICKEST6's conservative update with its boundary
// rules (:257-272, 457-477): a constant-value end takes its value, a
// constant-derivative end the update of its neighbour. (The periodic last
// point takes point 0's update; the caller copies it.)
template <int N, int K>
__device__ inline double upd_point(const double (&xr)[N], const double (&fc)[N], const double (&A)[N],
                                   const LineBC& bc) {
    if constexpr (K > 0 && K < N - 1) {
        return xr[K] - (fc[K + 1] * A[K + 1] - fc[K] * A[K]);
    } else if constexpr (K == 0) {
        if (bc.lt == VERB_BC_CONSTANT_VALUE) return bc.lbc;
        if (bc.lt == VERB_BC_CONSTANT_DERIVATIVE) return xr[1] - (fc[2] * A[2] - fc[1] * A[1]);
        return xr[0] - (fc[1] * A[1] - fc[0] * A[0]);
    } else {
        if (bc.ut == VERB_BC_CONSTANT_VALUE) return bc.ubc;
        return xr[N - 2] - (fc[N - 1] * A[N - 1] - fc[N - 2] * A[N - 2]);
    }
}

// One QUICKEST step on the line xr (N points, in registers), the result
// written to out[i * os] (LDS). The four passes of the reference's step
// (unlimited faces, unlimited values, limited faces, final update) run as
// one pipeline over the line, each pass a fixed number of points behind
// the one it reads, so only a window of each stays live: at stage s, the
// Courant number and unlimited value of face s+6, the unlimited value of
// point s+4, the limited value of face s and the final value of point s-2.
// Face i is left of point i (face 0 exists on periodic lines only); the
// Courant number of face i is (C[i-1] + C[i]) / 2 with C[i] = V[i] / dx *
// dt_inner the point's Courant number (QUICKEST6.cpp:148-161).
template <int N>
__device__ inline void line_step(const double (&xr)[N], const double (&C)[N], const LineBC& bc, double* out,
                                 int os) {
    // (addresses formed where they are used, not hoisted and kept live)
    asm volatile("" : "+v"(out), "+s"(os));
    const bool per = bc.lt == VERB_BC_PERIODIC;
    double fc[N], F[N], FL[N], y[N];
    double out0 = 0.0;  // point 0's final value (the periodic last point takes it)
    sfor<-6, N + 2>([&](auto sc) {
        constexpr int s = sc;
        constexpr int kf = s + 6;
        if constexpr (kf >= 0 && kf < N) {
            constexpr int ka = (kf == 0) ? N - 2 : kf - 1;
            fc[kf] = 0.0;
            F[kf] = 0.0;
            if (kf > 0 || per) {
                fc[kf] = (C[ka] + C[kf]) / 2;
                double W[10];
                sfor<0, 10>([&](auto kc) {
                    constexpr int k = kc;
                    W[k] = xo<N, kf - 5 + k>(xr, bc);
                });
                F[kf] = unlimited_face(W, fc[kf]);
            }
        }
        constexpr int ky = s + 4;
        if constexpr (ky >= 0 && ky < N - 1) y[ky] = upd_point<N, ky>(xr, fc, F, bc);
        if constexpr (ky == N - 1) y[ky] = (bc.ut == VERB_BC_PERIODIC) ? y[0] : upd_point<N, ky>(xr, fc, F, bc);
        if constexpr (s >= 0 && s < N) {
            FL[s] = 0.0;
            if (s > 0 || per) {
                const double cr = fc[s];
                double UW[10];
                sfor<0, 10>([&](auto kc) {
                    constexpr int k = kc;
                    UW[k] = xu<N, s - 5 + k>(xr, y, bc);
                });
                // original points s-2 .. s+1: (ou, oc, od) is (s-2, s-1, s)
                // for a non-negative Courant number, else (s+1, s, s-1)
                const double o0 = xo<N, s - 2>(xr, bc), o1 = xo<N, s - 1>(xr, bc);
                const double o2 = xo<N, s>(xr, bc), o3 = xo<N, s + 1>(xr, bc);
                const double ou = (cr >= 0) ? o0 : o3;
                const double oc = (cr >= 0) ? o1 : o2;
                const double od = (cr >= 0) ? o2 : o1;
                FL[s] = limited_face(UW, ou, oc, od, F[s], cr);
            }
        }
        constexpr int ko = s - 2;
        if constexpr (ko == 0) {
            out0 = upd_point<N, 0>(xr, fc, FL, bc);
            out[0] = out0;
        }
        if constexpr (ko > 0 && ko < N - 1) out[ko * os] = upd_point<N, ko>(xr, fc, FL, bc);
        if constexpr (ko == N - 1)
            out[ko * os] = (bc.ut == VERB_BC_PERIODIC) ? out0 : upd_point<N, ko>(xr, fc, FL, bc);
        // (one stage at a time: the windows above are all that stays live)
        __builtin_amdgcn_sched_barrier(0);
    });
}

// Point Courant numbers C[i] = V[i] / dx * dti of a line, V[i] at
// vel[i * vs]. Called once per QUICKEST step: the base pointer goes through
// an empty asm, so that neither the line's addresses nor the Courant numbers
// and the face polynomial coefficients that depend on them alone are
// hoisted out of the sub-step loops and kept live for the whole line.
template <int N>
__device__ inline void line_courant(const double* vel, int vs, double dx, double dti, double (&C)[N]) {
    asm volatile("" : "+v"(vel));
    sfor<0, N>([&](auto ic) {
        constexpr int i = ic;
        C[i] = vel[i * vs] / dx * dti;
    });
}

constexpr int kLineSlots = 32;  // lanes per plane (>= nP, nR)

// Inner sub-step count of a line whose largest |velocity| is umax
// (QUICKEST6.cpp:37-52); 0 when it cannot run (reported).
__device__ inline int inner_steps(double dtq, double umax, double dx, int* status) {
    int nin = 0;
    const double sd = dtq * umax / dx / 1.0;
    if (sd <= 1) nin = 1;
    else if (sd <= kMaxSubsteps) nin = (int)ceil(sd);
    else if (sd > kMaxSubsteps) raise_error(status, VERB_ERR_COURANT);  // incl. infinite velocity
    if (nin > 0) {
        const double dti = dtq / nin;
        if (fabs(umax / dx * dti) > 1) raise_error(status, VERB_ERR_COURANT);
    }
    return nin;
}

template <int NP, int NR>
__global__ __launch_bounds__(512) void convection_lines(const verb_grid g, double* __restrict__ psd,
                                                        int* __restrict__ status, int TQ, int ntiles, int per_xcd,
                                                        double* __restrict__ lost, int fast) {
    static_assert(NP <= kLineSlots && NR <= kLineSlots && NP >= 6 && NR >= 6, "line lengths");
    // XCD-aware tile order, as convection_kernel
    const int bid = blockIdx.x;
    const int tile = (bid % 8) * per_xcd + bid / 8;
    if (tile >= ntiles) return;

    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int tqs = __builtin_ctz((unsigned)TQ);
    const int nthreads = TQ * kLineSlots;
    const int npt = NP * NR * TQ;
    const long NQ = (long)g.nV * g.nK;
    double* X = lds;                         // [p][r][tq]
    double* redP = X + npt;                  // [tq][line] max |VP| of MLT line r
    double* redR = redP + TQ * kLineSlots;   // [tq][line] max |VR| of R line p
    double* dtq_s = redR + TQ * kLineSlots;  // [tq]
    int* ns_s = (int*)(dtq_s + TQ);          // [tq]
    Tile T;
    T.nP = NP; T.nR = NR; T.TQ = TQ; T.npt = npt; T.tq_shift = tqs; T.mR = 0u;
    T.NQ = NQ;
    T.q0 = (long)tile * TQ;
    T.bc = nullptr;

    const int tq = tid & (TQ - 1), li = tid >> tqs;
    const long q = T.q0 + tq;
    const bool qok = q < NQ;
    const size_t pstride = (size_t)NR * NQ;  // global stride between MLT points

    for (int t = tid; t < npt; t += nthreads) {
        const int tq2 = t & (TQ - 1);
        const int pr = t >> tqs;
        const int p = pr / NR, r = pr - p * NR;
        const long qq = T.q0 + tq2;
        X[t] = (qq < NQ) ? psd[((size_t)p * NR + r) * NQ + qq] : 0.0;
    }
    // largest |VP| of MLT line r = li (every r: the plane's maximum) and |VR|
    // of R line p = li
    double umP = 0.0, umR = 0.0;
    if (qok && li < NR) {
        const double* v = g.VP + (size_t)li * NQ + q;
#pragma unroll
        for (int p = 0; p < NP; p++) umP = fmax(umP, fabs(v[p * pstride]));
    }
    if (qok && li < NP) {
        const double* v = g.VR + (size_t)li * pstride + q;
#pragma unroll
        for (int r = 0; r < NR; r++) umR = fmax(umR, fabs(v[(size_t)r * NQ]));
    }
    redP[tq * kLineSlots + li] = umP;
    redR[tq * kLineSlots + li] = umR;
    __syncthreads();
    if (tid < TQ) {
        // outer sub-steps of plane tid (Convection_2D.cpp:57-86)
        double a = 0.0, b = 0.0;
        for (int i = 0; i < kLineSlots; i++) {
            a = fmax(a, redP[tid * kLineSlots + i]);
            b = fmax(b, redR[tid * kLineSlots + i]);
        }
        const long qp = T.q0 + tid;
        double nsP = 1, nsR = 1;
        if (qp < NQ) {
            if (NP >= 3) {
                const double dP = g.Pg[((size_t)1 * NR + 0) * NQ + qp] - g.Pg[qp];
                const double nd = (double)g.dt * a / dP / 0.25;
                nsP = nd <= 1 ? 1 : ceil(nd);
            }
            if (NR >= 3) {
                const double dR = g.Rg[((size_t)0 * NR + 1) * NQ + qp] - g.Rg[qp];
                const double nd = (double)g.dt * b / dR / 0.25;
                nsR = nd <= 1 ? 1 : ceil(nd);
            }
        }
        const double ns = fast ? ((nsR < nsP) ? nsR : nsP)   // std::min
                               : ((nsP < nsR) ? nsR : nsP);  // std::max
        int nst = 0;
        if (qp < NQ && ns <= kMaxSubsteps) nst = (int)ns;
        else if (qp < NQ && ns > kMaxSubsteps) raise_error(status, VERB_ERR_COURANT);
        ns_s[tid] = nst;
        dtq_s[tid] = g.dt / ns;
    }
    __syncthreads();
    int maxsteps = 0;
    for (int i = 0; i < TQ; i++) maxsteps = max(maxsteps, ns_s[i]);
    const int nsq = qok ? ns_s[tq] : 0;
    const double dtq = dtq_s[tq];

    // this lane's MLT line (r = li < NR-1) and R line (p = li): inner
    // sub-steps, boundary data, geometry
    const bool mline = qok && nsq > 0 && NP >= 3 && li < NR - 1;
    const bool rline = qok && nsq > 0 && li < NP;
    int ninP = 0, ninR = 0;
    double dxP = 1.0, dxR = 1.0;
    LineBC bcP{}, bcR{};
    if (mline) {
        dxP = g.Pg[((size_t)1 * NR + li) * NQ + q] - g.Pg[((size_t)0 * NR + li) * NQ + q];
        ninP = inner_steps(dtq, umP, dxP, status);
        bcP = line_bc(g, T, 0, 0, li, tq, g.G_conv);
    }
    if (rline && NR >= 3) {
        dxR = g.Rg[((size_t)li * NR + 1) * NQ + q] - g.Rg[((size_t)li * NR + 0) * NQ + q];
        ninR = inner_steps(dtq, umR, dxR, status);
        bcR = line_bc(g, T, 1, li, 0, tq, g.G_conv);
    }

    for (int it = 0; it < maxsteps; it++) {
        const bool act = it < nsq;
        if (act && mline && ninP > 0) {
            // MLT line r = li: points p at X[(p*NR + li)*TQ + tq]
            double* xl = X + li * TQ + tq;
            const int xs = NR * TQ;
            for (int k = 0; k < ninP; k++) {
                double C[NP];
                line_courant<NP>(g.VP + (size_t)li * NQ + q, (int)pstride, dxP, dtq / ninP, C);
                double xr[NP];
                const double* xp = xl;
                asm volatile("" : "+v"(xp));
                sfor<0, NP>([&](auto ic) {
                    constexpr int i = ic;
                    xr[i] = xp[i * xs];
                });
                line_step<NP>(xr, C, bcP, xl, xs);
            }
        }
        __syncthreads();
        if (act && rline) {
            // R line p = li: points r at X[(li*NR + r)*TQ + tq]
            double* xl = X + li * NR * TQ + tq;
            const size_t gb = (size_t)li * pstride + q;  // global index of (li, r = 0)
            const double* gc = g.G_conv + gb;
            asm volatile("" : "+v"(gc), "+v"(xl));
            if (NR >= 3) {
                // PSD_PR_G = PSD_PR.times(G_conv), R sweep, then / G_conv
#pragma unroll 1
                for (int i = 0; i < NR; i++) xl[i * TQ] = xl[i * TQ] * gc[(size_t)i * NQ];
                if (ninR > 0) {
                    for (int k = 0; k < ninR; k++) {
                        double C[NR];
                        line_courant<NR>(g.VR + gb, (int)NQ, dxR, dtq / ninR, C);
                        double xr[NR];
                        const double* xs = xl;
                        asm volatile("" : "+v"(xs));
                        sfor<0, NR>([&](auto ic) {
                            constexpr int i = ic;
                            xr[i] = xs[i * TQ];
                        });
                        line_step<NR>(xr, C, bcR, xl, TQ);
                    }
                }
#pragma unroll 1
                for (int i = 0; i < NR; i++) xl[i * TQ] = xl[i * TQ] / gc[(size_t)i * NQ];
            }
            // losses (exp(L*dt)) on every point, sources on interior R
#pragma unroll 1
            for (int i = 0; i < NR; i++) {
                const size_t gi = gb + (size_t)i * NQ;
                const double before = xl[i * TQ];
                double v = before * exp(g.losses_conv[gi] * dtq);
                if (lost) lost[gi] = lost[gi] + (before - v);
                if (i >= 1 && i < NR - 1) v = v + g.sources[gi] * dtq;
                xl[i * TQ] = v;
            }
        }
        __syncthreads();
    }
    for (int t = tid; t < npt; t += nthreads) {
        const int tq2 = t & (TQ - 1);
        const int pr = t >> tqs;
        const int p = pr / NR, r = pr - p * NR;
        const long qq = T.q0 + tq2;
        if (qq < NQ) psd[((size_t)p * NR + r) * NQ + qq] = X[t];
    }
}

size_t convection_lines_lds_bytes(int nP, int nR, int TQ) {
    return ((size_t)nP * nR * TQ + 2 * (size_t)TQ * kLineSlots + TQ) * sizeof(double) + TQ * sizeof(int);
}

// Line-kernel instantiations: (nP, nR) of the grids this build runs.
template <int NP, int NR>
int launch_lines_t(const verb_grid& g, double* psd, int* status, hipStream_t stream, int TQ, double* lost, int fast,
                   int max_lds) {
    const size_t lds = convection_lines_lds_bytes(NP, NR, TQ);
    if (lds > (size_t)max_lds) return VERB_ERR_ARG;
    hipFuncSetAttribute((const void*)convection_lines<NP, NR>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const long NQ = (long)g.nV * g.nK;
    const int ntiles = (int)((NQ + TQ - 1) / TQ);
    const int per_xcd = (ntiles + 7) / 8;
    hipLaunchKernelGGL((convection_lines<NP, NR>), dim3(per_xcd * 8), dim3(TQ * kLineSlots), lds, stream, g, psd,
                       status, TQ, ntiles, per_xcd, lost, fast);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

// Returns -1 when no instantiation covers (nP, nR).
int launch_lines(const verb_grid& g, double* psd, int* status, hipStream_t stream, int TQ, double* lost, int fast,
                 int max_lds) {
    if (g.nP == 24 && g.nR == 29) return launch_lines_t<24, 29>(g, psd, status, stream, TQ, lost, fast, max_lds);
    if (g.nP == 24 && g.nR == 10) return launch_lines_t<24, 10>(g, psd, status, stream, TQ, lost, fast, max_lds);
    return -1;
}

}  // namespace

size_t convection_lds_bytes(int nP, int nR, int TQ) {
    const size_t npt = (size_t)nP * nR * TQ;
    const size_t red_extra = 2 * (size_t)kThreads > 3 * npt ? 2 * (size_t)kThreads - 3 * npt : 0;
    size_t b = (5 * npt + red_extra + TQ) * sizeof(double);
    const size_t nint = (size_t)(nR + nP) * TQ + 2 * TQ;
    b += (nint + (nint & 1)) * sizeof(int);
    b += (size_t)(nR + nP) * TQ * 4 * sizeof(double);
    return b;
}

int convection_tq = 0;  // 0 = choose automatically

int launch_convection_mode(const verb_grid& g, double* psd, int* status, hipStream_t stream, int single_outer,
                           int tq, double* lost, int fast, int kernel) {
    if (g.nP < 1 || g.nR < 1) return VERB_ERR_ARG;
    if ((g.nP >= 3 && g.nP < 6) || (g.nR >= 3 && g.nR < 6)) return VERB_ERR_ARG;  // stencil needs 6 points
    if ((g.Pl_type == VERB_BC_PERIODIC) != (g.Pu_type == VERB_BC_PERIODIC)) return VERB_ERR_ARG;
    if ((g.Rl_type == VERB_BC_PERIODIC) != (g.Ru_type == VERB_BC_PERIODIC)) return VERB_ERR_ARG;
    for (int t : {g.Pl_type, g.Pu_type, g.Rl_type, g.Ru_type})
        if (t < 0 || t > 2) return VERB_ERR_BC;
    int dev = 0;
    hipGetDevice(&dev);
    int max_lds = 0;
    hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (kernel == 0 && !single_outer) {
        // the line kernel where it is instantiated for (nP, nR): TQ planes
        // per workgroup, 32 lanes each
        const int tql = tq > 0 ? tq : 8;
        if ((tql & (tql - 1)) || tql * kLineSlots > 512) return VERB_ERR_ARG;  // (launch bounds)
        const int rc = launch_lines(g, psd, status, stream, tql, lost, fast, max_lds);
        if (rc >= 0) return rc;
    }
    int TQ = tq > 0 ? tq : convection_tq;
    if (TQ <= 0) {
        TQ = 4;
        while (TQ > 1 && convection_lds_bytes(g.nP, g.nR, TQ) > (size_t)max_lds) TQ /= 2;
    }
    if (kThreads % TQ) return VERB_ERR_ARG;
    const size_t lds = convection_lds_bytes(g.nP, g.nR, TQ);
    if (lds > (size_t)max_lds) return VERB_ERR_ARG;
    hipFuncSetAttribute((const void*)convection_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const long NQ = (long)g.nV * g.nK;
    const int ntiles = (int)((NQ + TQ - 1) / TQ);
    const int per_xcd = (ntiles + 7) / 8;
    hipLaunchKernelGGL(convection_kernel, dim3(per_xcd * 8), dim3(kThreads), lds, stream, g, psd, status, TQ, ntiles,
                       per_xcd, single_outer, lost, fast);
    return hipGetLastError() == hipSuccess ? VERB_OK : VERB_ERR_HIP;
}

int launch_convection(const verb_grid& g, double* psd, int* status, hipStream_t stream, int tq, double* lost,
                      int fast, int kernel) {
    return launch_convection_mode(g, psd, status, stream, 0, tq, lost, fast, kernel);
}

}  // namespace verb
