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
                           int tq, double* lost, int fast) {
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
                      int fast) {
    return launch_convection_mode(g, psd, status, stream, 0, tq, lost, fast);
}

}  // namespace verb
