// Arguments of the register-window band LU (band_lu.hip).
#pragma once
#include <cstdint>

namespace verb {

// T: the precision of the factorisation (the operator copy, right-hand side,
// eliminated right-hand side and factors are all T).
template <typename T>
struct LuArgsT {
    const T* diag;       // [nodes][9][n] assembled operator, energy-fast rows
    const T* rhs;        // [nodes][n] B0*psd + C0 (energy-fast), or null
    T* y;                // [nodes][n] forward-eliminated right-hand side, or null
    T* ab;               // [nodes][ldab*n] LAPACK band factors
    int64_t* ipiv;       // [nodes][n] 1-based pivots
    int* uext;           // [nodes][n] last nonzero column of each U row, or null
    int n, nx, write_L;
    int* status;
    int nR, include_boundary;  // node skip rule (nR = 0: factor every node)
    const int* nn;             // per-node unknowns (null: n for every node); buffers keep stride n
    const int* order;          // dispatch order of the nodes (null: identity)
    const int* done = nullptr; // nodes with done[node] != 0 are left alone (null: none)
    int* queue = nullptr;      // persistent mode: workgroups take dispatch slots 0..qnodes-1 from this counter (zeroed)
    int qnodes = 0;
    // U in 16-row blocks instead of the LAPACK layout (write_L == 0 only):
    // block b of a node holds rows 16b..16b+15 and the ublk_cols columns
    // from 16b on, column-major (16 consecutive rows per column), at
    // ab + node*ublk_stride + b*16*ublk_cols; entries outside the band are 0
    int ublk_cols = 0;         // 0: LAPACK layout
    size_t ublk_stride = 0;
    // Fused back-substitution (FP64, persistent mode, blocked U, y given):
    // per-node flags set once a node's factors are written, a second
    // dispatch counter (both zeroed by the launch), and x = U^-1 y [nodes][n]
    int* bs_flags = nullptr;
    int* bs_queue = nullptr;
    T* bs_x = nullptr;
    // ... and, when bs_psd is given, x goes straight into the node's PSD
    // slice instead of bs_x: [P][R][V][K] with nK pitch-angle points per
    // energy, the node's unknowns starting at pitch-angle index bs_k0[node]
    // (null: 0), row r = iy*nx + ix at ix*bs_nK + k0 + iy
    T* bs_psd = nullptr;
    const int* bs_k0 = nullptr;
    int bs_nK = 0;
};

// Whether the register-window LU that launch_band_lu picks for kl
// (variant < 0: the smallest that covers it) runs the back-substitution in
// its persistent launch (it needs at least 256 threads per workgroup).
bool band_lu_fuses_backsolve(int variant, int kl);

// Blocked U geometry for n unknowns and bandwidths kl, ku: columns per block
// (kl+ku+16 rounded up to 16) and the per-node stride. Returns false when the
// blocked factors would not fit in the LAPACK band storage of the node.
inline bool ublk_geometry(int n, int kl, int ku, int* cols, size_t* stride) {
    const int c = (kl + ku + 16 + 15) / 16 * 16;
    const size_t st = (size_t)((n + 15) / 16) * 16 * c;
    if (st > (size_t)(2 * kl + ku + 1) * n) return false;
    *cols = c;
    *stride = st;
    return true;
}

using LuArgs = LuArgsT<double>;

}  // namespace verb
