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
};

using LuArgs = LuArgsT<double>;

}  // namespace verb
