// Arguments of the matrix-free GMRES local solve (gmres.hip).
#pragma once

namespace verb {

constexpr int kGmMaxM = 24;  // largest Krylov dimension per restart cycle

struct GmArgs {
    const double* diag;  // [nodes][9][nfull] assembled operator (energy-fast rows, stride n)
    const double* rhs;   // [nodes][nfull] b = B0*psd + C0
    double* x;           // [nodes][nfull] in: initial iterate, out: solution
    float* d32;          // [nodes][9][nfull] FP32 copy of the operator
    float* pf;           // [nodes][3][nfull] pitch-angle block factors (sub, multipliers, inverse pivots)
    float* V;            // [nodes][kGmMaxM+1][nfull] Krylov basis
    float* z;            // [nodes][2][nfull] preconditioned vector and combination scratch
    int* conv;           // [nodes] 1 when the FP64 residual met tol (or the node is skipped)
    int* iters;          // [nodes] inner iterations used
    double* resid;       // [nodes] max|b - A x| / max|b| at the last FP64 check (nullable)
    int nx, ny, nfull;
    int m, max_cycles;
    int max_iters;       // inner iterations per node at most, over all cycles (0: m * max_cycles)
    double tol;          // max|b - A x| <= tol * max|b|
    float inner_tol;     // residual reduction that ends an FP32 cycle early
    int nR, include_boundary;
};

int launch_gmres(const GmArgs& a, int nodes, hipStream_t stream);

}  // namespace verb
