/*
 * verb4d_hip.h - C-ABI of the MI355X-native VERB-4D operator-split step.
 *
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * Every entry returns an int status:
 *     0   success
 *    >0   LAPACK-style info: the 1-based column of the first exactly-zero
 *         pivot of a band factorisation (dgbtrf's info > 0)
 *    <0   a verb_status_t error below
 * The C++ drop-in translation units map status != 0 to the reference's
 * behaviour (Logger::error + exit(EXIT_FAILURE)); this layer never exits.
 *
 * Array conventions follow the reference's Matrix containers (row-major,
 * last index fastest): 2-D [q1][q2], 4-D [P][R][V][K] (src/Matrix.cpp:161,
 * 4266). Boundary types use the values of the reference's
 * BoundaryConditionType (src/BoundaryConditionType.hpp:53-88).
 *
 * Threading: a verb_ctx owns one HIP stream and its scratch; calls on one ctx
 * are serialised by an internal mutex, so the drop-in shims may be called
 * from OpenMP workers the way the reference time loop does
 * (src/VERB4D_Solver.cpp:885,1009,1101). Use one ctx per thread for overlap.
 */
#ifndef VERB4D_HIP_H
#define VERB4D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    VERB_OK = 0,
    VERB_ERR_ARG = -1,        /* bad argument / unsupported size */
    VERB_ERR_BC = -2,         /* unknown boundary type (MatrixSolver.cpp:57-58) */
    VERB_ERR_COURANT = -3,    /* Courant number > 1 (QUICKEST6.cpp:48-52) */
    VERB_ERR_HIP = -4,        /* HIP runtime failure */
    VERB_ERR_NOMEM = -5,      /* device allocation failed */
    VERB_ERR_NODEVICE = -6,   /* no HIP device visible */
    VERB_ERR_UNSTABLE = -7,   /* DVV*DKK - DVK^2 < 0 somewhere (VERB4D_Solver.cpp:747-753) */
    VERB_ERR_DOMAIN = -8      /* interpolation input the reference rejects: a non-finite point to
                                 evaluate or NaN setting the data's direction (the interpolator
                                 throws, MonotCubicInterpolator.cpp:178-180, 464), or a NaN grid
                                 value (no place in std::map's order) */,
    VERB_ERR_IO = -9,         /* file missing, unreadable, short or not writable */
    VERB_ERR_FORMAT = -10,    /* file content the reference's readers reject or misread: a token
                                 that is not a number, more data than the sizes ("Check grid
                                 size", Matrix.cpp:4957-4961), a grid mismatch in a checked read
                                 (Matrix.cpp:5390-5393) */
    VERB_ERR_SINGULAR = -11   /* a dense solve met a zero pivot (trans_solve's dgesv info > 0,
                                 MatrixOperations.cpp:779-786, exits) */
} verb_status_t;

/* Values of the reference's BoundaryConditionType enum class. */
enum { VERB_BC_PERIODIC = 0, VERB_BC_CONSTANT_VALUE = 1, VERB_BC_CONSTANT_DERIVATIVE = 2 };

typedef struct verb_ctx verb_ctx;

/* Context: binds the calling thread's current HIP device, creates a stream. */
int verb_ctx_create(verb_ctx** out);
int verb_ctx_destroy(verb_ctx* ctx);
/* The ctx's stream (a hipStream_t) for callers that enqueue their own work. */
void* verb_ctx_stream(verb_ctx* ctx);
int verb_ctx_synchronize(verb_ctx* ctx);
/* Human-readable text of the last error raised on this ctx. */
const char* verb_ctx_last_error(verb_ctx* ctx);
/* Bytes of device scratch the ctx currently holds. */
size_t verb_ctx_scratch_bytes(verb_ctx* ctx);

/* Kernel-phase timing with HIP events recorded on the ctx stream around each
 * batched launch (off by default). Phases: VERB_PHASE_*. get_timings
 * synchronises the stream, writes the summed milliseconds and launch counts
 * of every phase (arrays of VERB_NUM_PHASES) and clears the totals.
 * No reference counterpart: the reference only times the whole run
 * (src/VERB4D_Solver.cpp:429-431, 1311-1312). */
enum { VERB_PHASE_CONVECTION = 0, VERB_PHASE_RADIAL = 1, VERB_PHASE_LOCAL_ASSEMBLE = 2, VERB_PHASE_LOCAL_FACTOR = 3,
       VERB_PHASE_LOCAL_SOLVE = 4, VERB_PHASE_LOCAL_KRYLOV = 5, VERB_PHASE_LOCAL_WARMSTART = 6,
       VERB_NUM_PHASES = 7 };
int verb_ctx_set_timing(verb_ctx* ctx, int enable);
int verb_ctx_get_timings(verb_ctx* ctx, double* ms, long* counts);

/* ------------------------------------------------------------------------ */
/* Per-call drop-in entries (host pointers, synchronous).                    */
/* ------------------------------------------------------------------------ */

/* Replaces Diffusion_2D (src/Diffusion_2D.h:86-111, both the plain and the
 * LU_CACHING variants) including its AddBoundaries_2D /
 * SecondDerivativeApproximation_2D assembly (src/Diffusion_2D.cpp:43-130)
 * and the Lapack() banded solve (src/MatrixSolver.cpp:223-340).
 * 2-D arrays are [nx][ny]; xl/xu have ny entries, yl/yu nx entries.
 * lu_cache == NULL selects the plain variant. Otherwise lu_cache holds
 * (2kl+ku+1)*nx*ny doubles and ipiv_cache nx*ny int64 in LAPACK band
 * layout (kl = ku = nx+1), factored when recompute_lu != 0 and reused
 * otherwise, as the reference's caller-owned caches are. */
int verb_diffusion_2d(verb_ctx* ctx, double* psd, const double* x, const double* y, int nx, int ny,
                      const double* xl, const double* xu, const double* yl, const double* yu,
                      int xl_type, int xu_type, int yl_type, int yu_type,
                      const double* Dxx, const double* Dyy, const double* Dxy, const double* Dyx,
                      const double* G, const double* sources, const double* losses, double dt, double sub_dt,
                      double* lu_cache, int64_t* ipiv_cache, int recompute_lu);

/* verb_diffusion_2d without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results as verb_diffusion_2d;
 * every device buffer (staged inputs, operator, band factors, pivots,
 * status words) comes from one static 32 MiB device array per GPU, held
 * until the process ends. A node needing more (about nx*ny*(3nx+29)*8
 * bytes; the 101x91 production node takes 23.3 MiB) is rejected with
 * VERB_ERR_ARG. Host threads of one process are serialised per GPU; calls
 * that overlap on one GPU otherwise race on the array. */
int verb_diffusion_2d_noalloc(verb_ctx* ctx, double* psd, const double* x, const double* y, int nx, int ny,
                              const double* xl, const double* xu, const double* yl, const double* yu,
                              int xl_type, int xu_type, int yl_type, int yu_type,
                              const double* Dxx, const double* Dyy, const double* Dxy, const double* Dyx,
                              const double* G, const double* sources, const double* losses, double dt,
                              double sub_dt, double* lu_cache, int64_t* ipiv_cache, int recompute_lu);

/* Replaces Diffusion_1D (src/Diffusion_1D.h:69-77): assembly of
 * getCalculationMatrix1DDiffusion (src/Diffusion_1D.cpp:91-184), RHS
 * C + B*f and tridag (src/MatrixSolver.cpp:689-720). Lines of 3..64 points
 * (the Thomas recurrences are held in registers); longer lines return
 * VERB_ERR_ARG, which the reference's tridag would accept. */
int verb_diffusion_1d(verb_ctx* ctx, double* psd, const double* x, int n, double lbc, double ubc, int l_type,
                      int u_type, const double* Dxx, const double* G, const double* sources, const double* losses,
                      double dt);

/* Replaces Convection_2D (src/Convection_2D.h:98-108, SAVE_PSD_LOST_CONV
 * off): 2-D arrays [nP][nR]; Pl/Pu have nR entries, Rl/Ru nP entries. */
int verb_convection_2d(verb_ctx* ctx, double* psd, const double* P, const double* R, int nP, int nR,
                       const double* Pl, const double* Pu, const double* Rl, const double* Ru,
                       int Pl_type, int Pu_type, int Rl_type, int Ru_type,
                       const double* VP, const double* VR, const double* sources, const double* losses,
                       const double* G_conv, double dt);

/* The reference's two compile-time convection options, per call:
 *   psd_lost (NULL: off): SAVE_PSD_LOST_CONV (src/Convection_2D.h:84-96,
 *     src/Convection_2D.cpp:150-155). Every outer sub-step adds
 *     (PSD - PSD*exp(L*dt_sub)) to it, the PSD the exp losses took; it is
 *     read and written, the caller zeroes it once per time step as the
 *     reference's loop does (src/VERB4D_Solver.cpp:626-628). Same layout as
 *     psd: [nP][nR] for the per-call entry, [P][R][V][K] for the batched one.
 *   fast_convection != 0: FAST_CONVECTION (src/Convection_2D.cpp:80-84), the
 *     smaller of the MLT and R outer sub-step counts instead of the larger. */
typedef struct {
    double* psd_lost;
    int fast_convection;
} verb_convection_options;

/* Convection_2D with the options above (opt == NULL: verb_convection_2d);
 * with psd_lost it is the SAVE_PSD_LOST_CONV signature
 * Convection_2D(PSD_PR, PSD_lost_PR, ...) (src/Convection_2D.h:84-96). */
int verb_convection_2d_ex(verb_ctx* ctx, double* psd, const double* P, const double* R, int nP, int nR,
                          const double* Pl, const double* Pu, const double* Rl, const double* Ru,
                          int Pl_type, int Pu_type, int Rl_type, int Ru_type,
                          const double* VP, const double* VR, const double* sources, const double* losses,
                          const double* G_conv, double dt, const verb_convection_options* opt);

/* Replaces Convection_1D_ULTIMATE_QUICKEST6
 * (src/Convection_1D_ULTIMATE_QUICKEST6.h:71-79). */
int verb_convection_1d_quickest6(verb_ctx* ctx, double* psd, const double* x, int n, double lbc, double ubc,
                                 int l_type, int u_type, const double* Ux, double dt);

/* Replace the LAPACK calls the reference binds (src/MatrixSolver.h:243-248):
 * dgbtrf_ (band LU with partial pivoting, ipiv 1-based, return = info) and
 * dgbtrs_('N', nrhs = 1). Column-major band storage, ldab >= 2kl+ku+1. */
int verb_dgbtrf(verb_ctx* ctx, int64_t m, int64_t n, int64_t kl, int64_t ku, double* ab, int64_t ldab,
                int64_t* ipiv);
int verb_dgbtrs(verb_ctx* ctx, int64_t n, int64_t kl, int64_t ku, const double* ab, int64_t ldab,
                const int64_t* ipiv, double* b);

/* Replaces tridag (src/MatrixSolver.cpp:689-720) for `lines` independent
 * systems stored line-major (line l occupies [l*n, (l+1)*n)); n <= 64. */
int verb_tridag_batched(verb_ctx* ctx, const double* a, const double* b, const double* c, const double* r,
                        double* u, int n, int lines);

/* ------------------------------------------------------------------------ */
/* Batched full-grid entries (device pointers, asynchronous on ctx stream).  */
/* ------------------------------------------------------------------------ */

/* All fields the reference time loop hands to the three operators, as
 * device-resident arrays. 4-D fields are [P][R][V][K]. The boundary arrays
 * are the reference's Matrix3D boundary fields: Pl/Pu [R][V][K],
 * Rl/Ru [P][V][K] (convection, src/VERB4D_Solver.cpp:909-910),
 * Ll/Lu [P][V][K] (radial, :1041), Vl/Vu [P][R][K] and Kl/Ku [P][R][V]
 * (local, :1147-1148). */
typedef struct verb_grid {
    int nP, nR, nV, nK;
    const double *Pg, *Rg, *Vg, *Kg, *Lg;
    const double *DLL, *DVV, *DKK, *DVK, *VP, *VR;
    const double *G_local, *G_radial, *G_conv, *sources, *losses_local, *losses_radial, *losses_conv;
    const double *Pl, *Pu, *Rl, *Ru, *Ll, *Lu, *Vl, *Vu, *Kl, *Ku;
    int Pl_type, Pu_type, Rl_type, Ru_type, Ll_type, Lu_type, Vl_type, Vu_type, Kl_type, Ku_type;
    double dt, sub_dt;
    /* 0/1 multipliers routing the sources to radial or local diffusion
     * (src/VERB4D_Solver.cpp:420-426, 1037, 1151) */
    double radial_sources, local_sources;
    /* include_boundary == 0 skips local diffusion at iR = 0 and nR-1
     * (src/VERB4D_Solver.cpp:1117-1120) */
    int include_boundary;
} verb_grid;

/* Convection_2D over every (V,K) plane (src/VERB4D_Solver.cpp:860-958). */
int verb_convection_batched(verb_ctx* ctx, const verb_grid* g, double* psd);
/* The same with verb_convection_options (psd_lost a device [P][R][V][K]
 * array, PSD_lost_conv of src/VERB4D_Solver.cpp:335-337, 921-948). */
int verb_convection_batched_ex(verb_ctx* ctx, const verb_grid* g, double* psd, const verb_convection_options* opt);
/* Diffusion_1D over every (P,V,K) line (src/VERB4D_Solver.cpp:999-1055);
 * nR <= 64. */
int verb_radial_diffusion_batched(verb_ctx* ctx, const verb_grid* g, double* psd);
/* Diffusion_2D over every (P,R) node (src/VERB4D_Solver.cpp:1100-1210).
 * lu_cache/ipiv_cache: NULL for the plain path (factor every call), or
 * device buffers of verb_lu_cache_doubles / verb_lu_cache_ints elements
 * (LAPACK band layout per node) factored when recompute_lu != 0. */
int verb_local_diffusion_batched(verb_ctx* ctx, const verb_grid* g, double* psd, double* lu_cache,
                                 int64_t* ipiv_cache, int recompute_lu);
size_t verb_lu_cache_doubles(const verb_grid* g);

/* The operator Diffusion_2D assembles for every (P,R) node
 * (src/Diffusion_2D.cpp:43-130: AddBoundaries_2D on edge points and the
 * twelve SecondDerivativeApproximation_2D terms inside,
 * src/MatrixSolver.cpp:118-494) for one sub-step of dt / max(1,
 * int(dt/sub_dt)), as the batched local diffusion factors it. Device
 * outputs, energy-fast unknowns i = iK*nV + iV (src/Matrix.cpp:6479-6481):
 *   diag [nodes][9][n]: the CalculationMatrix diagonals in key order,
 *     offsets -nV-1, -nV, -nV+1, -1, 0, 1, nV-1, nV, nV+1, with
 *     diag[s][i] = A(i, i + offset_s) (src/Matrix.cpp:6423-6474);
 *   b0, c0 [nodes][n]: the diagonals of B and C.
 * Nodes skipped by include_boundary are not written. */
int verb_local_assembly(verb_ctx* ctx, const verb_grid* g, const double* psd, double* diag, double* b0,
                        double* c0);
size_t verb_lu_cache_ints(const verb_grid* g);

/* Local diffusion with an uneven active pitch-angle extent per node, as the
 * reference expresses it by calling Diffusion_2D with a trimmed y_size and
 * sliced arrays (src/Diffusion_2D.h:100-110). k0[node] (host array of
 * nP*nR ints, node = iP*nR + iR) is the first active K column: the node
 * solves the system of the slice [V][k0:K], whose lower pitch-angle row
 * takes the K-lower boundary condition; columns below k0 are left
 * untouched. Nodes are dispatched largest first. No reference counterpart
 * as a batched entry. */
int verb_local_diffusion_uneven(verb_ctx* ctx, const verb_grid* g, double* psd, const int* k0, double* lu_cache,
                                int64_t* ipiv_cache, int recompute_lu);
int verb_step_uneven(verb_ctx* ctx, const verb_grid* g, double* psd, const int* k0, int run_convection,
                     int run_radial, int run_local);

/* One full operator-split step: convection, radial, local
 * (src/VERB4D_Solver.cpp:851-1216), each enabled by its run_* flag. */
int verb_step(verb_ctx* ctx, const verb_grid* g, double* psd, int run_convection, int run_radial, int run_local,
              double* lu_cache, int64_t* ipiv_cache, int recompute_lu);

/* The step as the reference's time loop runs it around the three operators
 * (src/VERB4D_Solver.cpp:539-1282), on the device-resident grid:
 *   positive_psd != 0: PSD.max_of(1e-21) before the step and after each
 *     operator (:621, 980, 1072, 1218-1221), max_of being
 *     (v < 1e-21) ? 1e-21 : v (src/Matrix.cpp:5948-5953);
 *   vl_from_convection (device [P][R][K] buffer, or NULL): when convection
 *     runs and Vl_type is ConstantValue, it takes PSD[:, :, 0, :] after
 *     convection and is the V-lower boundary of this step's local diffusion
 *     (:965-966, Vl_BC = PSD.ySlice(0)); the caller keeps it as g->Vl for
 *     later steps as the reference keeps Vl_BC until its next update;
 *   negative_points (host pointer, or NULL): the number of PSD values below
 *     0 after the step (:1223-1238), 0 with positive_psd (:1218-1221);
 *     reading it synchronises the stream;
 *   psd_lost, fast_convection: the convection options of
 *     verb_convection_options.
 * opt == NULL is verb_step. */
typedef struct {
    int positive_psd;
    double* vl_from_convection;
    long long* negative_points;
    double* psd_lost;     /* convection's SAVE_PSD_LOST_CONV accumulator (NULL: off) */
    int fast_convection;  /* FAST_CONVECTION outer sub-steps */
    double* vu_from_convection; /* as vl_from_convection for the V-upper boundary:
                                   PSD[:, :, nV-1, :] when Vu_type is ConstantValue
                                   (src/VERB4D_Solver.cpp:970-975) */
} verb_step_options;
int verb_step_ex(verb_ctx* ctx, const verb_grid* g, double* psd, int run_convection, int run_radial,
                 int run_local, double* lu_cache, int64_t* ipiv_cache, int recompute_lu,
                 const verb_step_options* opt);

/* The scaling and limiting of UpdatableMatrix::update
 * (src/UpdatableMatrix.cpp:601-827), on device arrays of n points: dst =
 * original * scale (the multiplication only when scale != 1), then every
 * point whose coordinate grid[a][i] lies outside [from[a], to[a]] on any
 * given axis a (grid[a] != NULL) is set to 0 (MatrixLimit, :196-300). The
 * host keeps the time queues that decide scale and ranges (and which
 * original array a .lst update loaded). dst may equal original only when
 * scale == 1 and no axis is given. */
typedef struct {
    double scale;
    const double* grid[4];
    double from[4], to[4];
} verb_field_update_params;
int verb_field_update(verb_ctx* ctx, double* dst, const double* original, size_t n,
                      const verb_field_update_params* params);

/* The stability check of the time loop (src/VERB4D_Solver.cpp:747-753):
 * min over the grid of DVV*DKK - DVK*DVK (NaN entries passed over, as
 * Matrix4D::min does, src/Matrix.cpp:5881-5888) into *min_value (host);
 * returns VERB_ERR_UNSTABLE when it is negative. Synchronises the stream. */
int verb_check_stability(verb_ctx* ctx, const verb_grid* g, double* min_value);

/* Adiabatic remapping (src/VERB4D_Solver.cpp:557-619): every (P,V,K) line of
 * psd is interpolated from the old L* values of its points to the new ones
 * with Cubic1D (src/Interpolation.cpp:25-51: the monotone cubic Hermite
 * interpolator of src/MonotCubicInterpolator.cpp over the points (0, 0) and
 * (lstar_old[i], psd[i]), the extra point coming from Cubic1D's vectors
 * holding n zeros before the grid). psd, lstar_old and lstar_new are device
 * arrays [P][R][V][K]; nR <= 105. Asynchronous on the context's stream; a
 * VERB_ERR_DOMAIN line is left unchanged and reported by verb_check_status. */
int verb_remap_batched(verb_ctx* ctx, double* psd, const double* lstar_old, const double* lstar_new, int nP,
                       int nR, int nV, int nK);
/* Cubic1D on one line of host arrays (src/Interpolation.h:25): values (n)
 * replaced by the interpolant at x_new (n); synchronous. */
int verb_cubic1d(verb_ctx* ctx, int n, const double* x_old, double* values, const double* x_new);

/* Pulls the device-side status word of the last batched launches (Courant
 * violation, zero pivot) and synchronises the stream. */
int verb_check_status(verb_ctx* ctx);

/* ------------------------------------------------------------------------
 * Local diffusion by matrix-free GMRES. No counterpart in the reference
 * (src/Diffusion_2D.cpp:144-148 always factors with dgbtrf_); it solves the
 * same system A x = B0*psd + C0 per (P,R) node for a single sub-step
 * (dt <= sub_dt). FP32 restarted GMRES cycles, right-preconditioned by the
 * exact pitch-angle tridiagonal blocks, correct an FP64 iterate; a node is
 * converged when the FP64 residual satisfies max|b - A x| <= tol * max|b|.
 * Nodes that are not converged after max_cycles cycles are factored by the
 * band LU (the verb_local_diffusion_batched path) instead. The initial
 * iterate is the incoming PSD. stats (nullable) synchronises the stream.
 * ------------------------------------------------------------------------ */
typedef struct {
    double tol;         /* FP64 residual tolerance, relative to max|b| */
    int restart;        /* Krylov vectors per FP32 cycle, 1..24 */
    int max_cycles;     /* cycles before a node falls back to the band LU */
    double inner_tol;   /* residual reduction that ends an FP32 cycle early */
    int max_iterations; /* inner iterations per node over all cycles before
                           it falls back (0: restart * max_cycles); the
                           warm-started path uses 8 */
} verb_gmres_params;

typedef struct {
    int nodes;            /* systems solved (nodes skipped by include_boundary excluded) */
    int converged;        /* met tol by GMRES */
    int fallback;         /* solved by the band LU instead */
    long long iterations; /* inner iterations, summed over the solved nodes */
    int max_iterations;   /* inner iterations of the slowest node */
    double max_residual;  /* largest max|b - A x| / max|b| at a node's last FP64 check */
    double mean_log10_residual; /* mean over nodes of log10 of that residual (before any fallback) */
} verb_gmres_stats;

int verb_local_diffusion_gmres(verb_ctx* ctx, const verb_grid* g, double* psd, const verb_gmres_params* params,
                               verb_gmres_stats* stats);
/* The same solve from the initial iterate x0 (device, [P][R][V][K] like
 * psd) instead of the incoming PSD; x0 is only read. */
int verb_local_diffusion_gmres_x0(verb_ctx* ctx, const verb_grid* g, double* psd, const double* x0,
                                  const verb_gmres_params* params, verb_gmres_stats* stats);
/* convection + radial + verb_local_diffusion_gmres */
int verb_step_gmres(verb_ctx* ctx, const verb_grid* g, double* psd, const verb_gmres_params* params,
                    verb_gmres_stats* stats);

/* ------------------------------------------------------------------------
 * Launch shapes and the captured step. No counterpart in the reference (it
 * has no device launches); they change speed, not results.
 *   verb_set_tuning / verb_get_tuning: the ctx's launch shapes.
 *   verb_autotune: times every convection tile size and every register-window
 *     LU shape that covers nV (each with and without the persistent node
 *     queue) on this grid and PSD, keeps the fastest in the ctx and reports
 *     them; psd is restored before it returns.
 *   verb_step_graph: the plain step (convection, radial, local with LU caching
 *     off) of verb_step, or of verb_step_uneven when k0 is given (host array,
 *     one first active pitch-angle column per node). The first call for a
 *     grid/psd/tuning/k0 runs it eagerly and records it as a HIP graph; later
 *     calls replay the graph.
 * ------------------------------------------------------------------------ */
typedef struct {
    int convection_tq;    /* (V,K) planes per convection workgroup, 0 = automatic */
    int lu_rb, lu_cb;     /* register-window LU shape in 16-row / 16-column blocks, 0 = automatic */
    int lu_persistent;    /* 1: one LU workgroup per CU takes nodes from a queue */
    double convection_ms; /* verb_autotune: time of the chosen convection launch */
    double local_ms;      /* verb_autotune: time of the chosen local-diffusion launches */
    int convection_kernel; /* 0: a lane per (MLT or R) line, the line in registers, where the
                              build has that kernel for (nP, nR) (24x29, 24x10), else 1;
                              1: a lane per grid point over LDS tiles */
} verb_tuning;

int verb_set_tuning(verb_ctx* ctx, const verb_tuning* t);
int verb_get_tuning(verb_ctx* ctx, verb_tuning* t);
int verb_autotune(verb_ctx* ctx, const verb_grid* g, double* psd, verb_tuning* out);
int verb_step_graph(verb_ctx* ctx, const verb_grid* g, double* psd, const int* k0);

/* ------------------------------------------------------------------------
 * Local diffusion with an FP32 band LU and FP64 iterative refinement. No
 * counterpart in the reference (which factors in FP64 with dgbtrf_); it
 * solves the same system A x = B0*psd + C0 per (P,R) node for a single
 * sub-step: the operator is factored in FP32 (partial pivoting, register
 * window up to kl = nV+1 = 152), then from x = 0 each sweep forms the FP64
 * residual r = b - A x, stops the node when max|r| <= tol * max|b|, solves
 * L U d = P r in FP32 and adds d to x in FP64. stats (nullable)
 * synchronises the stream.
 * ------------------------------------------------------------------------ */
typedef struct {
    double tol;      /* FP64 residual tolerance, relative to max|b| */
    int max_sweeps;  /* FP32 solves per node at most */
} verb_mixed_params;

typedef struct {
    int nodes;           /* systems solved (nodes skipped by include_boundary excluded) */
    int converged;       /* met tol within max_sweeps */
    int max_sweeps_used; /* FP32 solves of the slowest node */
    double mean_sweeps;  /* FP32 solves per node */
    double max_residual; /* largest final max|r| / max|b| */
} verb_mixed_stats;

int verb_local_diffusion_mixed(verb_ctx* ctx, const verb_grid* g, double* psd, const verb_mixed_params* params,
                               verb_mixed_stats* stats);
/* convection + radial + verb_local_diffusion_mixed */
int verb_step_mixed(verb_ctx* ctx, const verb_grid* g, double* psd, const verb_mixed_params* params,
                    verb_mixed_stats* stats);

/* ------------------------------------------------------------------------
 * The learned warm start. No counterpart in the reference. A 4-level U-Net
 * (3x3 convolutions 16->32->32 | 2x2 average pool | 64 | pool | 96 | pool |
 * 128 | nearest 2x upsampling with skip concatenation: 96, 64, 32 | 1x1 head
 * to one channel; ReLU after every 3x3 convolution; 512,065 parameters, of
 * which 508,897 are live: the input layer's weights for the zero channels
 * 5..15 stay 0) maps five channels of a (P,R) node on its E x alpha grid
 * padded to multiples of 8 (Hp = round8(nV), Wp = round8(nK)):
 *   0: (log10 Daa + 5)/3   1: (log10 DEE + 5)/3     2: DaE/sqrt(Daa DEE)
 *   3: log10(PSD/s)/15 + 1  4: dt/600 s
 * with Daa = DKK, DEE = DVV/E^2, DaE/sqrt(Daa DEE) = DVK/sqrt(DVV DKK) and s
 * the node's max|PSD|, to delta = (PSD_out - PSD_in)/s; x0 = PSD_in + s delta
 * starts the GMRES local solve. BF16 operands on MFMA with FP32
 * accumulation; BF16 activations, FP32 gradients and master parameters.
 * Features: feat [count][Hp][Wp][16] BF16 (uint16 bit patterns), pin
 * [count][Hp][Wp] FP32 = PSD/s (0 in the padding), scale [count] FP64.
 * Parameters (FP32 vector of verb_unet_param_count()): per 3x3 convolution
 * W [9][cin][cout] (tap = 3(dh+1) + dw+1) then b [cout]; then the head
 * w [32], b [1]. Loss of a batch: sum over pixels of
 * (pin + delta - pout)^2 / (batch Hp Wp), pout = PSD_out/s.
 * All device pointers; work on the ctx's stream.
 * ------------------------------------------------------------------------ */
typedef struct verb_unet verb_unet;
size_t verb_unet_param_count(void);
/* A network for batches of up to max_batch nodes of height x width
 * (multiples of 8); train != 0 also allocates gradients and Adam state. */
int verb_unet_create(verb_ctx* ctx, int max_batch, int height, int width, int train, verb_unet** out);
int verb_unet_destroy(verb_unet* net);
/* host FP32 parameter vectors */
int verb_unet_set_params(verb_unet* net, const float* params);
int verb_unet_get_params(verb_unet* net, float* params);
/* features of nodes [node0, node0+count) of a device grid and PSD */
int verb_unet_features(verb_unet* net, const verb_grid* g, const double* psd, int node0, int count, uint16_t* feat,
                       float* pin, double* scale);
/* out = psd / scale on the padded grid (training targets pout) */
int verb_unet_normalize(verb_unet* net, const verb_grid* g, const double* psd, int node0, int count,
                        const double* scale, float* out);
/* delta [batch][height][width] */
int verb_unet_forward(verb_unet* net, const uint16_t* feat, int batch, int height, int width, float* delta);
/* the loss (host, nullable: then no synchronisation) and the gradient of
 * every parameter (device FP32 vector, nullable) */
int verb_unet_gradients(verb_unet* net, const uint16_t* feat, const float* pin, const float* pout, int batch,
                        int height, int width, float* grad, float* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) at learning rate lr */
int verb_unet_train_step(verb_unet* net, const uint16_t* feat, const float* pin, const float* pout, int batch,
                         int height, int width, float lr, float* loss);

/* Local diffusion from the network's warm start: x0 = PSD + s delta per node
 * (net created on the same ctx, max_batch nodes per forward launch), then
 * verb_local_diffusion_gmres_x0: the FP64 residual judges each node and the
 * band LU solves the nodes that miss params->tol within the iteration
 * budget (params->max_iterations = 8: the 8-iteration rule). Attaches where
 * src/Diffusion_2D.cpp:144-148 calls Lapack. */
int verb_local_diffusion_learned(verb_ctx* ctx, const verb_grid* g, double* psd, verb_unet* net,
                                 const verb_gmres_params* params, verb_gmres_stats* stats);
/* convection + radial + verb_local_diffusion_learned */
int verb_step_learned(verb_ctx* ctx, const verb_grid* g, double* psd, verb_unet* net,
                      const verb_gmres_params* params, verb_gmres_stats* stats);

/* Data assimilation along L: the Kalman update that
 * DataAssimilationManagerDiffusion::assimilate runs on every (V,K) line of
 * one P slice (src/DataAssimilationDiffusion.cpp:70-249, runKalmanFilter
 * src/DataAssimilation.cpp:18-73). Per line of n = nR points (3 <= n <= 32):
 * F = B A^-1 from the line's implicit 1-D diffusion matrices over g's L
 * grid, DLL and G_radial with zero sources and losses (Ll_type/Lu_type
 * boundaries, time step params->dt); covariance = 5 model_error I when the
 * line's is all zero; Pf = F Pa F^T + model_error I; H selects the line's
 * finite observations; K = Pf H^T (H Pf H^T + observation_error I)^-1;
 * psd += K (obs - H psd); Pa = Pf - K H Pf. psd and obs are device arrays
 * [P][R][V][K] (obs non-finite = not observed); covariance [V][K][n][n]
 * (device, in/out, zero before the first call); model_matrix [V][K][n][n]
 * (device, nullable) receives F. Asynchronous; VERB_ERR_SINGULAR (a zero
 * pivot) and VERB_ERR_DOMAIN (a NaN analysis, where the reference throws)
 * are reported by verb_check_status. */
typedef struct {
    double dt;
    double model_error;
    double observation_error;
    int p_index;
} verb_da_params;
int verb_da_radial(verb_ctx* ctx, const verb_grid* g, double* psd, const double* obs, double* covariance,
                   double* model_matrix, const verb_da_params* params);

/* Data assimilation in the (MLT, R) plane: the Kalman filter that
 * DataAssimilationManagerConvection::assimilate runs on every (V,K) plane
 * (src/DataAssimilationConvection.cpp:42-238), state = the plane's N = nP*nR
 * points. Per plane: the model matrix M (the one-step upwind transport
 * matrix of src/DataAssimilationHelper.cpp:25-102 with loss rate
 * loss/psd, zero where psd >= saturation, for the Courant-split sub-step,
 * raised to the sub-step count) is rebuilt into `model` when
 * recompute_model is set and read from it otherwise; M1d = M^ceil(3/24/dt);
 * Q = M1d diag(tmp) M1d^T + diag(model_error)/10 (tmp = model_error, (1 -
 * c^2) model_error at the R ends, c = exp(-dt/correlation_time), those Q
 * entries reset unless growing_Q_at_boundary); Pf = M Pa M^T + Q; S = H Pf
 * H^T + R (R = observation_error I with use_log, else diag(obs *
 * observation_error)); K = Pf H^T S^-1; x += K (obs - H x), in log space
 * with use_log; Pa = Pf - K H Pf. psd, obs, loss, saturation: device
 * [P][R][V][K] (saturation nullable: none saturated); covariance and model:
 * device [V][K][N][N] (covariance zero before the first call: it starts at
 * 5 model_error I). Planes plane_begin .. +plane_count (0: to the last).
 * process_noise (device [V][K][N][N], nullable): when given, Q is built into
 * it with the model matrix (recompute_model) and read from it otherwise, as
 * it depends only on M, dt and the parameters (the reference rebuilds it on
 * every call: the same values).
 * The dense products run on rocBLAS, S's LU and inverse on rocSOLVER, in
 * chunks of planes. VERB_ERR_DOMAIN (an unstable model matrix, a NaN analysis: the
 * reference throws) is reported by verb_check_status; VERB_ERR_SINGULAR
 * (a zero pivot in S, as for a plane without observations outside log
 * space) is returned. */
typedef struct {
    double dt;
    double correlation_time;
    double model_error;
    double observation_error;
    int use_log;
    int growing_Q_at_boundary;
    int recompute_model;
    long long plane_begin;
    long long plane_count;
    double* process_noise;
} verb_da_conv_params;
int verb_da_convection(verb_ctx* ctx, const verb_grid* g, double* psd, const double* obs, const double* loss,
                       const double* saturation, double* covariance, double* model, const verb_da_conv_params* params);

/* ------------------------------------------------------------------------
 * On-disk formats (host only; no device work). ndim is 3 (boundary arrays)
 * or 4 ([P][R][V][K] fields); sizes[0..ndim) in the array's index order,
 * arrays row-major with the last index fastest.
 * ---------------------------------------------------------------------- */
/* Matrix4D/3D::writeToFile(filename, info) (src/Matrix.cpp:4556-4567,
 * 3011-3024): VARIABLES line with `name` ("f" if empty), ZONE line with the
 * title `info` (the path if empty) and I (fastest), J, K[, L], then one
 * value per line in "%.6e". */
int verb_plt_write(const char* path, int ndim, const int* sizes, const double* data, const char* name,
                   const char* info);
/* writeToFile(filename, grids...) (src/Matrix.cpp:4833-4853, 3031-3050): every
 * line "\t" grid coordinates "\t" value; grid_names nullable ("w","x","y","z"). */
int verb_plt_write_grid(const char* path, int ndim, const int* sizes, const double* const* grids,
                        const char* const* grid_names, const double* data, const char* name);
/* The sizes of a .plt file from its ZONE line, in array index order
 * (sizes[0] = L or K, ...); returns the number of sizes (3 or 4) or an
 * error (the grid.plt parse of src/ReadInitialData.cpp:297-310). */
int verb_plt_sizes(const char* path, int* sizes);
/* readFromFile(filename, column) (src/Matrix.cpp:4906-4965, 3059-3096):
 * the column-th token of each point line; more data than the sizes is an
 * error. */
int verb_plt_read(const char* path, int column, int ndim, const int* sizes, double* out);
/* readFromFile(filename, grids...) (src/Matrix.cpp:5344-5413, 3431-3477):
 * each line's coordinates must agree with grids[a][i] within 1e-6 in
 * log10; the first point that does not is returned in *bad_index
 * (nullable) with VERB_ERR_FORMAT. */
int verb_plt_read_checked(const char* path, int ndim, const int* sizes, const double* const* grids, double* out,
                          long long* bad_index);
/* writeToBinaryFile / readFromBinaryFile (src/Matrix.cpp:5704-5726,
 * 5792-5818, 3781-3838): ndim int32 sizes, then the doubles. The reader
 * reads count(sizes) values and reports the file's own sizes (nullable),
 * which the reference does not check either. */
int verb_pltb_write(const char* path, int ndim, const int* sizes, const double* data);
int verb_pltb_read(const char* path, int ndim, const int* sizes, double* out, int* file_sizes);
/* The P, R, V, K grids of an input folder (folder ends with '/'):
 * grid.pltb (4 int32 sizes, then the four grids) if it opens, else the
 * four columns of grid.plt (src/ReadInitialData.cpp:253-346). With P ==
 * NULL only sizes[4] is filled; else each grid takes `capacity` doubles. */
int verb_grid_read(const char* folder, int* sizes, double* P, double* R, double* V, double* K, size_t capacity);
/* Matrix4D::writeToLstFile (src/Matrix.cpp:5764-5785): appends
 * "info<TAB>filename.plt|.pltb" (the output folder removed from the name)
 * to output_folder + "PSD.lst". */
int verb_psd_lst_append(const char* output_folder, const char* filename, int binary, const char* info);

/* One .tab file's first line as UpdatableMatrix::readFromIniFile /
 * readFromString split it (src/UpdatableMatrix.cpp:308-551): the data field
 * (data_kind 0: none or "-", 1: a data file, 2: a .lst update list, by the
 * reference's suffix test), then, while the line's stream has not ended,
 * the scale field and up to four from/to limit pairs (a field read past
 * trailing blanks is empty, as the reference's is). Each field is a number
 * (verb_is_number) or the name of a "time value" queue file. */
#define VERB_TAB_STRLEN 1024
typedef struct {
    int has_line;       /* 0: empty or comment first line (nothing read) */
    int data_kind;
    int scale_present;
    int nlimits;        /* limit pairs read (the array is limited if > 0) */
    char data[VERB_TAB_STRLEN];
    char scale[VERB_TAB_STRLEN];
    char from[4][VERB_TAB_STRLEN];
    char to[4][VERB_TAB_STRLEN];
} verb_tab_line;
int verb_tab_read(const char* tab_path, verb_tab_line* out);
/* is_number (src/UpdatableMatrix.cpp:1081-1093) and the value an istream
 * reads from such a string (0 when none). */
int verb_is_number(const char* s);
double verb_parse_number(const char* s);
/* fill_queue_from_file (src/UpdatableMatrix.cpp:349-370): "time value"
 * lines; values as numbers (values != NULL) or file names (names != NULL,
 * name_len bytes per entry). *count receives the number of lines with
 * content; call with capacity 0 to size the buffers. */
int verb_queue_read(const char* path, double* times, double* values, char* names, size_t name_len, size_t capacity,
                    size_t* count);

#ifdef __cplusplus
}
#endif
#endif /* VERB4D_HIP_H */
