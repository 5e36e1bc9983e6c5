/*
 * kmsd.h - C ABI of the MI355X (gfx950) stochastic KPM transport library.
 *
 * Drop-in boundary for the reference's compiled CPU kernel, the pybind11
 * module kubo_msd._native built from cpp/native.cpp (module definition at
 * cpp/native.cpp:613-720).  Each entry point below names the reference
 * interface it replaces.
 *
 * Conventions
 *   - Dense blocks are (N, nb) row-major with interleaved complex entries,
 *     exactly the reference's Dense layout (cpp/native.cpp:26-41).  Vector
 *     blocks passed to compute calls are DEVICE pointers owned by the caller
 *     (e.g. torch tensors); the library never frees caller memory.
 *   - dtype selects the vector precision: KMSD_C128 (complex128, the
 *     reference's only type) or KMSD_C64 (complex64 storage, FP64 arithmetic
 *     in registers and FP64 reductions).
 *   - Matrices and metric data are HOST arrays at creation time and are
 *     copied to device memory; nothing is retained from the caller.
 *   - Calls are ordered on the context's stream.  Calls that must report a
 *     numerical status (solves, series, chains, moments) synchronize that
 *     stream before returning.
 *   - Return codes: KMSD_OK, KMSD_EINVAL (reference: std::invalid_argument ->
 *     ValueError), KMSD_ENUMERIC (reference: std::runtime_error /
 *     MetricFailure -> RuntimeError with per-column diagnostics),
 *     KMSD_ERESOURCE (message starts with "RESOURCE_BLOCKED"), KMSD_EDEVICE.
 *     kmsd_last_error() gives the message of the last failure in the context.
 */
#ifndef KMSD_H
#define KMSD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { KMSD_OK = 0, KMSD_EINVAL = 1, KMSD_ENUMERIC = 2, KMSD_ERESOURCE = 3, KMSD_EDEVICE = 4 };
enum { KMSD_C64 = 1, KMSD_C128 = 2 };
/* Or-ed into kmsd_autotune's dtype: time the product with FP32 matrix values,
 * the form the inner solves of a complex64 metric solve use. */
enum { KMSD_TUNE_F32_VALUES = 0x100 };
/* Or-ed into kmsd_autotune's dtype: choice has three entries, and for a
 * 4x4-block matrix with a plane stride the tuner also times the plane-tiled
 * walk's tile size ("tile_bytes": 0 = natural order, 512 KiB, 1 MiB, 2 MiB)
 * with the chosen shape and writes it to choice[2]. */
enum { KMSD_TUNE_TILE = 0x200 };
enum { KMSD_CSR = 0, KMSD_BSR4 = 1, KMSD_AUTO = 2 };
enum { KMSD_PRECOND_DIAGONAL = 0, KMSD_PRECOND_BLOCK_JACOBI = 1 };
enum { KMSD_MU_FP32 = 0, KMSD_MU_BF16X3 = 1, KMSD_MU_FP64 = 2 };

typedef struct kmsd_ctx kmsd_ctx;
typedef struct kmsd_matrix kmsd_matrix;
typedef struct kmsd_metric kmsd_metric;
typedef struct kmsd_engine kmsd_engine;

/* Per-column metric-solve report (reference SolveReport, cpp/native.cpp:123-143).
 * Arrays are caller-owned host arrays of length nb; any may be NULL. */
typedef struct {
    uint8_t *converged;
    uint8_t *breakdown;
    int64_t *iterations;
    double *true_residual;
    double *relative_residual;
} kmsd_report;

/* ---- context --------------------------------------------------------- */
/* No reference counterpart (the reference CPU kernel has no device).  Binds
 * a HIP device and stream (NULL = null stream) and an HBM budget in bytes
 * (0 = 85% of free device memory) that every library allocation is charged
 * against, as the reference GPU edition's require_bytes does
 * (kubo_msd_gpu/src/kubo_msd/backends/gpu_context.py:128-136). */
/* Threads: every call on a context, or on a matrix, metric or engine made
 * from it, is serialized by the context's lock, so any thread may call (or
 * destroy) them.  All live contexts of one process belong to the host thread
 * that created the first of them: a create from another thread while those
 * live returns KMSD_EINVAL ("one host thread"); run concurrent device work in
 * separate processes. */
int kmsd_ctx_create(int device, uint64_t hbm_budget_bytes, void *stream, kmsd_ctx **out);
int kmsd_ctx_destroy(kmsd_ctx *ctx);
int kmsd_ctx_set_stream(kmsd_ctx *ctx, void *stream);
int kmsd_ctx_synchronize(kmsd_ctx *ctx);
/* bytes currently charged to the budget and the budget itself */
int kmsd_ctx_memory(const kmsd_ctx *ctx, uint64_t *used, uint64_t *budget);
/* Transfer ledger, the counterpart of the reference GPU edition's transfer
 * accounting (src/kubo_msd/backends/gpu.py:16-29 upload_bytes /
 * download_bytes; kubo_msd_gpu/src/kubo_msd/backends/gpu_context.py:75-120):
 * bytes the library copied on this context since it was created.
 * out3[0] host-to-device (matrix and metric uploads, coefficient tables,
 * solver state), out3[1] device-to-host array data (only host-returning
 * set-up calls such as kmsd_ic0_factor), out3[2] device-to-host status
 * (nonfinite flags, per-column solver state, reports and norms).  Compute
 * calls on device blocks copy no array data to the host. */
int kmsd_ctx_transfers(const kmsd_ctx *ctx, uint64_t *out3);
/* Launch tuning: key "groups" (workgroups per launch, multiple of 8; 0 =
 * automatic, the default: each fused step launches one workgroup per
 * resident slot of its kernel on the device), "cpl" (complex columns per
 * lane: 1, 2 or 4; 0 = automatic: 2 for complex64 and 1 for complex128
 * blocks, 4 and 2 for products with 4x4-block matrices; "cpl" sets both
 * precisions, "cpl_c64" and "cpl_c128" one of them),
 * "moment_chunk" (Chebyshev orders per partial-reduction launch),
 * "fp32_math" (complex64 blocks: FP32 values and row sums), "pcg_poll" (PCG
 * iterations between host polls), "graphs" (1: moment chains with an
 * identity metric replay whole chunks of orders as one HIP graph, and the
 * level-by-level IC(0) sweeps of a preconditioner application replay as one
 * graph; 0: launch every kernel; same results either way), "ell_kernel" (1, the default:
 * products whose off-diagonal part is ELL of width 3, 4 or 6 use the kernel
 * that prefetches the next row's indices; 0: the generic row kernel; same
 * results either way), "fuse_velocity" (1, the default: kmsd_chain_v forms
 * V T_k inside the chain's steps where it can; 0: always the regrouping path;
 * same results either way), "tile_bytes" (4x4-block products of matrices
 * with a plane stride walk the block rows in tiles of this many gathered
 * bytes across planes; 0: natural order; same results either way),
 * "bsr_mfma" (1, the default: FP32-valued 4x4-block products of complex64
 * blocks, the metric products of complex64 inner solves, run on the matrix
 * cores with FP32 row sums; 0: the VALU kernel with FP64 row sums),
 * "bsr_lds" (1, the default: FP64-valued 4x4-block products stage each
 * chunk's block columns and values in LDS; 0: the kernel that loads them
 * block by block; same results either way), "min_blocks" (1..64, default 1:
 * row-walking launches use at most one workgroup per this many row blocks,
 * so small operators get fewer, fuller workgroups and fewer partial sums),
 * "pcg_min_blocks" (1..64, default 4: the same for the vector passes of PCG
 * solves, keeping at least 512 workgroups where there are that many row
 * blocks; their S products then use no more workgroups than the passes;
 * both change the partial-sum partition, so sums round differently),
 * "cheb_inner" (1, the default: complex64 metric solves whose S is CSR/ELL
 * with the diagonal or no preconditioner and has a Gershgorin interval
 * [lo, hi] with lo > 0 and hi <= 4 lo take Chebyshev semi-iteration inner
 * solves on that interval inside the FP64 refinement, one fused product per
 * step and no reductions; uniform 4x4 block-Jacobi on a 4x4-block S: the first inner
 * solve is PCG, whose CG coefficients give Ritz bounds for the block-Jacobi
 * Chebyshev semi-iteration in the later ones; the acceptance stays the FP64 true residual, and
 * a series round that reduces some column's residual by less than the square
 * root of the reduction its interval promised returns that metric to PCG
 * inner solves; 0: PCG inner solves), "bcheb_lo_pct" (1..10000, default 96:
 * the block-Jacobi iteration interval's lower end in percent of the smallest
 * Ritz value;
 * its upper end is the largest Ritz value plus 2 %). */
int kmsd_ctx_set_tuning(kmsd_ctx *ctx, const char *key, int64_t value);
/* No reference counterpart (the reference's CPU kernel has no launch shape).
 * Times the fused Chebyshev step of a moment chain with product A on an
 * nb-column block of dtype for every columns-per-lane (1, 2, 4) and
 * workgroup count (automatic, 1024, 2048, 4096), keeps the default shape
 * (automatic workgroups) unless another is more than 3 % faster, applies
 * the choice to ctx ("cpl_c64" or "cpl_c128" for the tuned precision, and
 * "groups") and writes it to choice[0..1] (groups 0 = automatic).  dtype may carry KMSD_TUNE_F32_VALUES (tune for the metric
 * products of complex64 solves) and KMSD_TUNE_TILE (also the 4x4-block tile
 * size, into choice[2]; the default 1 MiB stays unless another is more than
 * 3 % faster).  Uses three scratch blocks of n*nb elements from the budget. */
int kmsd_autotune(kmsd_ctx *ctx, const kmsd_matrix *A, int64_t nb, int dtype, int64_t *choice);
/* No reference counterpart (the reference's sample batch is a fixed
 * setting).  Times one fused Chebyshev step of a moment chain with product A
 * on an nb-column block of dtype at the context's current launch shape
 * (median of three after a warm-up) and writes microseconds per step to *us;
 * the vector-block width tuner compares these per column across widths.
 * Uses three scratch blocks of n*nb elements from the budget. */
int kmsd_time_step(kmsd_ctx *ctx, const kmsd_matrix *A, int64_t nb, int dtype, double *us);
/* Message of the last failure in ctx (or of the last failed create when ctx
 * is NULL).  Valid until the next call on ctx. */
const char *kmsd_last_error(const kmsd_ctx *ctx);
/* Reference: build_info() (cpp/native.cpp:631-639).  Writes a short
 * NUL-terminated description into buf. */
int kmsd_build_info(char *buf, int64_t len);

/* ---- matrices -------------------------------------------------------- */
/* Reference: CSR(indptr:int64[n+1], indices:int64[nnz], data:c128[nnz], n,
 * threads) (cpp/native.cpp:54-74, 640-646).  Same validation: monotone
 * pointers starting at 0 and ending at nnz, indices in [0,n), finite values.
 * values: complex128 interleaved (2*nnz doubles).  Purely real or purely
 * imaginary matrices are detected and stored as one real array.
 * fmt is the device storage: KMSD_CSR (diagonal split off, ELL when rows
 * are near-uniform), KMSD_BSR4 (4x4 blocks; n a multiple of 4; the input is
 * still CSR and is blocked here) or KMSD_AUTO (BSR4 when n is a multiple of
 * 4 and the nonzero 4x4 blocks are at least 75% full, else CSR). */
int kmsd_matrix_create(kmsd_ctx *ctx, int fmt, int64_t n, int64_t nnz, const int64_t *rowptr,
                       const int64_t *colidx, const double *values, kmsd_matrix **out);
int kmsd_matrix_destroy(kmsd_matrix *m);
/* No reference counterpart (the reference's CSR exposes no layout): out[8] =
 * {storage format (KMSD_CSR/KMSD_BSR4), n, nnz as given, stored entries
 * (off-diagonal or block entries), device bytes, value kind (0 real,
 * 1 imaginary, 2 complex), uniform values (1: ELL storage whose stored
 * off-diagonal values, padding aside, are all equal: the width-3/4/6 ELL
 * kernel reads the value once), plane stride (BSR4: block rows between a
 * block row and its outermost stencil neighbour, which products walk in
 * tiles across planes; 0: natural order)}. */
int kmsd_matrix_info(const kmsd_matrix *m, int64_t *out);
/* Reference: CSR.matmul(B) (cpp/native.cpp:75-89, 642-646): out = A B. */
int kmsd_matrix_matmul(kmsd_matrix *m, const void *B, void *out, int64_t nb, int dtype);

/* ---- metric (overlap S) ---------------------------------------------- */
/* Reference: Metric(S, diag:f64[n], blocks:[(a,b,L)], threads)
 * (cpp/native.cpp:212-236, 647-648).  diag: host f64[n] (> 0, finite).
 * nblocks contiguous atom blocks [block_ptr[q], block_ptr[q+1]) with lower
 * Cholesky factors concatenated row-major (complex128 interleaved);
 * nblocks = 0 selects the diagonal preconditioner.  Blocks may have any size
 * (one block of n rows is the full Cholesky factor).  An S that is exactly
 * the identity with the diagonal preconditioner is detected; its solves
 * return the right-hand side (the reference's PCG converges there in one
 * step with alpha = 1). */
int kmsd_metric_create(kmsd_ctx *ctx, kmsd_matrix *S, const double *diag, int64_t nblocks,
                       const int64_t *block_ptr, const double *factors, kmsd_metric **out);
int kmsd_metric_destroy(kmsd_metric *m);
/* Reference: Metric.set_residual_policy(policy, interval, gap)
 * (cpp/native.cpp:206-210): adaptive 0 = "fixed", 1 = "adaptive". */
int kmsd_metric_set_residual_policy(kmsd_metric *m, int adaptive, int interval, double gap);
/* No counterpart in the reference (its solves are FP64 only); for callers
 * running in mixed precision.  enabled = 1: complex128 right-hand sides are
 * solved like complex64 ones, by FP32 inner solves inside the FP64
 * iterative refinement, whose acceptance is the same FP64 true-residual test
 * (max(rtol |q|, atol/safe) per column) as the FP64 PCG; the solutions then
 * differ from the FP64 PCG's within that tolerance.  0 (the default): FP64
 * PCG for complex128 blocks. */
int kmsd_metric_set_mixed(kmsd_metric *m, int enabled);
/* Reference: Metric.solve(B, rtol, atol, maxiter) -> (Y, report)
 * (cpp/native.cpp:313-452, 649-654).  Batched PCG, one independent system per
 * column, per-column RHS scaling, true-residual refresh every `interval`
 * iterations, final original-equation acceptance.  Returns KMSD_ENUMERIC with
 * the report filled when a column fails. */
int kmsd_metric_solve(kmsd_metric *m, const void *B, void *Y, int64_t nb, int dtype, double rtol, double atol,
                      int maxiter, kmsd_report *report);
/* Reference: Metric.sample(xi, coeff, center, half) (cpp/native.cpp:453-479):
 * out = Z p(W) xi with W = Z S Z, Z = diag^-1/2, p by Clenshaw. */
int kmsd_metric_sample(kmsd_metric *m, const void *xi, void *out, int64_t nb, int dtype, const double *coeff,
                       int64_t ncoeff, double center, double half);
/* Reference: OperatorBundle.inner(A, B) = sum(conj(A) * (S @ B), axis=0)
 * (src/kubo_msd/operators/velocity.py:117).  out: DEVICE complex128[nb]. */
int kmsd_metric_inner(kmsd_metric *m, const void *A, const void *B, double *out, int64_t nb, int dtype);
/* Reference: Metric.diagnostics() counters (cpp/native.cpp:666-675):
 * out[0] calls, [1] sum of max iterations, [2] refreshes, [3] residual
 * checks, [4] restarts; *max_relative. */
int kmsd_metric_counters(const kmsd_metric *m, int64_t *out5, double *max_relative);
/* Reference: Metric.set_index_blocks([(ids, L)]) (cpp/native.cpp:180-197).
 * Cluster-Jacobi blocks replacing the contiguous ones: block q holds the rows
 * rows[block_off[q] .. block_off[q+1]) (any order, each row of S exactly
 * once) and its lower Cholesky factor (m x m row-major complex128, factors
 * concatenated).  Errors: "invalid cluster factor", "invalid cluster
 * coverage", "incomplete cluster coverage". */
int kmsd_metric_set_index_blocks(kmsd_metric *m, int64_t nblocks, const int64_t *block_off, const int64_t *rows,
                                 const double *factors);
/* Reference: Metric.set_ic0(L, p, scaling) (cpp/native.cpp:163-179).  IC(0)
 * preconditioner M^-1 r = D P^T L^-H L^-1 P D r: L in CSR (n rows, host
 * arrays, complex128 values) with each row's strictly lower entries first
 * and its real positive diagonal last, perm[i] the original row of factor
 * row i, scaling D per original row.  Takes precedence over blocks and the
 * diagonal.  Errors: "invalid IC permutation/scaling", "invalid IC factor
 * diagonal", and (stricter than the reference, which the level schedules
 * need) "invalid IC factor pattern" for an entry right of the diagonal. */
int kmsd_metric_set_ic0(kmsd_metric *m, int64_t nnz, const int64_t *rowptr, const int64_t *colidx,
                        const double *values, const int64_t *perm, const double *scaling);
/* Reference: Metric.set_direct(enabled, steps) (cpp/native.cpp:198-205).
 * Direct solves with the single full-size block factor: y = M^-1 b plus up
 * to `steps` refinement corrections on the original equation; iterations
 * are reported as 0.  Enabling needs exactly one block covering all rows
 * ("invalid direct metric factor"). */
int kmsd_metric_set_direct(kmsd_metric *m, int enabled, int steps);
/* Reference: MetricSolver.precondition(R) (src/kubo_msd/metric/solve.py:59-77;
 * the compiled Metric applies it only inside solve): Z = M^-1 R with the
 * metric's preconditioner (IC(0), else the diagonal without blocks, else the
 * blocks), one application; R and Z are (n, nb) device blocks that must not
 * overlap. */
int kmsd_metric_precondition(kmsd_metric *m, const void *R, void *Z, int64_t nb, int dtype);
/* Reference: Metric.diagnostics()["refinement_steps"] (cpp/native.cpp:668-672). */
int kmsd_metric_refinement_steps(const kmsd_metric *m, int64_t *out);
/* Reference: ic0_factor(CSR, shift) (cpp/native.cpp:92-121, 614-618).  IC(0)
 * values on the given sorted lower pattern (host arrays in and out,
 * complex128 interleaved), computed on the device by dependency levels with
 * the reference's arithmetic (same values to the last bit).  Errors:
 * KMSD_EINVAL "invalid IC shift" / pattern messages; KMSD_ENUMERIC "IC
 * nonpositive/nonfinite pivot at row i" for the first failing row. */
int kmsd_ic0_factor(kmsd_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rowptr, const int64_t *colidx,
                    const double *values, double shift, double *out);

/* ---- engine ----------------------------------------------------------- */
/* Reference: Engine(H, Metric, [(kind, [CSR...]) x3], rtol, atol, maxiter,
 * hbar) (cpp/native.cpp:488-503, 676-677).  kinds[3] per Cartesian axis:
 * -1 absent, 0 explicit velocity V (1 matrix), 1 position X (1 matrix),
 * 2 composite bond/xbar/delta (3 matrices).  mats[3*axis + k]. */
int kmsd_engine_create(kmsd_ctx *ctx, kmsd_matrix *H, kmsd_metric *metric, const int *kinds,
                       kmsd_matrix *const *mats, double rtol, double atol, int maxiter, double hbar,
                       kmsd_engine **out);
int kmsd_engine_destroy(kmsd_engine *e);
/* Solve controls used by the engine's actions (reference
 * v_cov_with_controls copies the engine with new controls,
 * cpp/native.cpp:688-697). */
int kmsd_engine_set_controls(kmsd_engine *e, double rtol, double atol, int maxiter);
/* Reference: read-only h_calls / v_calls (cpp/native.cpp:718-719). */
int kmsd_engine_counters(const kmsd_engine *e, int64_t *h_calls, int64_t *v_calls);
/* Report of the engine's last failed metric solve (nb columns). */
int kmsd_engine_last_report(const kmsd_engine *e, kmsd_report *report, int64_t *nb);
/* Reference: Engine.h_action(B) = S^-1 H B (cpp/native.cpp:505, 678-682). */
int kmsd_h_action(kmsd_engine *e, const void *B, void *out, int64_t nb, int dtype);
/* Reference: Engine.v_cov(B, dir[3]) (cpp/native.cpp:506-533, 683-687). */
int kmsd_v_cov(kmsd_engine *e, const double *dir3, const void *B, void *out, int64_t nb, int dtype);
/* Reference: Engine.series(B, coeff, center, half, dirs)
 * (cpp/native.cpp:540-573, 698-705): out = sum_k c_k T_k(A) B where A is the
 * scaled action, augmented with the commutator source (-i hbar/half)
 * S^-1 V_d B0 in column block d+1 when ndirs > 0.  coeff: complex128
 * interleaved host array; dirs: host f64[ndirs][3]. */
int kmsd_series(kmsd_engine *e, const void *B, void *out, int64_t width, int dtype, const double *coeff,
                int64_t ncoeff, double center, double half, const double *dirs, int64_t ndirs);
/* Reference: Engine.chain(B, L, center, half) -> (L, N, nb)
 * (cpp/native.cpp:574-589, 706-711).  out: device (L, N, nb) block. */
int kmsd_chain(kmsd_engine *e, const void *B, void *out, int64_t L, int64_t nb, int dtype, double center,
               double half);
/* Reference: the spectral path's cov = V chain(S^-1 V c, K+1)
 * (src/kubo_msd/solvers/spectral.py:33-36, the v_cov of a stored chain):
 * out[k] = v_cov(dir, T_k(H~) B), (L, N, nb).  With an identity metric, H a
 * real ELL matrix of width 3, 4 or 6 and dir along one axis whose velocity
 * is one real or imaginary matrix with no diagonal and its entries in H's
 * pattern, each step that gathers T_k for T_(k+1) also forms V T_k (the
 * velocity's values in H's ELL slots, n*width doubles, are built on the
 * first call and kept by the engine; the chain itself uses two N*nb
 * blocks).  Otherwise one velocity product covers all orders (the chain is
 * regrouped to N rows of L*nb columns and back on the device; two scratch
 * blocks of L*N*nb elements come from the budget and stay with the engine).
 * Both give the same bits; the engine's blocks are released by
 * kmsd_engine_destroy. */
int kmsd_chain_v(kmsd_engine *e, const double *dir, const void *B, void *out, int64_t L, int64_t nb, int dtype,
                 double center, double half);
/* Reference: Engine.moments(B, L, center, half) -> f64 (L, nb)
 * (cpp/native.cpp:590-609, 712-717): mu[l][j] = Re <S B_j, T_l(A) B_j>.
 * mu: DEVICE f64 (L, nb). */
int kmsd_moments(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype, double center,
                 double half);

/* No reference counterpart: kmsd_moments for `count` engines of one context
 * at once (continuous batching of small transport jobs): each engine e[q]
 * with an identity metric, its block B[q] (n_q x nb, dtype), its output
 * mu[q] (DEVICE f64 (L, nb)) and its own scale (center[q], half[q]; host
 * arrays of count entries).  One width nb, length L and dtype for all.
 * Orders 0-2 run chain by chain; from order 3 on, every chain's step of one
 * order is one launch shared by all (each chain keeps its own matrix format
 * and the workgroup count, partial sums and reductions of its own
 * kmsd_moments call, so its moments are the same bits); chains whose
 * kernels cannot share a launch (4x4-block or complex matrices) run one
 * launch each, with the same result. */
int kmsd_moments_group(int64_t count, kmsd_engine *const *e, const void *const *B, double *const *mu, int64_t L,
                       int64_t nb, int dtype, const double *center, const double *half);

/* No reference counterpart: the same moments as kmsd_moments (S = I only),
 * from the Chebyshev doubling identities T_2k = 2 T_k^2 - 1 and
 * T_2k+1 = 2 T_k T_k+1 - T_1: mu_2k = 2<T_k B, T_k B> - mu_0 and
 * mu_2k+1 = 2 Re<T_k B, T_k+1 B> - mu_1.  About L/2 Chebyshev orders instead of
 * L - 1, and no projection-block stream.  Exact in arithmetic; rounding
 * differs from the direct recursion. */
int kmsd_moments_doubling(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype, double center,
                          double half);

/* No reference counterpart: kmsd_moments for callers that cannot allocate
 * device memory (serving loops, harnesses that forbid allocation).  Same
 * arguments and bitwise the same moments as kmsd_moments; needs an identity
 * metric (S = I with the diagonal preconditioner; else KMSD_EINVAL).  The two
 * chain blocks (n*nb elements each) and the partial sums live in a fixed
 * 256 MiB per-device array that the library holds from the first context
 * created on a device until the process ends; a block beyond it is rejected
 * with KMSD_EINVAL ("fixed scratch").  Allocates no device memory and charges
 * nothing to the budget.  Every call on one GPU shares that array, and the
 * call synchronizes before it returns: calls that overlap on one GPU (from
 * several processes' graphs or streams replaying it) race. */
int kmsd_moments_noalloc(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype,
                         double center, double half);

/* ---- spectral double moments -------------------------------------------- */
/* Reference: einsum('jnb,knb->bjk', left.conj(), cov) (src/kubo_msd/solvers/
 * spectral.py:37; GPU edition solvers/spectral.py:65).  left: (m, N, nb),
 * right: (k, N, nb) device blocks; table: device complex128 (nb, m, k),
 * table[b][j][q] = sum_i conj(left[j][i][b]) right[q][i][b].
 * accumulate != 0 adds into table (chunked N).  mode: KMSD_MU_FP64 (FP64
 * vector arithmetic), KMSD_MU_FP32 (MFMA FP32), KMSD_MU_BF16X3 (MFMA bf16
 * three-term split). */
int kmsd_mu_table(kmsd_ctx *ctx, const void *left, const void *right, void *table, int64_t n, int64_t m,
                  int64_t k, int64_t nb, int dtype, int mode, int accumulate);

/* ---- statistics of the stochastic estimator ------------------------------ */
/* Reference: the per-sample accumulation loops of solvers/stochastic.py:60-71
 * (dos_blocks[block] += dh[local]; accum[key][block] += vh[local], samples in
 * order; GPU edition solvers/stochastic_gpu.py:88-93).  Device FP64 buffers:
 * dst[e] += src[l*ld_src + e] for l = 0 .. nloc-1 in that order, e < nelem;
 * if any src value is not finite, flag_bits are OR-ed into the device int32
 * *flags (the caller clears and reads it).  Stream-ordered, asynchronous. */
int kmsd_block_accumulate(kmsd_ctx *ctx, double *dst, const double *src, int64_t nloc, int64_t nelem,
                          int64_t ld_src, int32_t *flags, int32_t flag_bits);
/* Reference: msd_gradient (observables/diffusion.py:4-6):
 * dst = scale * numpy.gradient(src, time, axis=1, edge_order=2 if nt > 2 else 1)
 * for device FP64 src/dst of shape (outer, nt, inner); time: host, nt >= 2
 * strictly increasing values.  numpy's formulas and operation order (uniform
 * spacing when all differences are equal), so results equal numpy's bit for
 * bit.  Synchronous. */
int kmsd_time_gradient(kmsd_ctx *ctx, const double *src, double *dst, int64_t outer, int64_t nt, int64_t inner,
                       const double *time, double scale);
/* Reference: estimate_ratio (observables/statistics.py:5-23; GPU edition
 * observables/statistics_gpu.py:6-32), paired block bootstrap of sums.
 * A: device (nblocks, nelem) block sums, element e paired with column e % ne
 * of D: device (nblocks, dcols), dcols = ne or 1 (one denominator per block).
 * value[e] = sum_b A / sum_b D (NaN when that is 0); stderr_out (optional,
 * needs nblocks >= 2): standard deviation (ddof 1) over the nres resampled
 * ratios with weights (device, nres x nblocks: the counts of each block in
 * each resample, drawn on the host as the reference does); NaN when any
 * ratio is undefined.  valid (optional, device, ne bytes): value and stderr
 * are NaN where valid[e % ne] == 0.  total (optional): sum_b A.  The block
 * sums run in block order (value and total equal numpy's bit for bit); the
 * resampled sums differ from a BLAS product only by rounding.  Asynchronous. */
int kmsd_bootstrap_ratio(kmsd_ctx *ctx, const double *A, const double *D, int64_t nblocks, int64_t nelem, int64_t ne,
                         int64_t dcols, const double *weights, int64_t nres, const uint8_t *valid, double *value,
                         double *stderr_out, double *total);
/* Reference: the DOS validity masks of _normalize_dos (solvers/stochastic.py:
 * 78-87): valid = finite(rho) & rho > threshold & !(stderr not finite or
 * rho - 1.96 stderr <= 0); reason = LOW_DOS (1) | UNCERTAIN_TRACE (2) |
 * INSUFFICIENT_BLOCKS (32, nblocks < 8).  stderr_in may be NULL (one block).
 * Device buffers of ne entries.  Asynchronous. */
int kmsd_dos_validity(kmsd_ctx *ctx, const double *rho, const double *stderr_in, int64_t ne, double threshold,
                      int64_t nblocks, uint8_t *valid, uint32_t *reason);

#ifdef __cplusplus
}
#endif
#endif
