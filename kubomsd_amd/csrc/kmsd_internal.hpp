// Internal host-side declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "kmsd.h"

struct kmsd_ctx {
    // Serializes every call on this context and on the objects created from
    // it, whichever host thread makes it (a destroy from a garbage-collector
    // thread waits for a running call instead of freeing its buffers or
    // touching its stream mid-call).
    mutable std::recursive_mutex mu;
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, used = 0;
    std::string err;
    int64_t groups = 2048;  // workgroups per row-walking launch (multiple of 8)
    bool auto_groups = true; // fused steps: one workgroup per resident slot of the kernel (else `groups`)
    int min_blocks = 1;      // row blocks per workgroup at least (small operators: fewer, fuller workgroups)
    int pcg_min_blocks = 4;  // the same for the vector passes and products of PCG solves
    bool cheb_inner = true;  // complex64 refinement: Chebyshev inner solves where S's Gershgorin interval allows
    int bcheb_lo_pct = 96;   // block-Jacobi Chebyshev iteration: lower end in percent of the smallest Ritz value
    int cpl[2] = {0, 0};    // columns per lane of complex64 [0] and complex128 [1] blocks; 0 = automatic
    int moment_chunk = 16;  // Chebyshev orders per partial-reduction launch
    bool fp32_math = false; // complex64 blocks: FP32 matrix values and row sums
    int pcg_poll = 4;       // PCG iterations between host status polls
    bool graphs = true;     // replay moment-chain chunks as HIP graphs (identity metric)
    bool ell_kernel = true; // fixed-width ELL products (widths 3, 4, 6) use the prefetching kernel
    bool bsr_mfma = true;   // FP32-valued 4x4-block products of complex64 blocks on the matrix cores
    bool bsr_lds = true;    // FP64-valued 4x4-block products stage block columns and values in LDS
    int64_t tile_bytes = 1 << 20; // BSR4 plane-tiled walk: gathered-block bytes per tile (0: natural order)
    bool fuse_velocity = true; // kmsd_chain_v forms V T_k inside the chain's steps where V fits H's ELL slots
    hipStream_t capture = nullptr;  // private non-blocking stream for graph capture (lazy)
    int32_t *flag = nullptr;  // device scratch flags (nonfinite detection)
    int cvt_slot = 0;         // flag[1 + cvt_slot]: the next launch_convert_check's flag
    void *pin = nullptr;      // pinned host staging for status reads (kmsd::kPinBytes, allocated at creation)
    void *mu_part = nullptr;  // partial tables of kmsd_mu_table (grow-only)
    size_t mu_part_bytes = 0;
    void *stats_tab = nullptr;  // coefficient table of kmsd_time_gradient (grow-only)
    size_t stats_tab_bytes = 0;
    // Transfer ledger (kmsd_ctx_transfers): bytes the library copied between
    // host and device on this context.
    uint64_t h2d_bytes = 0;         // uploads: matrices, metric data, coefficient tables, solver state
    uint64_t d2h_data_bytes = 0;    // downloads of array data (host-returning set-up calls)
    uint64_t d2h_status_bytes = 0;  // downloads of flags, per-column solver state and norms
    // Matrices, metrics and engines made from this context that are still
    // alive.  kmsd_ctx_destroy only closes a context that has any; the last
    // child's destroy frees it, so objects may be destroyed in any order.
    int64_t children = 0;
    bool closed = false;
};

// Value kinds of a stored matrix.
enum { VK_REAL = 0, VK_IMAG = 1, VK_COMPLEX = 2 };

struct kmsd_matrix {
    kmsd_ctx *ctx = nullptr;
    int fmt = KMSD_CSR;
    int64_t n = 0, nnz = 0;  // nnz of the full matrix as given
    // Device layout (KMSD_CSR): the diagonal as a dense array plus the
    // off-diagonal part as ELL (when rows have near-uniform length) or CSR.
    // KMSD_BSR4: ptr/idx are block-row pointers and block columns of the
    // nonzero 4x4 blocks, val* 16 row-major entries per block, no split diagonal.
    int64_t noff = 0;        // stored off-diagonal entries (ELL: n*ell incl. padding; BSR4: 16 * blocks)
    int32_t ell = 0;         // ELL width, 0 = CSR
    int64_t uniform_slot = -1;  // ELL whose off-diagonal values are all equal: a slot holding it (-1: not)
    int32_t *ptr = nullptr;  // CSR row pointers of the off-diagonal part (CSR only)
    int32_t *idx = nullptr;
    int vkind = VK_COMPLEX;
    void *val64 = nullptr;   // off-diagonal values: double[] or cx<double>[]
    void *val32 = nullptr;   // float[] or cx<float>[]
    bool has_diag = false;
    void *dval64 = nullptr;  // diagonal values (n)
    void *dval32 = nullptr;
    bool identity = false;   // exactly the identity matrix
    // BSR4: block-row stride of the matrix's plane neighbours (a 3D stencil's
    // outermost stride), 0 when there is none.  Products then walk the block
    // rows tile by tile across planes (kmsd::tile_row), so the +-plane rows a
    // block row gathers were fetched a tile earlier and are still in L2.
    int32_t plane = 0;
    size_t bytes = 0;        // device bytes held
};

namespace kmsd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                                           \
    do {                                                                                                      \
        hipError_t e_ = (x);                                                                                  \
        if (e_ != hipSuccess) throw ::kmsd::Error(KMSD_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

// HIPCHECK that first restores the context's stream (used around capture).
#define HIPCHECK_RESTORE(x, ctx, saved)                                                                     \
    do {                                                                                                    \
        hipError_t e_ = (x);                                                                                \
        if (e_ != hipSuccess) {                                                                             \
            (ctx)->stream = (saved);                                                                        \
            throw ::kmsd::Error(KMSD_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_));                \
        }                                                                                                   \
    } while (0)

// Host <-> device copies on the context's stream, counted in its transfer
// ledger.  status: the bytes are flags, solver reports or per-column norms
// (control information), not array data.
inline void copy_h2d(kmsd_ctx *c, void *dst, const void *src, size_t bytes) {
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    c->h2d_bytes += bytes;
}
inline void copy_d2h(kmsd_ctx *c, void *dst, const void *src, size_t bytes, bool status) {
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    (status ? c->d2h_status_bytes : c->d2h_data_bytes) += bytes;
}

// Status bytes (flags, per-column solver state, norms) to the host in one
// stream synchronization: each part is copied asynchronously into the
// context's pinned staging buffer, then to dst after the synchronization.  A
// part that does not fit the staging buffer takes a pageable copy.  The
// staging buffer exists from context creation, so no call allocates for it.
constexpr size_t kPinBytes = size_t(256) << 10;
struct StatusPart {
    void *dst;
    const void *src;
    size_t bytes;
};
void read_status(kmsd_ctx *c, std::initializer_list<StatusPart> parts);

extern thread_local std::string g_create_error;

// Child objects of a context (c->mu held): adopt on creation; orphan on
// destruction returns true when c was closed and this was its last child, and
// the caller then frees c with ctx_free after releasing c->mu.
inline void ctx_adopt(kmsd_ctx *c) { ++c->children; }
inline bool ctx_orphan(kmsd_ctx *c) { return --c->children == 0 && c->closed; }
void ctx_free(kmsd_ctx *c);

template <class F>
int guard(kmsd_ctx *c, F &&f) {
    std::unique_lock<std::recursive_mutex> lock;
    if (c) lock = std::unique_lock<std::recursive_mutex>(c->mu);
    try {
        f();
        if (c) c->err.clear();
        return KMSD_OK;
    } catch (const Error &e) {
        if (c) c->err = e.what();
        else g_create_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        if (c) c->err = "host allocation failed";
        return KMSD_ERESOURCE;
    } catch (const std::exception &e) {
        if (c) c->err = e.what();
        return KMSD_EDEVICE;
    }
}

void *dalloc(kmsd_ctx *c, size_t bytes);
void dfree(kmsd_ctx *c, void *p, size_t bytes);

// Grow-only device workspace charged to the context budget.
struct DBuf {
    kmsd_ctx *c = nullptr;
    void *p = nullptr;
    size_t bytes = 0;
    void *ensure(kmsd_ctx *ctx, size_t want);
    void release();
    ~DBuf() { release(); }
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

inline size_t elem_bytes(int dtype) { return dtype == KMSD_C64 ? 8 : 16; }
void check_block_args(const void *B, const void *out, int64_t nb, int dtype);

void matrix_upload(kmsd_ctx *c, kmsd_matrix *m, int fmt, int64_t n, int64_t nnz, const int64_t *rowptr,
                   const int64_t *colidx, const double *values);
void matrix_release(kmsd_matrix *m);

// ---- launches (kmsd_launch.hip) -------------------------------------------
struct Cplx {
    double re = 0, im = 0;
};

// Generic fused row step:
//   y   = A X (A != null) or Z
//   t   = alpha*y + beta*X + gamma*P
//   OUT = t (store) ; part += conj(W) t (W != null) ; ACC += coef*t (ACC != null)
// part: complex FP64 partial sums, [groups][w] per step slot.
struct Step {
    const kmsd_matrix *A = nullptr;
    const void *X = nullptr;
    int64_t ldx = 0;
    const void *Z = nullptr;
    int64_t ldz = 0;
    const void *P = nullptr;
    int64_t ldp = 0;
    void *OUT = nullptr;
    int64_t ldo = 0;
    const void *W = nullptr;
    int64_t ldw = 0;
    void *ACC = nullptr;
    int64_t lda = 0;
    double *part = nullptr;
    Cplx alpha{1, 0}, beta{0, 0}, coef{1, 0};
    double gamma = 0;
    bool use_beta = false, use_p = false;
    int64_t n = 0;
    int32_t w = 0;
    int cpl_cap = 4;            // columns-per-lane ceiling (keeps the shape fixed across a chain)
    const int32_t *gate = nullptr;  // device flag: the launch is skipped when *gate == 0
    bool doubling = false;          // partials (|X_i|^2, Re conj(X_i) t) instead of conj(W_i) t
    bool norm = false;              // partials (|t|^2, 0) of the stored result (CSR/ELL matrices, no W)
    bool rhs = false;               // W is a right-hand side added as coef * W_i (needs bfac32; no ACC)
    int64_t groups = 0;         // if nonzero, the launch uses exactly this many workgroups (a multiple of 8)
    int64_t max_groups = 0;     // if nonzero (a multiple of 8) and `groups` is not set: at most this many
    bool f32_matrix = false;    // complex64 blocks: FP32 matrix values (row sums stay FP64)
    // 4x4-block products on the matrix cores: block-Jacobi solve of each block
    // row's result with these FP32 factors before the epilogue (F_BETA and a
    // right-hand side, `rhs`, allowed; no accumulation or partial sums); the
    // launch fails if the matrix-core kernel cannot run it
    const float *bfac32 = nullptr;
    // Fused second product (real width-3/4/6 ELL A with FP64 values): VOUT =
    // valpha * V X, V's real values in A's ELL slots.
    const double *vval = nullptr;
    void *VOUT = nullptr;
    int64_t ldv = 0;
    Cplx valpha{1, 0};
};
// Launches one fused step; returns the number of workgroups used (the
// partial-sum row count) so callers can size/reduce partials.
int64_t launch_step(kmsd_ctx *c, int dtype, const Step &s);
// Steady-state steps of several chains in shared launches (kmsd_moments_group):
// each chain's step (X, P, OUT and part are taken from buf, part0 per
// launch; groups pinned), its two alternating blocks and its partial sums.
struct GroupChain {
    Step step;
    void *buf[2] = {nullptr, nullptr};
    double *part0 = nullptr;
    int64_t slot_doubles = 0;
};
struct GroupPlan {
    const void *fn = nullptr;  // grouped kernel
    int gy = 0;
    int64_t total = 0;         // workgroups per launch (the sum of the chains' own counts)
    int32_t njobs = 0;
    DBuf table;                // device table of the chains
};
// False when the chains' kernels differ or have no grouped form (then each
// runs alone); else uploads the chain table.
bool group_prepare(kmsd_ctx *c, int dtype, const std::vector<GroupChain> &chains, GroupPlan &plan);
// One shared launch: buf[parity] of every chain is T_k, partial slot `slot`.
void group_launch(kmsd_ctx *c, const GroupPlan &plan, int parity, int slot);
// Number of workgroups launch_step would choose for (n, w, dtype) with
// product operator A (null: no product).  Chains pass it as Step::groups to
// every launch that writes partial rows of one reduction.
// Workgroups wanted for nblk row blocks: one per `per` blocks (default: the
// context's min_blocks), rounded up to a multiple of 8 (the XCD count).
int64_t groups_for_blocks(const kmsd_ctx *c, int64_t nblk, int per = 0);
int64_t step_groups(kmsd_ctx *c, int64_t n, int32_t w, int dtype, int cpl_cap, const kmsd_matrix *A,
                    bool f32_matrix = false);
// Automatic columns per lane of products with A (null: no product) on dtype blocks.
int default_cpl(int dtype, const kmsd_matrix *A);
// Columns-per-lane ceiling supported by a set of blocks (alignment and leading dimensions).
int cpl_cap_for(int dtype, std::initializer_list<std::pair<const void *, int64_t>> blocks);
// out[s*ldo + j] = sum_g part[s][g][j]; mode 0 real part (double), 1 complex.
void launch_reduce(kmsd_ctx *c, const double *part, int64_t groups, int32_t w, int64_t nsteps, double *out,
                   int64_t ldo, int mode);
void launch_scale(kmsd_ctx *c, int dtype, const void *X, int64_t ldx, void *O, int64_t ldo, int64_t n, int32_t w,
                  Cplx coef);
// (R, C) grid of pixels of `words` 8-byte words -> (C, R) grid (out != in).
void launch_pixel_transpose(kmsd_ctx *c, const void *in, void *out, int64_t R, int64_t C, int64_t words);
bool launch_nonfinite(kmsd_ctx *c, int dtype, const void *X, int64_t ldx, int64_t n, int32_t w);
// Conversion between vector precisions with the finiteness check of the
// converted block in the same pass, and no fill: launches alternate between
// flag[1] and flag[2], each clearing the other one for the next launch.
// Returns the flag (nonzero if some converted element is not finite) to read
// before the next call.
const int32_t *launch_convert_check(kmsd_ctx *c, int dtype_in, const void *X, int64_t ldx, int dtype_out, void *O,
                                    int64_t ldo, int64_t n, int32_t w);
void launch_convert(kmsd_ctx *c, int dtype_in, const void *X, int64_t ldx, int dtype_out, void *O, int64_t ldo,
                    int64_t n, int32_t w);

// out = alpha * (A X) [+ out if accumulate]
void spmm(kmsd_ctx *c, int dtype, const kmsd_matrix *A, const void *X, int64_t ldx, void *out, int64_t ldo,
          int32_t w, double alpha_re, double alpha_im, bool accumulate);
void require_finite(kmsd_ctx *c, int dtype, const void *X, int64_t ldx, int64_t n, int32_t w);
// Device copy of S scaled by congruence with diag^-1/2 (the sampler's W).
kmsd_matrix *matrix_congruence_copy(kmsd_ctx *c, const kmsd_matrix *S, const double *diag);
void launch_doubling_finish(kmsd_ctx *c, const double *ab, double *mu, int64_t L, int32_t w);
void launch_sample_finish(kmsd_ctx *c, int dtype, void *out, const void *xi, int64_t ld, int64_t n, int32_t w,
                          const double *diag, double c0);

}  // namespace kmsd

// ---- metric / engine (kmsd_metric.hip, kmsd_engine.hip) --------------------
struct kmsd_metric {
    kmsd_ctx *ctx = nullptr;
    kmsd_matrix *S = nullptr;
    bool identity = false;
    int64_t n = 0;
    double *diag = nullptr;  // device f64[n]
    // block-Jacobi: contiguous blocks [block_ptr[q], block_ptr[q+1]), or
    // cluster blocks whose rows are block_rows[block_ptr[q] .. block_ptr[q+1])
    int64_t nblocks = 0;
    int32_t max_block = 0;           // largest block of at most SMALL_BLOCK rows
    int32_t *block_ptr = nullptr;    // device int32[nblocks+1]
    int64_t *factor_off = nullptr;   // device int64[nblocks] offsets into factors (complex entries)
    double *factors = nullptr;       // device cx<double>[sum m^2]
    int32_t *block_rows = nullptr;   // device int32[n] (null: contiguous blocks)
    int32_t *large = nullptr;        // device int32[nlarge]: blocks of more than SMALL_BLOCK rows
    int64_t nlarge = 0;
    size_t block_bytes = 0;          // device bytes of the block arrays
    std::vector<int64_t> host_block_ptr;
    bool uniform4 = false;           // every block is 4x4 aligned to multiples of 4
    // uniform 4x4 blocks with real factors: float[16] per block, L below the
    // diagonal and 1/L_ii on it (the fused FP32 inner iterations' copy)
    float *factors32 = nullptr;
    struct IcData;
    IcData *ic = nullptr;            // IC(0) preconditioner (takes precedence over blocks)
    bool direct = false;             // direct solve with the one full-size block plus refinement
    int refine_maxiter = 3;
    int64_t refinement_steps = 0;
    size_t bytes = 0;
    bool adaptive = false;
    int check_interval = 32;
    double gap_ratio = 0.1;
    int64_t calls = 0, iterations = 0, refreshes = 0, residual_checks = 0, restarts = 0;
    double max_relative = 0;
    // last report (host)
    std::vector<uint8_t> rep_conv, rep_brk;
    std::vector<int64_t> rep_its;
    std::vector<double> rep_res, rep_rel;
    struct Work;
    Work *work = nullptr;
    // Chebyshev inner solves of the FP64 refinement: S's Gershgorin interval
    // [cheb_lo, cheb_hi]; cheb_state 0 not yet examined, 1 positive and
    // narrow enough, 2 not usable.
    int cheb_state = 0;
    double cheb_lo = 0, cheb_hi = 0;
    // Block-Jacobi Chebyshev inner solves (uniform 4x4 blocks, 4x4-block S):
    // interval [bcheb_lo, bcheb_hi] from the Ritz values of the first PCG
    // inner solve, with margins; bcheb_state 0 not yet measured, 1 usable, 2 not
    int bcheb_state = 0;
    double bcheb_lo = 0, bcheb_hi = 0;
    // complex128 right-hand sides take the complex64 path's FP64 refinement
    // (FP32 inner solves) instead of FP64 PCG (kmsd_metric_set_mixed)
    bool mixed_c128 = false;
    kmsd_matrix *W = nullptr;  // Z S Z for the sampler (built on first use)
    kmsd::DBuf sb1, sb2;       // sampler Clenshaw buffers
    kmsd::DBuf pinner;         // partial sums of kmsd_metric_inner
};

struct kmsd_engine {
    kmsd_ctx *ctx = nullptr;
    kmsd_matrix *H = nullptr;
    kmsd_metric *metric = nullptr;
    int kind[3] = {-1, -1, -1};
    kmsd_matrix *mats[3][3] = {};
    double rtol = 1e-10, atol = 0, hbar = 0.6582119569509067;
    int maxiter = 1000;
    int64_t h_calls = 0, v_calls = 0;
    kmsd::DBuf t0, t1, t2, t3, tmp, tmp2, tmp3, part, wbuf;
    kmsd::DBuf cv1, cv2;  // kmsd_chain_v regrouped blocks (grow-only, like the other scratch)
    // Per axis: the velocity matrix's values in H's ELL slots (fused kmsd_chain_v);
    // state 0 = not built yet, 1 = built, -1 = the velocity does not fit.
    kmsd::DBuf fv[3];
    int fv_state[3] = {0, 0, 0};
};

namespace kmsd {
// Blocks of at most this many rows are solved one thread per (block, column)
// in registers; larger ones one workgroup per (block, column).
constexpr int32_t SMALL_BLOCK = 64;
// Replaces the metric's blocks: block q holds rows off[q] .. off[q+1] of the
// contiguous partition (rows == null) or rows[off[q] .. off[q+1]) (cluster
// blocks); factors concatenated m x m row-major complex.  Validated by the caller.
void metric_upload_blocks(kmsd_metric *m, int64_t nblocks, const int64_t *off, const int64_t *rows,
                          const double *factors);
void metric_release_blocks(kmsd_metric *m);
void metric_release_ic(kmsd_metric *m);
// z = M^-1 r with the IC(0) factor; R, Z (n x w) of dtype, internal leading dimension w.
void ic_apply(kmsd_metric *m, int dtype, const void *R, void *Z, int32_t w, const void *flags);
// Blocks larger than SMALL_BLOCK rows (the small ones are the PCG's own kernel).
void large_blocks_apply(kmsd_metric *m, int dtype, const void *R, void *Z, int32_t w, const void *flags);
// Buffers of one moment chain (kmsd_moments): two chain blocks of
// block_bytes, partial sums of `groups` workgroups for `chunk` orders.
struct MomentScratch {
    void *t1 = nullptr, *t2 = nullptr;
    double *part = nullptr;
    size_t block_bytes = 0, part_bytes = 0;
    int64_t groups = 0, chunk = 0;
};
// Sizes a moment chain of L orders on nb columns needs, with at most
// chunk_max orders per partial reduction.
MomentScratch moment_scratch_need(kmsd_engine *e, const void *B, int64_t L, int64_t nb, int dtype,
                                  int64_t chunk_max);
// kmsd_moments' chain; sc: caller-provided buffers (identity metric only,
// eager launches) or null for the engine's own.  Call inside guard().
void moment_chain(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype, double center,
                  double half, const MomentScratch *sc);
// Device address of kmsd_moments_noalloc's fixed scratch on `device`, looked
// up once per device (kmsd_noalloc.hip); kmsd_ctx_create calls it.
void noalloc_setup(int device);
// Metric solve on device blocks (any dtype); throws Error(KMSD_ENUMERIC) on
// failure with the report stored in the metric.
void metric_solve(kmsd_metric *m, int dtype, const void *B, int64_t ldb, void *Y, int64_t ldy, int32_t nb,
                  double rtol, double atol, int maxiter);
}  // namespace kmsd
