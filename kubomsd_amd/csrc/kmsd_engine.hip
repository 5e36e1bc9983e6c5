// Engine drivers: scaled action, velocity actions, the augmented Chebyshev
// series, stored chains and fused moment chains (reference Engine,
// cpp/native.cpp:488-610).  Everything stays on the device; the host only
// sequences launches.  With an identity metric every Chebyshev order is ONE
// fused launch (CSR product + three-term update + store + moment partials).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kmsd_internal.hpp"

using namespace kmsd;

namespace {

int64_t n_of(const kmsd_engine *e) { return e->H->n; }

// OUT = alpha*(S^-1 H X) + beta*X + gamma*P  (+ moment partials / accumulate)
// `s` carries everything except the operator.
void cheb(kmsd_engine *e, int dtype, Step s) {
    kmsd_ctx *c = e->ctx;
    ++e->h_calls;
    if (e->metric->identity) {
        s.A = e->H;
        launch_step(c, dtype, s);
        return;
    }
    const size_t bytes = size_t(s.n) * size_t(s.w) * elem_bytes(dtype);
    void *hx = e->tmp.ensure(c, bytes);
    void *z = e->tmp2.ensure(c, bytes);
    spmm(c, dtype, e->H, s.X, s.ldx, hx, s.w, s.w, 1.0, 0.0, false);
    metric_solve(e->metric, dtype, hx, s.w, z, s.w, s.w, e->rtol, e->atol, e->maxiter);
    s.A = nullptr;
    s.Z = z;
    s.ldz = s.w;
    launch_step(c, dtype, s);
}

// The fused steps of one chunk of a moment chain captured as a HIP graph and
// replayed for later chunks (the chunk's partial reduction is an ordinary
// launch after each replay, since only its output row changes).  With an
// identity metric every step is one kernel with fixed buffers: after the
// first two orders the three-term recursion alternates two buffers, so an
// even-length chunk starts from the same buffer assignment every time.
// Replays run the same kernels with the same arguments as eager launches
// (bitwise the same results) at a fraction of the per-launch host cost.
// Capture uses the context's private non-blocking stream (the caller's
// stream may be the legacy default stream, which cannot be captured).
struct ChainGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~ChainGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

hipStream_t capture_stream(kmsd_ctx *c) {
    if (!c->capture) HIPCHECK(hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking));
    return c->capture;
}

// out (+)= coef * Y where Y = V_axis-action(X) ; helpers for v_cov
void add_scaled(kmsd_ctx *c, int dtype, const void *Y, int64_t ldy, void *out, int64_t ldo, int64_t n, int32_t w,
                Cplx coef, bool accumulate) {
    Step s;
    s.Z = Y;
    s.ldz = ldy;
    s.alpha = coef;
    s.n = n;
    s.w = w;
    if (accumulate) {
        s.ACC = out;
        s.lda = ldo;
    } else {
        s.OUT = out;
        s.ldo = ldo;
    }
    launch_step(c, dtype, s);
}

// out (+)= sum_a dir[a] * v_a X   (reference Engine::v_cov / v_cart)
// Scratch: e->tmp, tmp2 (solves inside), tmp3.
void vcov(kmsd_engine *e, int dtype, const double *dir, const void *X, int64_t ldx, void *out, int64_t ldo, int32_t w,
          Cplx outer, bool accumulate) {
    kmsd_ctx *c = e->ctx;
    const int64_t n = n_of(e);
    ++e->v_calls;
    bool wrote = accumulate;
    const size_t bytes = size_t(n) * size_t(w) * elem_bytes(dtype);
    for (int ax = 0; ax < 3; ++ax) {
        if (dir[ax] == 0.0) continue;
        const int kind = e->kind[ax];
        kmsd_matrix *const *m = e->mats[ax];
        const Cplx coef{outer.re * dir[ax], outer.im * dir[ax]};
        if (kind == 0 && m[0]) {
            spmm(c, dtype, m[0], X, ldx, out, ldo, w, coef.re, coef.im, wrote);
            wrote = true;
            continue;
        }
        if (!((kind == 1 && m[0]) || (kind == 2 && m[0] && m[1] && m[2])))
            throw Error(KMSD_EINVAL, "missing Cartesian velocity");
        void *hx = e->tmp.ensure(c, bytes);
        void *sx = e->tmp2.ensure(c, bytes);
        void *acc = e->tmp3.ensure(c, bytes);
        const double ih = 1.0 / e->hbar;  // (i/hbar) folded into the final coefficient
        if (kind == 1) {
            // (i/hbar)(H S^-1 X B - X S^-1 H B)
            spmm(c, dtype, m[0], X, ldx, hx, w, w, 1.0, 0.0, false);
            metric_solve(e->metric, dtype, hx, w, sx, w, w, e->rtol, e->atol, e->maxiter);
            spmm(c, dtype, e->H, sx, w, acc, w, w, 1.0, 0.0, false);
            spmm(c, dtype, e->H, X, ldx, hx, w, w, 1.0, 0.0, false);
            metric_solve(e->metric, dtype, hx, w, sx, w, w, e->rtol, e->atol, e->maxiter);
            spmm(c, dtype, m[0], sx, w, acc, w, w, -1.0, 0.0, true);
        } else {
            // (i/hbar)(bond B - xbar S^-1 H B + H S^-1 delta B)
            spmm(c, dtype, m[0], X, ldx, acc, w, w, 1.0, 0.0, false);
            spmm(c, dtype, e->H, X, ldx, hx, w, w, 1.0, 0.0, false);
            metric_solve(e->metric, dtype, hx, w, sx, w, w, e->rtol, e->atol, e->maxiter);
            spmm(c, dtype, m[1], sx, w, acc, w, w, -1.0, 0.0, true);
            spmm(c, dtype, m[2], X, ldx, hx, w, w, 1.0, 0.0, false);
            metric_solve(e->metric, dtype, hx, w, sx, w, w, e->rtol, e->atol, e->maxiter);
            spmm(c, dtype, e->H, sx, w, acc, w, w, 1.0, 0.0, true);
        }
        // coef * (i/hbar) * acc
        add_scaled(c, dtype, acc, w, out, ldo, n, w, Cplx{-coef.im * ih, coef.re * ih}, wrote);
        wrote = true;
    }
    if (!wrote) launch_scale(c, dtype, X, ldx, out, ldo, n, w, Cplx{0, 0});
}

// Augmented action (reference Engine::augmented, cpp/native.cpp:540-554), in
// three-term form: OUT = f*(A X) - f*c/a X [+ gamma P] where block d+1 also
// receives f*(-i hbar/a) S^-1 v_d X0.  f = 1 for the first order, 2 after.
void augmented_step(kmsd_engine *e, int dtype, const void *X, const void *P, void *OUT, int64_t width, int32_t nb,
                    double f, double center, double half, const double *dirs, int64_t ndirs) {
    kmsd_ctx *c = e->ctx;
    const int64_t n = n_of(e);
    Step s;
    s.X = X;
    s.ldx = width;
    s.OUT = OUT;
    s.ldo = width;
    s.alpha = Cplx{f / half, 0};
    s.beta = Cplx{-f * center / half, 0};
    s.use_beta = true;
    if (P) {
        s.P = P;
        s.ldp = width;
        s.gamma = -1.0;
        s.use_p = true;
    }
    s.n = n;
    s.w = int32_t(width);
    cheb(e, dtype, s);
    const Cplx src{0, -f * e->hbar / half};
    for (int64_t d = 0; d < ndirs; ++d) {
        void *blk = static_cast<char *>(OUT) + size_t(d + 1) * nb * elem_bytes(dtype);
        const double *dir = dirs + 3 * d;
        bool simple = e->metric->identity;
        for (int ax = 0; ax < 3; ++ax)
            if (dir[ax] != 0.0 && e->kind[ax] != 0) simple = false;
        if (simple) {
            // S = I and explicit velocities: accumulate V X0 straight into the block.
            vcov(e, dtype, dir, X, width, blk, width, nb, src, true);
            continue;
        }
        const size_t bytes = size_t(n) * size_t(nb) * elem_bytes(dtype);
        void *v = e->t2.ensure(c, bytes);
        void *sv = e->t3.ensure(c, bytes);
        vcov(e, dtype, dir, X, width, v, nb, nb, Cplx{1, 0}, false);
        metric_solve(e->metric, dtype, v, nb, sv, nb, nb, e->rtol, e->atol, e->maxiter);
        add_scaled(c, dtype, sv, nb, blk, width, n, nb, src, true);
    }
}

void check_half(double half) {
    if (!(half > 0)) throw Error(KMSD_EINVAL, "spectral half-width must be positive");
}

}  // namespace

extern "C" int kmsd_engine_create(kmsd_ctx *c, kmsd_matrix *H, kmsd_metric *metric, const int *kinds,
                                  kmsd_matrix *const *mats, double rtol, double atol, int maxiter, double hbar,
                                  kmsd_engine **out) {
    if (!c || !out) return KMSD_EINVAL;
    *out = nullptr;
    kmsd_engine *e = new kmsd_engine();
    const int rc = guard(c, [&] {
        if (!H || !metric || !kinds || H->n != metric->n) throw Error(KMSD_EINVAL, "invalid engine operators");
        if (H->fmt != KMSD_CSR && H->fmt != KMSD_BSR4) throw Error(KMSD_EINVAL, "engine H must be CSR or BSR4");
        e->ctx = c;
        e->H = H;
        e->metric = metric;
        for (int a = 0; a < 3; ++a) {
            e->kind[a] = kinds[a];
            const int cnt = kinds[a] == 2 ? 3 : (kinds[a] == 0 || kinds[a] == 1 ? 1 : 0);
            if (kinds[a] < -1 || kinds[a] > 2) throw Error(KMSD_EINVAL, "invalid velocity kind");
            for (int k = 0; k < cnt; ++k) {
                kmsd_matrix *m = mats ? mats[3 * a + k] : nullptr;
                if (!m || m->n != H->n) throw Error(KMSD_EINVAL, "invalid engine operators");
                e->mats[a][k] = m;
            }
        }
        e->rtol = rtol;
        e->atol = atol;
        e->maxiter = maxiter;
        e->hbar = hbar;
        ctx_adopt(c);
    });
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return KMSD_OK;
}

extern "C" int kmsd_engine_destroy(kmsd_engine *e) {
    if (!e) return KMSD_OK;
    kmsd_ctx *c = e->ctx;
    bool free_ctx = false;
    {
        std::lock_guard<std::recursive_mutex> lock(c->mu);
        delete e;
        free_ctx = ctx_orphan(c);
    }
    if (free_ctx) ctx_free(c);
    return KMSD_OK;
}

extern "C" int kmsd_engine_set_controls(kmsd_engine *e, double rtol, double atol, int maxiter) {
    if (!e) return KMSD_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(e->ctx->mu);
    e->rtol = rtol;
    e->atol = atol;
    e->maxiter = maxiter;
    return KMSD_OK;
}

extern "C" int kmsd_engine_counters(const kmsd_engine *e, int64_t *h, int64_t *v) {
    if (!e) return KMSD_EINVAL;
    if (h) *h = e->h_calls;
    if (v) *v = e->v_calls;
    return KMSD_OK;
}

extern "C" int kmsd_engine_last_report(const kmsd_engine *e, kmsd_report *r, int64_t *nb) {
    if (!e) return KMSD_EINVAL;
    const kmsd_metric *m = e->metric;
    const size_t b = m->rep_conv.size();
    if (nb) *nb = int64_t(b);
    if (r) {
        if (r->converged) std::copy(m->rep_conv.begin(), m->rep_conv.end(), r->converged);
        if (r->breakdown) std::copy(m->rep_brk.begin(), m->rep_brk.end(), r->breakdown);
        if (r->iterations) std::copy(m->rep_its.begin(), m->rep_its.end(), r->iterations);
        if (r->true_residual) std::copy(m->rep_res.begin(), m->rep_res.end(), r->true_residual);
        if (r->relative_residual) std::copy(m->rep_rel.begin(), m->rep_rel.end(), r->relative_residual);
    }
    return KMSD_OK;
}

extern "C" int kmsd_h_action(kmsd_engine *e, const void *B, void *out, int64_t nb, int dtype) {
    if (!e) return KMSD_EINVAL;
    kmsd_ctx *c = e->ctx;
    return guard(c, [&] {
        check_block_args(B, out, nb, dtype);
        const int64_t n = n_of(e);
        require_finite(c, dtype, B, nb, n, int32_t(nb));
        ++e->h_calls;
        if (e->metric->identity) {
            spmm(c, dtype, e->H, B, nb, out, nb, int32_t(nb), 1.0, 0.0, false);
        } else {
            void *hx = e->tmp.ensure(c, size_t(n) * nb * elem_bytes(dtype));
            spmm(c, dtype, e->H, B, nb, hx, nb, int32_t(nb), 1.0, 0.0, false);
            metric_solve(e->metric, dtype, hx, nb, out, nb, int32_t(nb), e->rtol, e->atol, e->maxiter);
        }
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int kmsd_v_cov(kmsd_engine *e, const double *dir, const void *B, void *out, int64_t nb, int dtype) {
    if (!e) return KMSD_EINVAL;
    kmsd_ctx *c = e->ctx;
    return guard(c, [&] {
        check_block_args(B, out, nb, dtype);
        if (!dir) throw Error(KMSD_EINVAL, "direction must have three components");
        require_finite(c, dtype, B, nb, n_of(e), int32_t(nb));
        vcov(e, dtype, dir, B, nb, out, nb, int32_t(nb), Cplx{1, 0}, false);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int kmsd_series(kmsd_engine *e, const void *B, void *out, int64_t width, int dtype, const double *coeff,
                           int64_t ncoeff, double center, double half, const double *dirs, int64_t ndirs) {
    if (!e) return KMSD_EINVAL;
    kmsd_ctx *c = e->ctx;
    return guard(c, [&] {
        check_block_args(B, out, width, dtype);
        if (ncoeff < 1 || !coeff) throw Error(KMSD_EINVAL, "empty Chebyshev coefficients");
        const int64_t n = n_of(e);
        require_finite(c, dtype, B, width, n, int32_t(width));
        launch_scale(c, dtype, B, width, out, width, n, int32_t(width), Cplx{coeff[0], coeff[1]});
        if (ncoeff == 1) {
            HIPCHECK(hipStreamSynchronize(c->stream));
            return;
        }
        check_half(half);
        if (ndirs < 0 || (ndirs && (!dirs || width % (ndirs + 1))))
            throw Error(KMSD_EINVAL, "invalid augmented block columns");
        const int32_t nb = int32_t(width / (ndirs + 1));
        const size_t bytes = size_t(n) * size_t(width) * elem_bytes(dtype);
        void *u0 = e->t0.ensure(c, bytes);
        void *u1 = e->t1.ensure(c, bytes);
        augmented_step(e, dtype, B, nullptr, u0, width, nb, 1.0, center, half, dirs, ndirs);
        add_scaled(c, dtype, u0, width, out, width, n, int32_t(width), Cplx{coeff[2], coeff[3]}, true);
        const void *prev = B;
        void *cur = u0;
        void *spare = u1;
        for (int64_t k = 2; k < ncoeff; ++k) {
            void *nxt = (prev == B) ? spare : const_cast<void *>(prev);
            augmented_step(e, dtype, cur, prev, nxt, width, nb, 2.0, center, half, dirs, ndirs);
            add_scaled(c, dtype, nxt, width, out, width, n, int32_t(width), Cplx{coeff[2 * k], coeff[2 * k + 1]}, true);
            prev = cur;
            cur = nxt;
        }
        if (launch_nonfinite(c, dtype, out, width, n, int32_t(width)))
            throw Error(KMSD_ENUMERIC, "nonfinite Chebyshev propagation; check bounds and solves");
    });
}

namespace {

// out[k] = T_k(H~) B for k < L, (L, N, nb) blocks.
void chain_into(kmsd_engine *e, int dtype, const void *B, void *out, int64_t L, int64_t nb, double center,
                double half) {
    kmsd_ctx *c = e->ctx;
    {
        const int64_t n = n_of(e);
        const size_t stride = size_t(n) * size_t(nb) * elem_bytes(dtype);
        char *o = static_cast<char *>(out);
        launch_scale(c, dtype, B, nb, o, nb, n, int32_t(nb), Cplx{1, 0});
        if (L > 1) {
            check_half(half);
            Step s;
            s.X = B;
            s.ldx = nb;
            s.OUT = o + stride;
            s.ldo = nb;
            s.alpha = Cplx{1 / half, 0};
            s.beta = Cplx{-center / half, 0};
            s.use_beta = true;
            s.n = n;
            s.w = int32_t(nb);
            cheb(e, dtype, s);
        }
        for (int64_t k = 2; k < L; ++k) {
            Step s;
            s.X = o + size_t(k - 1) * stride;
            s.ldx = nb;
            s.P = o + size_t(k - 2) * stride;
            s.ldp = nb;
            s.OUT = o + size_t(k) * stride;
            s.ldo = nb;
            s.alpha = Cplx{2 / half, 0};
            s.beta = Cplx{-2 * center / half, 0};
            s.gamma = -1.0;
            s.use_beta = s.use_p = true;
            s.n = n;
            s.w = int32_t(nb);
            cheb(e, dtype, s);
        }
    }
}

}  // namespace

extern "C" int kmsd_chain(kmsd_engine *e, const void *B, void *out, int64_t L, int64_t nb, int dtype, double center,
                          double half) {
    if (!e) return KMSD_EINVAL;
    kmsd_ctx *c = e->ctx;
    return guard(c, [&] {
        check_block_args(B, out, nb, dtype);
        if (L < 1) throw Error(KMSD_EINVAL, "Chebyshev chain must be nonempty");
        require_finite(c, dtype, B, nb, n_of(e), int32_t(nb));
        chain_into(e, dtype, B, out, L, nb, center, half);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

namespace {

// V's values in the ELL slots of H (zero where V has no entry), one thread
// per row: each nonzero entry of V's row is matched, in column order, to the
// next slot of H's row with the same column.  *fail is set when an entry has
// no such slot (V's pattern is not inside H's, or its column order differs).
__global__ __launch_bounds__(256) void slot_values_kernel(const int32_t *hidx, int64_t n, int32_t w,
                                                          const int32_t *vptr, int32_t vell, const int32_t *vidx,
                                                          const double *vval, double *slots, int32_t *fail) {
    for (int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x; r < n; r += int64_t(gridDim.x) * 256) {
        const int64_t k0 = vell ? r * vell : vptr[r], k1 = vell ? (r + 1) * vell : vptr[r + 1];
        double *row = slots + r * w;
        const int32_t *hrow = hidx + r * w;
        for (int32_t s = 0; s < w; ++s) row[s] = 0.0;
        int32_t pos = 0;
        for (int64_t k = k0; k < k1; ++k) {
            const double v = vval[k];
            if (v == 0.0) continue;  // ELL padding or a stored zero: adds nothing to a finite row sum
            const int32_t col = vidx[k];
            while (pos < w && hrow[pos] != col) ++pos;
            if (pos == w) {
                atomicOr(fail, 1);
                break;
            }
            row[pos++] = v;
        }
    }
}

// Builds the slot values of V on the device; false when V does not fit H's
// slots.  Only the 4-byte failure flag comes back to the host.
bool build_slot_values(kmsd_engine *e, const kmsd_matrix *H, const kmsd_matrix *V, kmsd::DBuf &out) {
    kmsd_ctx *c = e->ctx;
    const int64_t n = H->n;
    const int32_t w = H->ell;
    double *d = static_cast<double *>(out.ensure(c, size_t(n) * size_t(w) * sizeof(double)));
    int32_t zero = 0, fail = 0;
    copy_h2d(c, c->flag, &zero, sizeof(zero));
    const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
    hipLaunchKernelGGL(slot_values_kernel, dim3(g), dim3(256), 0, c->stream, H->idx, n, w, V->ptr, V->ell, V->idx,
                       static_cast<const double *>(V->val64), d, c->flag);
    HIPCHECK(hipGetLastError());
    read_status(c, {{&fail, c->flag, sizeof(fail)}});
    if (fail) out.release();
    return fail == 0;
}

// The slot values of the one velocity matrix that makes up direction `dir`
// when kmsd_chain_v can form V T_k inside the chain's steps, with the scale
// vcov's product applies (dir times i for imaginary values, as launch_step
// rotates alpha); null when it cannot: a nonidentity metric (each order is a
// solve, not one launch), H not a real width-3/4/6 ELL matrix, FP32 matrix
// arithmetic, several axes, a commutator or composite velocity, a complex or
// BSR velocity, a nonzero velocity diagonal, or entries outside H's slots.
const double *fused_velocity(kmsd_engine *e, const double *dir, Cplx *valpha) {
    const kmsd_ctx *c = e->ctx;
    const kmsd_matrix *H = e->H;
    if (!c->fuse_velocity || !c->ell_kernel || c->fp32_math || !e->metric->identity) return nullptr;
    if (H->fmt != KMSD_CSR || H->vkind == VK_COMPLEX || (H->ell != 3 && H->ell != 4 && H->ell != 6)) return nullptr;
    int ax = -1;
    for (int a = 0; a < 3; ++a) {
        if (dir[a] == 0.0) continue;
        if (ax >= 0) return nullptr;
        ax = a;
    }
    if (ax < 0 || e->kind[ax] != 0 || !e->mats[ax][0]) return nullptr;
    const kmsd_matrix *V = e->mats[ax][0];
    if (V->fmt != KMSD_CSR || V->vkind == VK_COMPLEX || V->has_diag || V->n != H->n) return nullptr;
    if (e->fv_state[ax] == 0) e->fv_state[ax] = build_slot_values(e, H, V, e->fv[ax]) ? 1 : -1;
    if (e->fv_state[ax] < 0) return nullptr;
    const Cplx coef{1.0 * dir[ax], 0.0 * dir[ax]};
    *valpha = V->vkind == VK_IMAG ? Cplx{-coef.im, coef.re} : coef;
    return static_cast<const double *>(e->fv[ax].p);
}

// kmsd_chain_v with V T_(k-1) formed by the step that gathers T_(k-1) to make
// T_k: the chain lives in two engine buffers (T_k overwrites T_(k-2) in
// place), `out` receives only V T_k, and the last order's V T_(L-1) is one
// ordinary velocity product.  Same arithmetic as chain_into + vcov.
void fused_chain_v(kmsd_engine *e, int dtype, const double *dir, const double *vval, Cplx valpha, const void *B,
                   void *out, int64_t L, int64_t nb, double center, double half) {
    kmsd_ctx *c = e->ctx;
    const int64_t n = n_of(e);
    const size_t stride = size_t(n) * size_t(nb) * elem_bytes(dtype);
    char *o = static_cast<char *>(out);
    const void *prev = nullptr, *cur = B;
    if (L > 1) check_half(half);
    void *ta = e->t0.ensure(c, stride);
    void *tb = e->t1.ensure(c, stride);
    for (int64_t k = 1; k < L; ++k) {
        void *nxt = k == 1 ? ta : (prev == B ? tb : const_cast<void *>(prev));
        Step s;
        s.X = cur;
        s.ldx = nb;
        s.OUT = nxt;
        s.ldo = nb;
        if (k == 1) {
            s.alpha = Cplx{1 / half, 0};
            s.beta = Cplx{-center / half, 0};
            s.use_beta = true;
        } else {
            s.P = prev;
            s.ldp = nb;
            s.alpha = Cplx{2 / half, 0};
            s.beta = Cplx{-2 * center / half, 0};
            s.gamma = -1.0;
            s.use_beta = s.use_p = true;
        }
        s.n = n;
        s.w = int32_t(nb);
        s.vval = vval;
        s.VOUT = o + size_t(k - 1) * stride;
        s.ldv = nb;
        s.valpha = valpha;
        cheb(e, dtype, s);
        prev = cur;
        cur = nxt;
    }
    vcov(e, dtype, dir, cur, nb, o + size_t(L - 1) * stride, nb, int32_t(nb), Cplx{1, 0}, false);
}

}  // namespace

extern "C" int kmsd_chain_v(kmsd_engine *e, const double *dir, const void *B, void *out, int64_t L, int64_t nb,
                            int dtype, double center, double half) {
    if (!e) return KMSD_EINVAL;
    kmsd_ctx *c = e->ctx;
    return guard(c, [&] {
        check_block_args(B, out, nb, dtype);
        if (!dir || !std::isfinite(dir[0]) || !std::isfinite(dir[1]) || !std::isfinite(dir[2]))
            throw Error(KMSD_EINVAL, "direction must have three finite components");
        if (L < 1) throw Error(KMSD_EINVAL, "Chebyshev chain must be nonempty");
        if (L * nb >= (int64_t(1) << 31)) throw Error(KMSD_EINVAL, "chain too wide for one velocity product");
        const int64_t n = n_of(e);
        require_finite(c, dtype, B, nb, n, int32_t(nb));
        Cplx valpha;
        if (const double *vval = fused_velocity(e, dir, &valpha)) {
            fused_chain_v(e, dtype, dir, vval, valpha, B, out, L, nb, center, half);
            HIPCHECK(hipStreamSynchronize(c->stream));
            return;
        }
        chain_into(e, dtype, B, out, L, nb, center, half);
        // One velocity product over all orders: (L, N, nb) -> (N, L*nb) rows,
        // v_cov on L*nb columns, and back.
        // (grow-only engine scratch: large device allocations cost more than
        // the regrouping itself)
        const size_t total = size_t(L) * size_t(n) * size_t(nb) * elem_bytes(dtype);
        const int64_t words = nb * int64_t(elem_bytes(dtype) / 8);
        void *flat = e->cv1.ensure(c, total);
        void *vflat = e->cv2.ensure(c, total);
        launch_pixel_transpose(c, out, flat, L, n, words);
        vcov(e, dtype, dir, flat, L * nb, vflat, L * nb, int32_t(L * nb), Cplx{1, 0}, false);
        launch_pixel_transpose(c, vflat, out, n, L, words);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

namespace kmsd {

MomentScratch moment_scratch_need(kmsd_engine *e, const void *B, int64_t L, int64_t nb, int dtype,
                                  int64_t chunk_max) {
    kmsd_ctx *c = e->ctx;
    MomentScratch need;
    const int64_t n = n_of(e);
    need.block_bytes = size_t(n) * size_t(nb) * elem_bytes(dtype);
    // columns-per-lane ceiling of the chain: B's alignment and the width (the
    // scratch blocks are 256-byte aligned, so only their width counts)
    const int cap = cpl_cap_for(dtype, {{B, nb}});
    need.groups = step_groups(c, n, int32_t(nb), dtype, cap, e->metric->identity ? e->H : nullptr);
    need.chunk = std::min<int64_t>(std::min<int64_t>(c->moment_chunk, chunk_max), L);
    need.part_bytes = sizeof(double) * size_t(need.groups) * size_t(nb) * 2 * size_t(need.chunk);
    return need;
}

// The moment chain of kmsd_moments.  Buffers: the engine's own grow-only
// blocks (sc == null), or the caller's scratch (two chain blocks and the
// partial sums at the given addresses; identity metric only, eager launches).
void moment_chain(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype, double center,
                  double half, const MomentScratch *sc) {
    kmsd_ctx *c = e->ctx;
    {
        check_block_args(B, mu, nb, dtype);
        if (L < 1) throw Error(KMSD_EINVAL, "moment count must be positive");
        const int64_t n = n_of(e);
        const int32_t w = int32_t(nb);
        require_finite(c, dtype, B, nb, n, w);
        const size_t bytes = size_t(n) * size_t(nb) * elem_bytes(dtype);
        const void *W = B;
        if (!e->metric->identity) {
            if (sc) throw Error(KMSD_EINVAL, "caller scratch needs an identity metric");
            void *sb = e->wbuf.ensure(c, bytes);
            spmm(c, dtype, e->metric->S, B, nb, sb, nb, w, 1.0, 0.0, false);
            W = sb;
        }
        void *t1 = sc ? sc->t1 : e->t0.ensure(c, bytes);
        void *spare = sc ? sc->t2 : e->t1.ensure(c, bytes);
        const int cap = cpl_cap_for(dtype, {{B, nb}, {W, nb}, {t1, nb}, {spare, nb}});
        const int64_t G = step_groups(c, n, w, dtype, cap, e->metric->identity ? e->H : nullptr);
        const int64_t chunk = sc ? sc->chunk : std::min<int64_t>(c->moment_chunk, L);
        const size_t slot_doubles = size_t(G) * size_t(w) * 2;
        if (sc && (G != sc->groups || sizeof(double) * slot_doubles * size_t(chunk) > sc->part_bytes))
            throw Error(KMSD_EDEVICE, "internal error: moment scratch layout");
        double *part = sc ? sc->part
                          : static_cast<double *>(e->part.ensure(c, sizeof(double) * slot_doubles * size_t(chunk)));
        int64_t slot = 0, base = 0;
        auto flush = [&] {
            launch_reduce(c, part, G, w, slot, mu + base * nb, nb, 0);
            base += slot;
            slot = 0;
        };
        auto next_part = [&] {
            if (slot == chunk) flush();
            return part + size_t(slot++) * slot_doubles;
        };
        {
            Step s;
            s.Z = B;
            s.ldz = nb;
            s.W = W;
            s.ldw = nb;
            s.n = n;
            s.w = w;
            s.cpl_cap = cap;
            s.groups = G;
            s.part = next_part();
            launch_step(c, dtype, s);
        }
        if (L > 1) {
            check_half(half);
            Step s;
            s.X = B;
            s.ldx = nb;
            s.OUT = t1;
            s.ldo = nb;
            s.alpha = Cplx{1 / half, 0};
            s.beta = Cplx{-center / half, 0};
            s.use_beta = true;
            s.W = W;
            s.ldw = nb;
            s.n = n;
            s.w = w;
            s.cpl_cap = cap;
            s.groups = G;
            s.part = next_part();
            cheb(e, dtype, s);
            const void *prev = B;
            void *cur = t1;
            auto step = [&](int64_t) {
                void *o = (prev == B) ? spare : const_cast<void *>(prev);
                Step q;
                q.X = cur;
                q.ldx = nb;
                q.P = prev;
                q.ldp = nb;
                q.OUT = o;
                q.ldo = nb;
                q.alpha = Cplx{2 / half, 0};
                q.beta = Cplx{-2 * center / half, 0};
                q.gamma = -1.0;
                q.use_beta = q.use_p = true;
                q.W = W;
                q.ldw = nb;
                q.n = n;
                q.w = w;
                q.cpl_cap = cap;
                q.groups = G;
                q.part = next_part();
                cheb(e, dtype, q);
                prev = cur;
                cur = o;
            };
            // Graph replay of whole chunks: identity metric (one kernel per
            // order, no host round trips), even chunk length, chunk-aligned.
            // A caller's scratch runs eager launches (instantiating a graph
            // allocates).
            const bool graphs = !sc && c->graphs && e->metric->identity && chunk % 2 == 0 && chunk >= 2;
            ChainGraph cg;
            const void *g_prev = nullptr;
            const void *g_cur = nullptr;
            int64_t k = 2;
            while (k < L) {
                if (graphs && k >= 4 && k + chunk <= L && (slot == chunk || slot == 0)) {
                    if (slot == chunk) flush();  // eager partials before the graph
                    if (base != k) throw Error(KMSD_EDEVICE, "internal error: chain chunk misaligned");
                    if (!cg.exec) {
                        g_prev = prev;
                        g_cur = cur;
                        const uint64_t h0 = e->h_calls;
                        hipStream_t user = c->stream;
                        c->stream = capture_stream(c);
                        HIPCHECK_RESTORE(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), c, user);
                        try {
                            for (int64_t j = 0; j < chunk; ++j) step(k + j);  // partial slots 0 .. chunk-1
                        } catch (...) {
                            hipGraph_t g = nullptr;
                            (void)hipStreamEndCapture(c->stream, &g);
                            if (g) (void)hipGraphDestroy(g);
                            c->stream = user;
                            throw;
                        }
                        const hipError_t ec = hipStreamEndCapture(c->stream, &cg.graph);
                        c->stream = user;
                        HIPCHECK(ec);
                        HIPCHECK(hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0));
                        e->h_calls = h0;
                    }
                    if (prev != g_prev || cur != g_cur)
                        throw Error(KMSD_EDEVICE, "internal error: chain buffers do not repeat per chunk");
                    HIPCHECK(hipGraphLaunch(cg.exec, c->stream));
                    launch_reduce(c, part, G, w, chunk, mu + base * nb, nb, 0);
                    e->h_calls += uint64_t(chunk);
                    base += chunk;
                    slot = 0;
                    k += chunk;
                    continue;
                }
                step(k);
                ++k;
            }
        }
        flush();
        HIPCHECK(hipStreamSynchronize(c->stream));
    }
}

}  // namespace kmsd

extern "C" int kmsd_moments(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype,
                            double center, double half) {
    if (!e) return KMSD_EINVAL;
    return guard(e->ctx, [&] { moment_chain(e, B, mu, L, nb, dtype, center, half, nullptr); });
}

namespace {

// The moment chains of several engines (identity metrics, one dtype, width
// and length) with their steady-state orders in shared launches.  Each
// chain keeps the steps, workgroup count, partial slots and reductions of
// its own kmsd_moments call, so its moments are the same bits.
void moment_chain_group(const std::vector<kmsd_engine *> &es, const void *const *Bs, double *const *mus, int64_t L,
                        int64_t nb, int dtype, const double *centers, const double *halves) {
    kmsd_ctx *c = es[0]->ctx;
    const int32_t w = int32_t(nb);
    const size_t count = es.size();
    struct Job {
        kmsd_engine *e;
        const void *B;
        double *mu;
        void *t1, *spare;
        double *part;
        int cap;
        int64_t G;
        size_t slot_doubles;
        double c, a;
    };
    std::vector<Job> jobs(count);
    const int64_t chunk = std::min<int64_t>(c->moment_chunk, L);
    for (size_t q = 0; q < count; ++q) {
        Job &j = jobs[q];
        j.e = es[q];
        j.B = Bs[q];
        j.mu = mus[q];
        j.c = centers[q];
        j.a = halves[q];
        check_block_args(j.B, j.mu, nb, dtype);
        if (L > 1) check_half(j.a);
        const int64_t n = n_of(j.e);
        require_finite(c, dtype, j.B, nb, n, w);
        const size_t bytes = size_t(n) * size_t(nb) * elem_bytes(dtype);
        j.t1 = j.e->t0.ensure(c, bytes);
        j.spare = j.e->t1.ensure(c, bytes);
        j.cap = cpl_cap_for(dtype, {{j.B, nb}, {j.t1, nb}, {j.spare, nb}});
        j.G = step_groups(c, n, w, dtype, j.cap, j.e->H);
        j.slot_doubles = size_t(j.G) * size_t(w) * 2;
        j.part = static_cast<double *>(j.e->part.ensure(c, sizeof(double) * j.slot_doubles * size_t(chunk)));
    }
    int64_t slot = 0, base = 0;
    auto flush = [&] {
        for (Job &j : jobs) launch_reduce(c, j.part, j.G, w, slot, j.mu + base * nb, nb, 0);
        base += slot;
        slot = 0;
    };
    auto step_of = [&](const Job &j, int64_t k, const void *X, const void *P, void *O) {
        Step s;
        s.A = j.e->H;
        s.W = j.B;
        s.ldw = nb;
        s.n = n_of(j.e);
        s.w = w;
        s.cpl_cap = j.cap;
        s.groups = j.G;
        if (k == 0) {  // mu_0 = <B, B>
            s.A = nullptr;
            s.Z = j.B;
            s.ldz = nb;
        } else if (k == 1) {  // T_1 = (H B - c B) / a
            s.X = j.B;
            s.ldx = nb;
            s.OUT = O;
            s.ldo = nb;
            s.alpha = Cplx{1 / j.a, 0};
            s.beta = Cplx{-j.c / j.a, 0};
            s.use_beta = true;
        } else {  // T_(k+1) = 2 (H T_k - c T_k) / a - T_(k-1)
            s.X = X;
            s.ldx = nb;
            s.P = P;
            s.ldp = nb;
            s.OUT = O;
            s.ldo = nb;
            s.alpha = Cplx{2 / j.a, 0};
            s.beta = Cplx{-2 * j.c / j.a, 0};
            s.gamma = -1.0;
            s.use_beta = s.use_p = true;
        }
        return s;
    };
    // orders 0, 1 and 2 chain by chain (orders 0 and 1 read B; order 2 writes
    // the second block), then T_k alternates between t1 and spare
    for (int64_t k = 0; k < std::min<int64_t>(L, 3); ++k) {
        if (slot == chunk) flush();
        for (Job &j : jobs) {
            Step s = k == 0 ? step_of(j, 0, nullptr, nullptr, nullptr)
                            : k == 1 ? step_of(j, 1, nullptr, nullptr, j.t1) : step_of(j, 2, j.t1, j.B, j.spare);
            s.part = j.part + size_t(slot) * j.slot_doubles;
            launch_step(c, dtype, s);
            if (k) ++j.e->h_calls;
        }
        ++slot;
    }
    if (L > 3) {
        std::vector<GroupChain> chains(count);
        for (size_t q = 0; q < count; ++q) {
            const Job &j = jobs[q];
            chains[q].step = step_of(j, 3, j.spare, j.t1, j.t1);
            chains[q].step.part = j.part;
            chains[q].buf[0] = j.t1;
            chains[q].buf[1] = j.spare;
            chains[q].part0 = j.part;
            chains[q].slot_doubles = int64_t(j.slot_doubles);
        }
        GroupPlan plan;
        const bool shared = count > 1 && group_prepare(c, dtype, chains, plan);
        for (int64_t k = 3; k < L; ++k) {
            if (slot == chunk) flush();
            const int parity = int(k & 1);  // buf[parity] holds T_k
            if (shared) {
                group_launch(c, plan, parity, int(slot));
            } else {
                for (size_t q = 0; q < count; ++q) {
                    const Job &j = jobs[q];
                    void *cur = parity ? j.spare : j.t1, *other = parity ? j.t1 : j.spare;
                    Step s = step_of(j, k, cur, other, other);
                    s.part = j.part + size_t(slot) * j.slot_doubles;
                    launch_step(c, dtype, s);
                }
            }
            for (Job &j : jobs) ++j.e->h_calls;
            ++slot;
        }
        flush();
        HIPCHECK(hipStreamSynchronize(c->stream));  // the plan's table is freed on return
        return;
    }
    flush();
    HIPCHECK(hipStreamSynchronize(c->stream));
}

}  // namespace

extern "C" int kmsd_moments_group(int64_t count, kmsd_engine *const *engines, const void *const *B, double *const *mu,
                                  int64_t L, int64_t nb, int dtype, const double *center, const double *half) {
    if (count < 1 || !engines || !B || !mu || !center || !half || !engines[0]) return KMSD_EINVAL;
    kmsd_ctx *c = engines[0]->ctx;
    return guard(c, [&] {
        if (L < 1) throw Error(KMSD_EINVAL, "moment count must be positive");
        std::vector<kmsd_engine *> es;
        for (int64_t q = 0; q < count; ++q) {
            kmsd_engine *e = engines[q];
            if (!e || e->ctx != c) throw Error(KMSD_EINVAL, "grouped moment chains need engines of one context");
            if (!e->metric->identity) throw Error(KMSD_EINVAL, "grouped moment chains need identity metrics");
            for (kmsd_engine *o : es)
                if (o == e) throw Error(KMSD_EINVAL, "grouped moment chains need distinct engines");
            es.push_back(e);
        }
        moment_chain_group(es, B, mu, L, nb, dtype, center, half);
    });
}

extern "C" int kmsd_moments_doubling(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype,
                                     double center, double half) {
    if (!e) return KMSD_EINVAL;
    kmsd_ctx *c = e->ctx;
    return guard(c, [&] {
        check_block_args(B, mu, nb, dtype);
        if (L < 1) throw Error(KMSD_EINVAL, "moment count must be positive");
        if (!e->metric->identity) throw Error(KMSD_EINVAL, "moment doubling requires an identity metric");
        check_half(half);
        const int64_t n = n_of(e);
        const int32_t w = int32_t(nb);
        require_finite(c, dtype, B, nb, n, w);
        const size_t bytes = size_t(n) * size_t(nb) * elem_bytes(dtype);
        void *t1 = e->t0.ensure(c, bytes);
        void *spare = e->t1.ensure(c, bytes);
        const int cap = cpl_cap_for(dtype, {{B, nb}, {t1, nb}, {spare, nb}});
        const int64_t G = step_groups(c, n, w, dtype, cap, e->H);
        const int64_t steps = (L - 1) / 2 + 1;  // k = 0 .. floor((L-1)/2)
        const int64_t chunk = std::min<int64_t>(c->moment_chunk, steps);
        const size_t slot_doubles = size_t(G) * size_t(w) * 2;
        double *part = static_cast<double *>(e->part.ensure(c, sizeof(double) * slot_doubles * size_t(chunk)));
        double *ab = static_cast<double *>(e->wbuf.ensure(c, sizeof(double) * 2 * size_t(steps) * size_t(w)));
        int64_t slot = 0, base = 0;
        auto flush = [&] {
            launch_reduce(c, part, G, w, slot, ab + 2 * base * nb, nb, 1);
            base += slot;
            slot = 0;
        };
        auto next_part = [&] {
            if (slot == chunk) flush();
            return part + size_t(slot++) * slot_doubles;
        };
        const void *prev = nullptr;
        const void *cur = B;
        for (int64_t k = 0; k < steps; ++k) {
            void *o = (k == 0) ? t1 : (prev == B ? spare : const_cast<void *>(prev));
            Step q;
            q.X = cur;
            q.ldx = nb;
            q.OUT = o;
            q.ldo = nb;
            q.alpha = Cplx{(k == 0 ? 1.0 : 2.0) / half, 0};
            q.beta = Cplx{-(k == 0 ? 1.0 : 2.0) * center / half, 0};
            q.use_beta = true;
            if (k > 0) {
                q.P = prev;
                q.ldp = nb;
                q.gamma = -1.0;
                q.use_p = true;
            }
            q.doubling = true;
            q.n = n;
            q.w = w;
            q.cpl_cap = cap;
            q.groups = G;
            q.part = next_part();
            cheb(e, dtype, q);
            prev = cur;
            cur = o;
        }
        flush();
        launch_doubling_finish(c, ab, mu, L, w);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}
