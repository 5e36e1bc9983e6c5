// Overlap metric: device-resident S, diagonal / atom-block Jacobi factors,
// batched PCG (reference Metric, cpp/native.cpp:145-480).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kmsd_internal.hpp"
#include "kmsd_pcg.hpp"

using namespace kmsd;

namespace {
void metric_free(kmsd_metric *m);
}

extern "C" int kmsd_metric_create(kmsd_ctx *c, kmsd_matrix *S, const double *diag, int64_t nblocks,
                                  const int64_t *block_ptr, const double *factors, kmsd_metric **out) {
    if (!c || !out) return KMSD_EINVAL;
    *out = nullptr;
    kmsd_metric *m = new kmsd_metric();
    m->ctx = c;
    {
        std::lock_guard<std::recursive_mutex> lock(c->mu);
        ctx_adopt(c);  // the failure path below destroys m, which orphans it
    }
    const int rc = guard(c, [&] {
        if (!S || !diag) throw Error(KMSD_EINVAL, "invalid metric diagonal/threads");
        const int64_t n = S->n;
        m->S = S;
        m->n = n;
        for (int64_t i = 0; i < n; ++i)
            if (!(diag[i] > 0) || !std::isfinite(diag[i])) throw Error(KMSD_EINVAL, "invalid metric diagonal");
        if (nblocks < 0 || (nblocks && (!block_ptr || !factors)))
            throw Error(KMSD_EINVAL, "invalid block-Jacobi factor partition");
        size_t nfac = 0;
        for (int64_t q = 0; q < nblocks; ++q) {
            const int64_t a = block_ptr[q], b = block_ptr[q + 1];
            if ((q == 0 && a != 0) || b <= a || b > n) throw Error(KMSD_EINVAL, "invalid block-Jacobi factor partition");
            nfac += size_t((b - a) * (b - a));
        }
        if (nblocks && block_ptr[nblocks] != n) throw Error(KMSD_EINVAL, "incomplete block-Jacobi partition");
        for (size_t e = 0; e < 2 * nfac; ++e)
            if (!std::isfinite(factors[e])) throw Error(KMSD_EINVAL, "invalid block-Jacobi factor partition");
        bool diag_one = true;
        for (int64_t i = 0; i < n && diag_one; ++i) diag_one = diag[i] == 1.0;
        m->identity = S->identity && diag_one && nblocks == 0;
        m->diag = static_cast<double *>(dalloc(c, sizeof(double) * n));
        m->bytes = sizeof(double) * n;
        copy_h2d(c, m->diag, diag, sizeof(double) * n);
        if (nblocks) metric_upload_blocks(m, nblocks, block_ptr, nullptr, factors);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
    if (rc) {
        kmsd_metric_destroy(m);
        return rc;
    }
    *out = m;
    return KMSD_OK;
}

extern "C" int kmsd_metric_destroy(kmsd_metric *m) {
    if (!m) return KMSD_OK;
    kmsd_ctx *c = m->ctx;
    bool free_ctx = false;
    {
        std::lock_guard<std::recursive_mutex> lock(c->mu);
        metric_free(m);
        free_ctx = ctx_orphan(c);
    }
    if (free_ctx) ctx_free(c);
    return KMSD_OK;
}

namespace {
void metric_free(kmsd_metric *m) {
    kmsd_ctx *c = m->ctx;
    (void)hipStreamSynchronize(c->stream);
    pcg_release(m);
    if (m->W) {
        matrix_release(m->W);
        delete m->W;
        m->W = nullptr;
    }
    metric_release_blocks(m);
    metric_release_ic(m);
    const size_t nfac_bytes = m->bytes;  // the diagonal
    if (m->diag) (void)hipFree(m->diag);
    c->used -= std::min<uint64_t>(c->used, nfac_bytes);
    delete m;
}
}  // namespace

extern "C" int kmsd_metric_set_residual_policy(kmsd_metric *m, int adaptive, int interval, double gap) {
    if (!m) return KMSD_EINVAL;
    return guard(m->ctx, [&] {
        if (interval < 1 || !(gap > 0 && gap < 1)) throw Error(KMSD_EINVAL, "invalid residual policy");
        m->adaptive = adaptive != 0;
        m->check_interval = interval;
        m->gap_ratio = gap;
    });
}

extern "C" int kmsd_metric_set_mixed(kmsd_metric *m, int enabled) {
    if (!m) return KMSD_EINVAL;
    return guard(m->ctx, [&] { m->mixed_c128 = enabled != 0; });
}

static void copy_report(const kmsd_metric *m, kmsd_report *r) {
    if (!r) return;
    if (r->converged) std::copy(m->rep_conv.begin(), m->rep_conv.end(), r->converged);
    if (r->breakdown) std::copy(m->rep_brk.begin(), m->rep_brk.end(), r->breakdown);
    if (r->iterations) std::copy(m->rep_its.begin(), m->rep_its.end(), r->iterations);
    if (r->true_residual) std::copy(m->rep_res.begin(), m->rep_res.end(), r->true_residual);
    if (r->relative_residual) std::copy(m->rep_rel.begin(), m->rep_rel.end(), r->relative_residual);
}

extern "C" int kmsd_metric_solve(kmsd_metric *m, const void *B, void *Y, int64_t nb, int dtype, double rtol,
                                 double atol, int maxiter, kmsd_report *report) {
    if (!m) return KMSD_EINVAL;
    kmsd_ctx *c = m->ctx;
    const int rc = guard(c, [&] {
        check_block_args(B, Y, nb, dtype);
        require_finite(c, dtype, B, nb, m->n, int32_t(nb));
        metric_solve(m, dtype, B, nb, Y, nb, int32_t(nb), rtol, atol, maxiter);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
    copy_report(m, report);
    return rc;
}

extern "C" int kmsd_metric_counters(const kmsd_metric *m, int64_t *o, double *max_relative) {
    if (!m) return KMSD_EINVAL;
    if (o) {
        o[0] = m->calls;
        o[1] = m->iterations;
        o[2] = m->refreshes;
        o[3] = m->residual_checks;
        o[4] = m->restarts;
    }
    if (max_relative) *max_relative = m->max_relative;
    return KMSD_OK;
}

namespace kmsd {

void metric_solve(kmsd_metric *m, int dtype, const void *B, int64_t ldb, void *Y, int64_t ldy, int32_t nb,
                  double rtol, double atol, int maxiter) {
    if (!std::isfinite(rtol) || !std::isfinite(atol) || rtol < 0 || atol < 0 || maxiter < 1)
        throw Error(KMSD_EINVAL, "invalid metric RHS shape/tolerances");
    if (m->identity) {
        // S = I: the reference's PCG converges in one iteration with alpha = 1
        // and returns B up to the /safe*safe rescale; the block is copied.
        if (B != Y || ldb != ldy) launch_scale(m->ctx, dtype, B, ldb, Y, ldy, m->n, nb, Cplx{1, 0});
        m->rep_conv.assign(nb, 1);
        m->rep_brk.assign(nb, 0);
        m->rep_its.assign(nb, 0);
        m->rep_res.assign(nb, 0.0);
        m->rep_rel.assign(nb, 0.0);
        ++m->calls;
        ++m->residual_checks;  // the reference's final true-residual acceptance
        return;
    }
    pcg_solve(m, dtype, B, ldb, Y, ldy, nb, rtol, atol, maxiter);
}

}  // namespace kmsd

extern "C" int kmsd_metric_sample(kmsd_metric *m, const void *xi, void *out, int64_t nb, int dtype, const double *coeff,
                                  int64_t ncoeff, double center, double half) {
    if (!m) return KMSD_EINVAL;
    kmsd_ctx *c = m->ctx;
    return guard(c, [&] {
        check_block_args(xi, out, nb, dtype);
        if (ncoeff < 1 || !coeff) throw Error(KMSD_EINVAL, "invalid sampling dimensions");
        const int64_t n = m->n;
        const int32_t w = int32_t(nb);
        require_finite(c, dtype, xi, nb, n, w);
        if (ncoeff == 1) {
            launch_scale(c, dtype, xi, nb, out, nb, n, w, Cplx{0, 0});
        } else {
            if (!(half > 0)) throw Error(KMSD_EINVAL, "sampling half-width must be positive");
            if (!m->W) m->W = matrix_congruence_copy(c, m->S, m->diag);
            // Clenshaw: b_k = 2 A b_{k+1} - b_{k+2} + c_k xi, A = (W - center)/half
            const size_t bytes = size_t(n) * nb * elem_bytes(dtype);
            void *u = m->sb1.ensure(c, bytes);
            void *v = m->sb2.ensure(c, bytes);
            launch_scale(c, dtype, xi, nb, u, nb, n, w, Cplx{coeff[ncoeff - 1], 0});  // b_{K-1}
            launch_scale(c, dtype, xi, nb, v, nb, n, w, Cplx{0, 0});                 // b_K = 0
            void *cur = u, *prev = v;  // cur = b_{k+1}, prev = b_{k+2}
            for (int64_t k = ncoeff - 2; k >= 1; --k) {
                Step s;  // new = 2 A cur - prev, written over prev
                s.A = m->W;
                s.X = cur;
                s.ldx = nb;
                s.alpha = Cplx{2 / half, 0};
                s.beta = Cplx{-2 * center / half, 0};
                s.use_beta = true;
                s.P = prev;
                s.ldp = nb;
                s.gamma = -1.0;
                s.use_p = true;
                s.OUT = prev;
                s.ldo = nb;
                s.n = n;
                s.w = w;
                launch_step(c, dtype, s);
                Step t;  // new += c_k xi
                t.Z = xi;
                t.ldz = nb;
                t.alpha = Cplx{coeff[k], 0};
                t.ACC = prev;
                t.lda = nb;
                t.n = n;
                t.w = w;
                launch_step(c, dtype, t);
                std::swap(cur, prev);
            }
            // out = A b_1 - b_2 ; then out = z (out + c_0 xi)
            Step s;
            s.A = m->W;
            s.X = cur;
            s.ldx = nb;
            s.alpha = Cplx{1 / half, 0};
            s.beta = Cplx{-center / half, 0};
            s.use_beta = true;
            s.P = prev;
            s.ldp = nb;
            s.gamma = -1.0;
            s.use_p = true;
            s.OUT = out;
            s.ldo = nb;
            s.n = n;
            s.w = w;
            launch_step(c, dtype, s);
        }
        launch_sample_finish(c, dtype, out, xi, nb, n, w, m->diag, coeff[0]);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int kmsd_metric_inner(kmsd_metric *m, const void *A, const void *B, double *out, int64_t nb, int dtype) {
    if (!m) return KMSD_EINVAL;
    kmsd_ctx *c = m->ctx;
    return guard(c, [&] {
        check_block_args(A, B, nb, dtype);
        if (!out) throw Error(KMSD_EINVAL, "null output");
        const int64_t n = m->n;
        const int32_t w = int32_t(nb);
        const int cap = cpl_cap_for(dtype, {{A, nb}, {B, nb}});
        const int64_t G = step_groups(c, n, w, dtype, cap, m->identity ? nullptr : m->S);
        double *part = static_cast<double *>(m->pinner.ensure(c, sizeof(double) * 2 * size_t(G) * size_t(w)));
        Step s;
        if (m->identity) {
            s.Z = B;
            s.ldz = nb;
        } else {
            s.A = m->S;
            s.X = B;
            s.ldx = nb;
        }
        s.W = A;
        s.ldw = nb;
        s.part = part;
        s.n = n;
        s.w = w;
        s.cpl_cap = cap;
        s.groups = G;
        launch_step(c, dtype, s);
        launch_reduce(c, part, G, w, 1, out, nb, 1);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}
