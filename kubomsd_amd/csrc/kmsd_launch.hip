// Launch-shape policy, kernel dispatch over (vector precision, matrix value
// kind, columns per lane), device allocation and matrix upload.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kmsd_internal.hpp"
#include "kmsd_kernels.hpp"

namespace kmsd {

void *dalloc(kmsd_ctx *c, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (c->used + bytes > c->budget)
        throw Error(KMSD_ERESOURCE, "RESOURCE_BLOCKED: device allocation of " + std::to_string(bytes) +
                                        " bytes exceeds the HBM budget (" + std::to_string(c->used) + " of " +
                                        std::to_string(c->budget) + " bytes in use)");
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw Error(KMSD_ERESOURCE, std::string("RESOURCE_BLOCKED: hipMalloc failed: ") + hipGetErrorString(e));
    c->used += bytes;
    return p;
}

void dfree(kmsd_ctx *c, void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    if (bytes == 0) bytes = 16;
    c->used -= std::min<uint64_t>(c->used, bytes);
}

void *DBuf::ensure(kmsd_ctx *ctx, size_t want) {
    if (p && bytes >= want && c == ctx) return p;
    // Outstanding work may still read the old buffer.
    if (p) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        dfree(c, p, bytes);
        p = nullptr;
        bytes = 0;
    }
    c = ctx;
    p = dalloc(ctx, want);
    bytes = want;
    return p;
}

void DBuf::release() {
    if (p && c) {
        (void)hipStreamSynchronize(c->stream);
        dfree(c, p, bytes);
    }
    p = nullptr;
    bytes = 0;
}

void check_block_args(const void *B, const void *out, int64_t nb, int dtype) {
    if (dtype != KMSD_C64 && dtype != KMSD_C128) throw Error(KMSD_EINVAL, "dtype must be KMSD_C64 or KMSD_C128");
    if (nb < 1 || nb > (int64_t(1) << 20)) throw Error(KMSD_EINVAL, "expected nonempty finite (N,batch) array");
    if (!B || !out) throw Error(KMSD_EINVAL, "null device block");
}

// ---------------------------------------------------------------------------
// matrix upload

void matrix_release(kmsd_matrix *m) {
    kmsd_ctx *c = m->ctx;
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    const size_t vb = m->fmt == KMSD_BSR4 ? 16 : 1;
    const size_t vals = size_t(m->nnz) * vb;
    const size_t mult = m->vkind == VK_COMPLEX ? 2 : 1;
    const int64_t rows = m->fmt == KMSD_BSR4 ? m->n / 4 : m->n;
    dfree(c, m->ptr, sizeof(int32_t) * size_t(rows + 1));
    dfree(c, m->idx, sizeof(int32_t) * size_t(m->nnz));
    dfree(c, m->val64, sizeof(double) * vals * mult);
    dfree(c, m->val32, sizeof(float) * vals * mult);
    m->ptr = m->idx = nullptr;
    m->val64 = m->val32 = nullptr;
}

void matrix_upload(kmsd_ctx *c, kmsd_matrix *m, int fmt, int64_t n, int64_t nnz, const int64_t *rowptr,
                   const int64_t *colidx, const double *values) {
    if (fmt != KMSD_CSR && fmt != KMSD_BSR4) throw Error(KMSD_EINVAL, "unknown matrix format");
    if (n < 1 || nnz < 0 || !rowptr || (nnz && (!colidx || !values)))
        throw Error(KMSD_EINVAL, "invalid CSR dimensions/threads");
    if (fmt == KMSD_BSR4 && n % 4) throw Error(KMSD_EINVAL, "BSR4 requires n divisible by 4");
    const int64_t rows = fmt == KMSD_BSR4 ? n / 4 : n;
    if (n >= (int64_t(1) << 31) || nnz >= (int64_t(1) << 31))
        throw Error(KMSD_EINVAL, "matrix exceeds 32-bit device indexing");
    if (rowptr[0] != 0 || rowptr[rows] != nnz) throw Error(KMSD_EINVAL, "invalid CSR row pointers");
    for (int64_t r = 0; r < rows; ++r)
        if (rowptr[r] > rowptr[r + 1]) throw Error(KMSD_EINVAL, "nonmonotone CSR row pointers");
    const int64_t vpb = fmt == KMSD_BSR4 ? 16 : 1;
    bool any_re = false, any_im = false;
    for (int64_t k = 0; k < nnz; ++k) {
        if (colidx[k] < 0 || colidx[k] >= rows) throw Error(KMSD_EINVAL, "invalid CSR index or nonfinite entry");
        for (int64_t q = 0; q < vpb; ++q) {
            const double re = values[2 * (k * vpb + q)], im = values[2 * (k * vpb + q) + 1];
            if (!std::isfinite(re) || !std::isfinite(im))
                throw Error(KMSD_EINVAL, "invalid CSR index or nonfinite entry");
            any_re |= re != 0.0;
            any_im |= im != 0.0;
        }
    }
    m->fmt = fmt;
    m->n = n;
    m->nnz = nnz;
    m->vkind = any_im ? (any_re ? VK_COMPLEX : VK_IMAG) : VK_REAL;
    // identity detection (CSR only)
    if (fmt == KMSD_CSR && nnz == n) {
        bool id = true;
        for (int64_t r = 0; r < n && id; ++r)
            id = rowptr[r + 1] - rowptr[r] == 1 && colidx[rowptr[r]] == r && values[2 * rowptr[r]] == 1.0 &&
                 values[2 * rowptr[r] + 1] == 0.0;
        m->identity = id;
    }
    std::vector<int32_t> p32(rows + 1), i32(nnz);
    for (int64_t r = 0; r <= rows; ++r) p32[r] = int32_t(rowptr[r]);
    for (int64_t k = 0; k < nnz; ++k) i32[k] = int32_t(colidx[k]);
    const size_t nv = size_t(nnz * vpb);
    std::vector<double> v64;
    std::vector<float> v32;
    if (m->vkind == VK_COMPLEX) {
        v64.assign(values, values + 2 * nv);
        v32.resize(2 * nv);
        for (size_t e = 0; e < 2 * nv; ++e) v32[e] = float(v64[e]);
    } else {
        const int off = m->vkind == VK_IMAG ? 1 : 0;
        v64.resize(nv);
        v32.resize(nv);
        for (size_t e = 0; e < nv; ++e) {
            v64[e] = values[2 * e + off];
            v32[e] = float(v64[e]);
        }
    }
    m->ptr = static_cast<int32_t *>(dalloc(c, sizeof(int32_t) * p32.size()));
    m->idx = static_cast<int32_t *>(dalloc(c, sizeof(int32_t) * std::max<size_t>(1, i32.size())));
    m->val64 = dalloc(c, sizeof(double) * std::max<size_t>(1, v64.size()));
    m->val32 = dalloc(c, sizeof(float) * std::max<size_t>(1, v32.size()));
    HIPCHECK(hipMemcpyAsync(m->ptr, p32.data(), sizeof(int32_t) * p32.size(), hipMemcpyHostToDevice, c->stream));
    if (nnz) {
        HIPCHECK(hipMemcpyAsync(m->idx, i32.data(), sizeof(int32_t) * i32.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipMemcpyAsync(m->val64, v64.data(), sizeof(double) * v64.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipMemcpyAsync(m->val32, v32.data(), sizeof(float) * v32.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
}

// ---------------------------------------------------------------------------
// launch shapes

namespace {

struct Shape {
    int cpl = 1;
    int lpr_log2 = 6;
    int gy = 1;
    int64_t gx = 8;
};

int ilog2_ceil(int64_t v) {
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

Shape shape_for(kmsd_ctx *c, int64_t n, int32_t w, int dtype) {
    Shape s;
    if (dtype == KMSD_C64) {
        s.cpl = c->cpl ? c->cpl : 2;
        if (w % s.cpl) s.cpl = 1;
    }
    const int64_t lanes = (w + s.cpl - 1) / s.cpl;
    s.lpr_log2 = std::min(6, ilog2_ceil(lanes));
    const int64_t lpr = int64_t(1) << s.lpr_log2;
    s.gy = int((lanes + lpr - 1) / lpr);
    const int64_t rpb = 4 * (64 / lpr);
    const int64_t nblk = (n + rpb - 1) / rpb;
    const int64_t want = ((nblk + 7) / 8) * 8;
    s.gx = std::max<int64_t>(8, std::min<int64_t>(c->groups, want));
    return s;
}

template <class VT, class MR, bool MC, class AT, int CPL>
void go(kmsd_ctx *c, const Shape &sh, const StepArgs &a) {
    hipLaunchKernelGGL((step_kernel<VT, MR, MC, AT, CPL>), dim3(unsigned(sh.gx), unsigned(sh.gy)), dim3(256), 0,
                       c->stream, a);
}

template <class VT, class MR, class AT, int CPL>
void go_vk(kmsd_ctx *c, const Shape &sh, const StepArgs &a, bool complex_values) {
    if (complex_values)
        go<VT, MR, true, AT, CPL>(c, sh, a);
    else
        go<VT, MR, false, AT, CPL>(c, sh, a);
}

int64_t blocks_for(int64_t total) { return std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192)); }

}  // namespace

int64_t step_groups(kmsd_ctx *c, int64_t n, int32_t w, int dtype) { return shape_for(c, n, w, dtype).gx; }

int64_t launch_step(kmsd_ctx *c, int dtype, const Step &s) {
    if (s.n < 1 || s.w < 1) return 0;
    const Shape sh = shape_for(c, s.n, s.w, dtype);
    StepArgs a{};
    uint32_t fl = 0;
    Cplx alpha = s.alpha;
    const kmsd_matrix *A = s.A;
    if (A) {
        if (A->fmt != KMSD_CSR) throw Error(KMSD_EINVAL, "block-sparse products use the BSR path");
        fl |= F_SPMM;
        a.ptr = A->ptr;
        a.idx = A->idx;
        const bool f32 = dtype == KMSD_C64 && c->fp32_math;
        a.val = f32 ? A->val32 : A->val64;
        if (A->vkind == VK_IMAG) alpha = Cplx{-s.alpha.im, s.alpha.re};  // A = i * stored
    }
    if (s.use_beta) fl |= F_BETA;
    if (s.use_p) fl |= F_P;
    if (s.OUT) fl |= F_STORE;
    if (s.W) fl |= F_MOM;
    if (s.ACC) fl |= F_ACC;
    a.X = s.X;
    a.Z = s.Z;
    a.P = s.P;
    a.OUT = s.OUT;
    a.W = s.W;
    a.ACC = s.ACC;
    a.part = s.part;
    a.ldx = s.ldx;
    a.ldz = s.ldz;
    a.ldp = s.ldp;
    a.ldo = s.ldo;
    a.ldw = s.ldw;
    a.lda = s.lda;
    a.n = s.n;
    a.alpha_re = alpha.re;
    a.alpha_im = alpha.im;
    a.beta_re = s.beta.re;
    a.beta_im = s.beta.im;
    a.gamma = s.gamma;
    a.coef_re = s.coef.re;
    a.coef_im = s.coef.im;
    a.w = s.w;
    a.lpr_log2 = sh.lpr_log2;
    a.flags = fl;
    const bool cv = A && A->vkind == VK_COMPLEX;
    if (dtype == KMSD_C128) {
        go_vk<double, double, double, 1>(c, sh, a, cv);
    } else if (c->fp32_math) {
        if (sh.cpl == 2)
            go_vk<float, float, float, 2>(c, sh, a, cv);
        else
            go_vk<float, float, float, 1>(c, sh, a, cv);
    } else {
        if (sh.cpl == 2)
            go_vk<float, double, double, 2>(c, sh, a, cv);
        else
            go_vk<float, double, double, 1>(c, sh, a, cv);
    }
    HIPCHECK(hipGetLastError());
    return sh.gx;
}

void launch_reduce(kmsd_ctx *c, const double *part, int64_t groups, int32_t w, int64_t nsteps, double *out,
                   int64_t ldo, int mode) {
    if (nsteps < 1 || w < 1) return;
    hipLaunchKernelGGL(reduce_partials, dim3(unsigned(w), unsigned(nsteps)), dim3(256), 0, c->stream, part, groups,
                       w, groups * int64_t(w), out, ldo, mode);
    HIPCHECK(hipGetLastError());
}

void launch_scale(kmsd_ctx *c, int dtype, const void *X, int64_t ldx, void *O, int64_t ldo, int64_t n, int32_t w,
                  Cplx coef) {
    const int64_t g = blocks_for(n * w);
    if (dtype == KMSD_C128)
        hipLaunchKernelGGL(scale_kernel<double>, dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, O, ldo, n, w,
                           coef.re, coef.im);
    else
        hipLaunchKernelGGL(scale_kernel<float>, dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, O, ldo, n, w,
                           coef.re, coef.im);
    HIPCHECK(hipGetLastError());
}

void launch_convert(kmsd_ctx *c, int din, const void *X, int64_t ldx, int dout, void *O, int64_t ldo, int64_t n,
                    int32_t w) {
    const int64_t g = blocks_for(n * w);
    if (din == KMSD_C128 && dout == KMSD_C64)
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, O, ldo,
                           n, w);
    else if (din == KMSD_C64 && dout == KMSD_C128)
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, O, ldo,
                           n, w);
    else if (din == KMSD_C64)
        hipLaunchKernelGGL((convert_kernel<float, float>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, O, ldo,
                           n, w);
    else
        hipLaunchKernelGGL((convert_kernel<double, double>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, O,
                           ldo, n, w);
    HIPCHECK(hipGetLastError());
}

bool launch_nonfinite(kmsd_ctx *c, int dtype, const void *X, int64_t ldx, int64_t n, int32_t w) {
    int32_t zero = 0, got = 0;
    HIPCHECK(hipMemcpyAsync(c->flag, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
    const int64_t g = blocks_for(n * w);
    if (dtype == KMSD_C128)
        hipLaunchKernelGGL(nonfinite_kernel<double>, dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, n, w,
                           c->flag);
    else
        hipLaunchKernelGGL(nonfinite_kernel<float>, dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, n, w,
                           c->flag);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(&got, c->flag, sizeof(got), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return got != 0;
}

void require_finite(kmsd_ctx *c, int dtype, const void *X, int64_t ldx, int64_t n, int32_t w) {
    if (launch_nonfinite(c, dtype, X, ldx, n, w)) throw Error(KMSD_EINVAL, "expected finite (N,batch) array");
}

void spmm(kmsd_ctx *c, int dtype, const kmsd_matrix *A, const void *X, int64_t ldx, void *out, int64_t ldo, int32_t w,
          double are, double aim, bool accumulate) {
    Step s;
    s.A = A;
    s.X = X;
    s.ldx = ldx;
    s.alpha = Cplx{are, aim};
    s.n = A->n;
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

}  // namespace kmsd
