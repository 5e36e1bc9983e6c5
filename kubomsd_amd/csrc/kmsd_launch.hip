// Launch-shape policy, kernel dispatch over (vector precision, matrix value
// kind, columns per lane), device allocation and matrix upload.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <unordered_map>

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
    for (void *p : {static_cast<void *>(m->ptr), static_cast<void *>(m->idx), m->val64, m->val32, m->dval64,
                    m->dval32})
        if (p) (void)hipFree(p);
    c->used -= std::min<uint64_t>(c->used, m->bytes);
    m->bytes = 0;
    m->ptr = m->idx = nullptr;
    m->val64 = m->val32 = m->dval64 = m->dval32 = nullptr;
}

namespace {
template <class T>
T *upload(kmsd_ctx *c, kmsd_matrix *m, const std::vector<T> &h) {
    const size_t bytes = sizeof(T) * std::max<size_t>(1, h.size());
    T *d = static_cast<T *>(dalloc(c, bytes));
    m->bytes += bytes;
    if (!h.empty()) copy_h2d(c, d, h.data(), sizeof(T) * h.size());
    return d;
}
}  // namespace

namespace {

// Blocks a validated CSR matrix into 4x4 blocks.  Returns false (and builds
// nothing) when want_auto is set and the blocks would be less than 75% full.
bool bsr4_upload(kmsd_ctx *c, kmsd_matrix *m, bool want_auto, int64_t n, int64_t nnz, const int64_t *rowptr,
                 const int64_t *colidx, const double *values, int comps, int off) {
    const int64_t nq = n / 4;
    std::vector<int32_t> bptr(nq + 1, 0), bidx;
    bidx.reserve(size_t(nnz / 8 + 16));
    std::vector<int32_t> cols;
    for (int64_t q = 0; q < nq; ++q) {
        cols.clear();
        for (int64_t k = rowptr[4 * q]; k < rowptr[4 * q + 4]; ++k) cols.push_back(int32_t(colidx[k] >> 2));
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        bidx.insert(bidx.end(), cols.begin(), cols.end());
        bptr[q + 1] = int32_t(bidx.size());
    }
    const int64_t nnzb = int64_t(bidx.size());
    if (want_auto && 4 * nnz < 3 * 16 * nnzb) return false;
    if (16 * nnzb >= (int64_t(1) << 31)) throw Error(KMSD_EINVAL, "matrix exceeds 32-bit device indexing");
    std::vector<double> v64(size_t(std::max<int64_t>(1, 16 * nnzb)) * comps, 0.0);
    for (int64_t q = 0; q < nq; ++q) {
        const int32_t *b0 = bidx.data() + bptr[q], *b1 = bidx.data() + bptr[q + 1];
        for (int r = 0; r < 4; ++r)
            for (int64_t k = rowptr[4 * q + r]; k < rowptr[4 * q + r + 1]; ++k) {
                const int32_t jb = int32_t(colidx[k] >> 2);
                const int64_t pos = bptr[q] + (std::lower_bound(b0, b1, jb) - b0);
                const size_t e = size_t(pos * 16 + r * 4 + (colidx[k] & 3)) * comps;
                for (int t = 0; t < comps; ++t) v64[e + t] += values[2 * k + off + t];
            }
    }
    std::vector<float> v32(v64.size());
    for (size_t e = 0; e < v64.size(); ++e) v32[e] = float(v64[e]);
    m->fmt = KMSD_BSR4;
    m->noff = 16 * nnzb;
    m->ell = 0;
    m->has_diag = false;
    m->identity = false;
    m->ptr = upload(c, m, bptr);
    m->idx = upload(c, m, bidx);
    m->val64 = upload(c, m, v64);
    m->val32 = upload(c, m, v32);
    // Plane stride: the largest positive block offset that at least a quarter
    // of the block rows have and that divides the block rows into planes (a
    // 3D stencil's outermost neighbour).  Small strides need no tiling.
    std::unordered_map<int32_t, int64_t> offsets;
    for (int64_t q = 0; q < nq; ++q)
        for (int32_t k = bptr[q]; k < bptr[q + 1]; ++k)
            if (bidx[k] > q) ++offsets[int32_t(bidx[k] - q)];
    m->plane = 0;
    for (const auto &o : offsets)
        if (o.first >= 256 && 4 * o.second >= nq && nq % o.first == 0 && o.first > m->plane) m->plane = o.first;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return true;
}

}  // namespace

void matrix_upload(kmsd_ctx *c, kmsd_matrix *m, int fmt, int64_t n, int64_t nnz, const int64_t *rowptr,
                   const int64_t *colidx, const double *values) {
    if (fmt != KMSD_CSR && fmt != KMSD_BSR4 && fmt != KMSD_AUTO) throw Error(KMSD_EINVAL, "unknown matrix format");
    if (n < 1 || nnz < 0 || !rowptr || (nnz && (!colidx || !values)))
        throw Error(KMSD_EINVAL, "invalid CSR dimensions/threads");
    if (n >= (int64_t(1) << 31) || nnz >= (int64_t(1) << 31))
        throw Error(KMSD_EINVAL, "matrix exceeds 32-bit device indexing");
    if (rowptr[0] != 0 || rowptr[n] != nnz) throw Error(KMSD_EINVAL, "invalid CSR row pointers");
    for (int64_t r = 0; r < n; ++r)
        if (rowptr[r] > rowptr[r + 1]) throw Error(KMSD_EINVAL, "nonmonotone CSR row pointers");
    bool any_re = false, any_im = false;
    for (int64_t k = 0; k < nnz; ++k) {
        const double re = values[2 * k], im = values[2 * k + 1];
        if (colidx[k] < 0 || colidx[k] >= n || !std::isfinite(re) || !std::isfinite(im))
            throw Error(KMSD_EINVAL, "invalid CSR index or nonfinite entry");
        any_re |= re != 0.0;
        any_im |= im != 0.0;
    }
    m->fmt = KMSD_CSR;
    m->n = n;
    m->nnz = nnz;
    m->vkind = any_im ? (any_re ? VK_COMPLEX : VK_IMAG) : VK_REAL;
    const int comps = m->vkind == VK_COMPLEX ? 2 : 1;
    const int off = m->vkind == VK_IMAG ? 1 : 0;
    if (fmt == KMSD_BSR4 && n % 4) throw Error(KMSD_EINVAL, "BSR4 storage needs a multiple of 4 rows");
    // A 75%-full 4x4 block holds 12 entries, three per row: below 3 entries
    // per row on average AUTO cannot choose BSR4.
    if ((fmt == KMSD_BSR4 || (fmt == KMSD_AUTO && n % 4 == 0 && nnz >= 3 * n)) &&
        bsr4_upload(c, m, fmt == KMSD_AUTO, n, nnz, rowptr, colidx, values, comps, off))
        return;
    // Split: dense diagonal (duplicates summed) + off-diagonal entries.
    std::vector<double> dg(size_t(n) * comps, 0.0);
    std::vector<int64_t> cnt(n, 0);
    int64_t maxlen = 0, noff = 0;
    for (int64_t r = 0; r < n; ++r) {
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            if (colidx[k] == r) {
                for (int q = 0; q < comps; ++q) dg[size_t(r) * comps + q] += values[2 * k + off + q];
            } else {
                ++cnt[r];
            }
        }
        maxlen = std::max(maxlen, cnt[r]);
        noff += cnt[r];
    }
    m->has_diag = std::any_of(dg.begin(), dg.end(), [](double v) { return v != 0.0; });
    if (nnz == n) {
        bool id = true;
        for (int64_t r = 0; r < n && id; ++r)
            id = rowptr[r + 1] - rowptr[r] == 1 && colidx[rowptr[r]] == r && values[2 * rowptr[r]] == 1.0 &&
                 values[2 * rowptr[r] + 1] == 0.0;
        m->identity = id;
    }
    // ELL when padding adds at most ~25% entries.
    const bool use_ell = maxlen > 0 && maxlen * n <= noff + noff / 4 + n && maxlen <= 64;
    m->ell = use_ell ? int32_t(maxlen) : 0;
    const int64_t stored = use_ell ? maxlen * n : noff;
    m->noff = stored;
    std::vector<int32_t> i32(stored), p32;
    std::vector<double> v64(size_t(stored) * comps, 0.0);
    if (!use_ell) p32.resize(n + 1);
    int64_t pos = 0;
    for (int64_t r = 0; r < n; ++r) {
        if (!use_ell) p32[r] = int32_t(pos);
        int64_t q = use_ell ? r * maxlen : pos;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            if (colidx[k] == r) continue;
            i32[q] = int32_t(colidx[k]);
            for (int t = 0; t < comps; ++t) v64[size_t(q) * comps + t] = values[2 * k + off + t];
            ++q;
        }
        if (use_ell) {
            for (; q < (r + 1) * maxlen; ++q) i32[q] = int32_t(r);  // zero-valued padding
        } else {
            pos = q;
        }
    }
    if (!use_ell) p32[n] = int32_t(pos);
    // One off-diagonal value (a lattice with one hopping): the prefetching
    // ELL kernel reads that value once instead of streaming n * ell copies of
    // it, and uses zero in padding slots (column = own row, which no stored
    // off-diagonal entry has).  The array stays for the other kernels.
    if (use_ell && noff > 0) {
        int64_t first = -1;
        bool same = true;
        for (int64_t q = 0; q < stored && same; ++q) {
            if (i32[q] == q / maxlen) continue;  // padding
            if (first < 0) first = q;
            for (int t = 0; t < comps; ++t) same = same && v64[size_t(q) * comps + t] == v64[size_t(first) * comps + t];
        }
        m->uniform_slot = same ? first : -1;
    }
    std::vector<float> v32(v64.size()), d32(dg.size());
    for (size_t e = 0; e < v64.size(); ++e) v32[e] = float(v64[e]);
    for (size_t e = 0; e < dg.size(); ++e) d32[e] = float(dg[e]);
    if (!use_ell) m->ptr = upload(c, m, p32);
    m->idx = upload(c, m, i32);
    m->val64 = upload(c, m, v64);
    m->val32 = upload(c, m, v32);
    m->dval64 = upload(c, m, dg);
    m->dval32 = upload(c, m, d32);
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

}  // namespace

// Automatic columns per lane: 2 (complex64) / 1 (complex128); 4 / 2 for 4x4
// block products, whose lanes reuse each gathered row for four output rows.
int default_cpl(int dtype, const kmsd_matrix *A) {
    const bool bsr = A && A->fmt == KMSD_BSR4;
    if (dtype == KMSD_C64) return bsr ? 4 : 2;
    return bsr ? 2 : 1;
}

namespace {

int ilog2_ceil(int64_t v) {
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

using KernelFn = void (*)(const StepArgs);

// The kernel one fused step runs, from its operand kinds.
struct KernelKey {
    int dtype = KMSD_C128;
    bool f32 = false;    // complex64 blocks with FP32 matrix values (FP64 row sums)
    int cpl = 1;
    bool cv = false;     // complex matrix values
    bool bsr = false;
    int ell = 0;         // ELL width with the prefetching kernel, 0 = generic kernel
    int lpr_log2 = 6;
    bool vout = false;   // the ELL kernel also forms V X (FP64 values and row sums, real A)
    bool mfma = false;   // bsr4_mfma_kernel (FP32-valued 4x4 blocks on the matrix cores)
    bool blk = false;    // bsr4_mfma_blk_kernel (the same with a block-Jacobi solve of each block row)
    bool lds = false;    // bsr4_lds_kernel (FP64-valued 4x4 blocks, LDS-staged)
};

template <class VT, class MR, bool MC, class AT, int CPL>
KernelFn pick(const KernelKey &k) {
    if (k.vout) {
        if constexpr (std::is_same<MR, double>::value && std::is_same<AT, double>::value && !MC) {
            if (k.ell == 3) return ell_step_kernel<VT, MR, MC, AT, CPL, 3, true>;
            if (k.ell == 4) return ell_step_kernel<VT, MR, MC, AT, CPL, 4, true>;
            if (k.ell == 6) return ell_step_kernel<VT, MR, MC, AT, CPL, 6, true>;
        }
        throw Error(KMSD_EDEVICE, "internal error: no fused velocity kernel for these operands");
    }
    if (k.ell == 3) return ell_step_kernel<VT, MR, MC, AT, CPL, 3>;
    if (k.ell == 4) return ell_step_kernel<VT, MR, MC, AT, CPL, 4>;
    if (k.ell == 6) return ell_step_kernel<VT, MR, MC, AT, CPL, 6>;
    if (k.bsr && k.lpr_log2 >= 4) return bsr4_step_kernel<VT, MR, MC, AT, CPL, true>;
    if (k.bsr) return bsr4_step_kernel<VT, MR, MC, AT, CPL, false>;
    return step_kernel<VT, MR, MC, AT, CPL>;
}

template <class VT, class MR, class AT, int CPL>
KernelFn pick_vk(const KernelKey &k) {
    return k.cv ? pick<VT, MR, true, AT, CPL>(k) : pick<VT, MR, false, AT, CPL>(k);
}

KernelFn select_kernel(const kmsd_ctx *c, const KernelKey &k) {
    if (k.dtype == KMSD_C128) return k.cpl == 2 ? pick_vk<double, double, double, 2>(k)
                                                : pick_vk<double, double, double, 1>(k);
    if (c->fp32_math) {
        if (k.cpl == 4) return pick_vk<float, float, float, 4>(k);
        return k.cpl == 2 ? pick_vk<float, float, float, 2>(k) : pick_vk<float, float, float, 1>(k);
    }
    if (k.f32) {
        if (k.cpl == 4) return pick_vk<float, float, double, 4>(k);
        return k.cpl == 2 ? pick_vk<float, float, double, 2>(k) : pick_vk<float, float, double, 1>(k);
    }
    if (k.cpl == 4) return pick_vk<float, double, double, 4>(k);
    return k.cpl == 2 ? pick_vk<float, double, double, 2>(k) : pick_vk<float, double, double, 1>(k);
}

// Workgroups of `fn` (256 threads) resident per CU, from its register and
// LDS use (cached per kernel).
int resident_per_cu(KernelFn fn) {
    static std::mutex mu;
    static std::unordered_map<const void *, int> cache;
    const void *key = reinterpret_cast<const void *>(fn);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int blocks = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, key, 256, 0));
    blocks = std::max(1, blocks);
    cache.emplace(key, blocks);
    return blocks;
}

// Launch shape and kernel of one fused step with product operator A (null:
// no product) on n x w blocks.  Workgroups: the explicit `groups` tuning, or
// (automatic) exactly the kernel's resident capacity, one workgroup per
// resident slot, so the grid-stride row walk runs as a single wave of
// workgroups with no partially filled second round.
Shape shape_for(kmsd_ctx *c, int64_t n, int32_t w, int dtype, int cpl_max, const kmsd_matrix *A, bool f32,
                KernelFn *fn_out = nullptr, bool vout = false, KernelKey *key_out = nullptr) {
    Shape s;
    const int tuned = c->cpl[dtype == KMSD_C64 ? 0 : 1];
    const int want = tuned ? tuned : default_cpl(dtype, A);
    s.cpl = std::min(want, std::min(cpl_max, dtype == KMSD_C64 ? 4 : 2));
    while (s.cpl > 1 && w % s.cpl) s.cpl >>= 1;
    const int64_t lanes = (w + s.cpl - 1) / s.cpl;
    s.lpr_log2 = std::min(6, ilog2_ceil(lanes));
    const int64_t lpr = int64_t(1) << s.lpr_log2;
    s.gy = int((lanes + lpr - 1) / lpr);
    const int64_t rpb = 4 * (64 / lpr);
    const int64_t nblk = (n + rpb - 1) / rpb;
    const int64_t want_g = groups_for_blocks(c, nblk);
    KernelKey k;
    k.dtype = dtype;
    k.f32 = f32;
    k.cpl = s.cpl;
    k.lpr_log2 = s.lpr_log2;
    k.vout = vout;
    if (A) {
        k.bsr = A->fmt == KMSD_BSR4;
        k.cv = A->vkind == VK_COMPLEX;
        if (!k.bsr && c->ell_kernel && (A->ell == 3 || A->ell == 4 || A->ell == 6)) k.ell = A->ell;
    }
    const KernelFn fn = select_kernel(c, k);
    if (fn_out) *fn_out = fn;
    if (key_out) *key_out = k;
    int64_t g = c->groups;
    if (c->auto_groups) {
        const int64_t resident = int64_t(c->cus) * resident_per_cu(fn);
        g = std::max<int64_t>(8, (resident / s.gy) / 8 * 8);
    }
    s.gx = std::max<int64_t>(8, std::min<int64_t>(g, want_g));
    return s;
}

// Largest columns-per-lane every operand supports (pointer alignment and
// leading dimensions multiples of cpl).
int cpl_limit(const Step &s, int dtype) {
    const size_t eb = elem_bytes(dtype);
    int lim = 4;
    auto fit = [&](const void *p, int64_t ld) {
        if (!p) return;
        while (lim > 1 && ((reinterpret_cast<uintptr_t>(p) % (eb * lim)) || (ld % lim))) lim >>= 1;
    };
    fit(s.X, s.ldx);
    fit(s.Z, s.ldz);
    fit(s.P, s.ldp);
    fit(s.OUT, s.ldo);
    fit(s.W, s.ldw);
    fit(s.ACC, s.lda);
    fit(s.VOUT, s.ldv);
    return lim;
}

bool f32_values(const kmsd_ctx *c, int dtype, const kmsd_matrix *A, bool f32_matrix) {
    return A && dtype == KMSD_C64 && (c->fp32_math || f32_matrix);
}

int64_t blocks_for(int64_t total) { return std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 8192)); }

}  // namespace

int64_t groups_for_blocks(const kmsd_ctx *c, int64_t nblk, int per) {
    const int64_t k = per > 0 ? per : c->min_blocks;
    const int64_t g = (nblk + k - 1) / k;
    return ((g + 7) / 8) * 8;
}

int64_t step_groups(kmsd_ctx *c, int64_t n, int32_t w, int dtype, int cpl_cap, const kmsd_matrix *A, bool f32_matrix) {
    return shape_for(c, n, w, dtype, cpl_cap, A, f32_values(c, dtype, A, f32_matrix)).gx;
}

int cpl_cap_for(int dtype, std::initializer_list<std::pair<const void *, int64_t>> blocks) {
    const size_t eb = elem_bytes(dtype);
    int lim = 4;
    for (const auto &b : blocks) {
        if (!b.first) continue;
        while (lim > 1 && ((reinterpret_cast<uintptr_t>(b.first) % (eb * lim)) || (b.second % lim))) lim >>= 1;
    }
    return lim;
}

namespace {

// The matrix-core 4x4-block kernel for an FP32-valued BSR4 product of a
// complex64 block when its flags, widths and byte extents allow it (buffer
// offsets are 32-bit); otherwise the launch keeps the VALU kernel.
void mfma_shape(kmsd_ctx *c, const Step &s, const kmsd_matrix *A, uint32_t fl, StepArgs &a, Shape &sh, KernelFn &fn,
                KernelKey &key) {
    const uint32_t allowed = F_SPMM | F_STORE | F_P | (s.bfac32 ? (F_BETA | F_RHS) : F_MOM);
    if (fl & ~allowed) return;
    if (s.w % 4 || cpl_limit(s, KMSD_C64) < 4) return;
    const int64_t lim = (int64_t(1) << 31) - (int64_t(1) << 16);
    const int64_t nnzb = A->noff / 16;
    const int64_t xbytes = s.n * s.ldx * 8;
    if (xbytes > lim || nnzb * 64 > lim) return;
    const int lg = std::max(2, std::min(4, ilog2_ceil((s.w + 3) / 4)));
    const int64_t lpr = int64_t(1) << lg;
    fn = s.bfac32 ? bsr4_mfma_blk_kernel : bsr4_mfma_kernel;
    key.mfma = true;
    key.blk = s.bfac32 != nullptr;
    a.bfac32 = s.bfac32;
    sh.cpl = 4;
    sh.lpr_log2 = lg;
    sh.gy = int((s.w + 4 * lpr - 1) / (4 * lpr));
    if (!s.groups) {
        const int64_t qpb = 4 * (64 / lpr);
        const int64_t nblk = (s.n / 4 + qpb - 1) / qpb;
        const int64_t want_g = groups_for_blocks(c, nblk);
        int64_t g = c->groups;
        if (c->auto_groups) g = std::max<int64_t>(8, (int64_t(c->cus) * resident_per_cu(fn) / sh.gy) / 8 * 8);
        sh.gx = std::max<int64_t>(8, std::min<int64_t>(g, want_g));
    }
    a.lpr_log2 = lg;
    a.xbytes = uint32_t(xbytes);
    a.vbytes = uint32_t(nnzb * 64);
    a.ibytes = uint32_t(nnzb * 4);
}

// The LDS-staged FP64-valued 4x4-block kernel (bsr4_lds_kernel) when the
// flags, width and byte extents allow it; otherwise bsr4_step_kernel.
void lds_shape(kmsd_ctx *c, int dtype, const Step &s, const kmsd_matrix *A, uint32_t fl, StepArgs &a, Shape &sh,
               KernelFn &fn, KernelKey &key) {
    if (fl & ~(F_SPMM | F_BETA | F_P | F_STORE | F_MOM)) return;
    const int cpl = dtype == KMSD_C64 ? 4 : 2;
    if (s.w % cpl || cpl_limit(s, dtype) < cpl) return;
    const int64_t lanes = s.w / cpl;
    const int lg = std::max(3, std::min(5, ilog2_ceil(lanes)));
    const int64_t lpr = int64_t(1) << lg;
    const int64_t lim = (int64_t(1) << 31) - (int64_t(1) << 16);
    const int64_t nnzb = A->noff / 16;
    const int64_t xbytes = s.n * s.ldx * int64_t(elem_bytes(dtype));
    if (xbytes > lim || nnzb * 128 > lim) return;
    if (dtype == KMSD_C64)
        fn = lg >= 4 ? bsr4_lds_kernel<float, 4, true> : bsr4_lds_kernel<float, 4, false>;
    else
        fn = lg >= 4 ? bsr4_lds_kernel<double, 2, true> : bsr4_lds_kernel<double, 2, false>;
    key.lds = true;
    sh.cpl = cpl;
    sh.lpr_log2 = lg;
    sh.gy = int((lanes + lpr - 1) / lpr);
    if (!s.groups) {
        const int64_t qpb = 4 * (64 / lpr);
        const int64_t nblk = (s.n / 4 + qpb - 1) / qpb;
        const int64_t want_g = groups_for_blocks(c, nblk);
        int64_t g = c->groups;
        if (c->auto_groups) g = std::max<int64_t>(8, (int64_t(c->cus) * resident_per_cu(fn) / sh.gy) / 8 * 8);
        sh.gx = std::max<int64_t>(8, std::min<int64_t>(g, want_g));
    }
    a.lpr_log2 = lg;
    a.xbytes = uint32_t(xbytes);
    a.vbytes = uint32_t(nnzb * 128);
    a.ibytes = uint32_t(nnzb * 4);
}

// Kernel arguments, launch shape and kernel of one fused step.
void prepare_step(kmsd_ctx *c, int dtype, const Step &s, StepArgs &a, Shape &sh, KernelFn &fn, KernelKey &key) {
    const kmsd_matrix *A = s.A;
    const bool f32 = f32_values(c, dtype, A, s.f32_matrix);
    fn = nullptr;
    if (s.VOUT && (!A || A->fmt == KMSD_BSR4 || !c->ell_kernel || f32 || A->vkind == VK_COMPLEX || !s.vval ||
                   (A->ell != 3 && A->ell != 4 && A->ell != 6)))
        throw Error(KMSD_EDEVICE, "internal error: fused velocity output needs a real width-3/4/6 ELL operator");
    sh = shape_for(c, s.n, s.w, dtype, std::min(s.cpl_cap, cpl_limit(s, dtype)), A, f32, &fn, s.VOUT != nullptr,
                   &key);
    if (s.groups) {
        if (s.groups < 8 || s.groups % 8) throw Error(KMSD_EDEVICE, "internal error: pinned workgroup count");
        sh.gx = s.groups;
    }
    a = StepArgs{};
    uint32_t fl = 0;
    Cplx alpha = s.alpha;
    const bool bsr = A && A->fmt == KMSD_BSR4;
    if (A) {
        fl |= F_SPMM;
        a.ptr = A->ptr;
        a.idx = A->idx;
        a.ell = bsr ? 0 : A->ell;
        if (a.ell) fl |= F_ELL;
        if (!bsr && A->has_diag) fl |= F_DIAG;
        a.val = f32 ? A->val32 : A->val64;
        a.dval = f32 ? A->dval32 : A->dval64;
        a.uniform_slot = (!bsr && A->ell) ? A->uniform_slot : -1;
        if (bsr && A->plane && c->tile_bytes > 0) {
            const int64_t rows = c->tile_bytes / (4 * int64_t(s.w) * int64_t(elem_bytes(dtype)));
            if (rows >= 1 && rows < A->plane) {
                a.tile_plane = uint32_t(A->plane);
                a.tile_rows = uint32_t(rows);
            }
        }
        if (s.VOUT) {
            fl |= F_VOUT;
            a.vval = s.vval;
            a.VOUT = s.VOUT;
            a.ldv = s.ldv;
            a.valpha_re = s.valpha.re;
            a.valpha_im = s.valpha.im;
        }
        if (A->vkind == VK_IMAG) alpha = Cplx{-s.alpha.im, s.alpha.re};  // A = i * stored
    }
    if (s.use_beta) fl |= F_BETA;
    if (s.use_p) fl |= F_P;
    if (s.OUT) fl |= F_STORE;
    if (s.rhs) {
        if ((bsr && !s.bfac32) || (!bsr && !f32) || !s.W || s.ACC || s.doubling || s.norm || s.VOUT)
            throw Error(KMSD_EDEVICE, "internal error: a right-hand-side operand needs a CSR/ELL or block-Jacobi step");
        fl |= F_RHS;
    } else if (s.W) {
        fl |= F_MOM;
    }
    if (s.doubling) fl |= F_MOM2;
    if (s.norm) {
        if (bsr || s.W || s.doubling || !s.OUT)
            throw Error(KMSD_EDEVICE, "internal error: norm partials need a CSR/ELL step with a stored result");
        fl |= F_NORM;
    }
    if (s.ACC) fl |= F_ACC;
    if (bsr && f32 && A->vkind != VK_COMPLEX && (c->bsr_mfma || s.bfac32)) mfma_shape(c, s, A, fl, a, sh, fn, key);
    if (s.bfac32 && !key.blk)
        throw Error(KMSD_EDEVICE, "internal error: a block-Jacobi product needs the matrix-core 4x4-block kernel");
    if (bsr && !f32 && A->vkind != VK_COMPLEX && c->bsr_lds) lds_shape(c, dtype, s, A, fl, a, sh, fn, key);
    if (!s.groups && s.max_groups) {
        if (s.max_groups < 8 || s.max_groups % 8) throw Error(KMSD_EDEVICE, "internal error: workgroup cap");
        sh.gx = std::min<int64_t>(sh.gx, s.max_groups);
    }
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
    a.gate = s.gate;
}

using GroupFn = void (*)(const GroupJob *, int32_t, int32_t, int32_t);

template <class VT, class AT, int CPL>
GroupFn pick_group(int ell) {
    if (ell == 3) return step_group_kernel<VT, double, false, AT, CPL, 3>;
    if (ell == 4) return step_group_kernel<VT, double, false, AT, CPL, 4>;
    if (ell == 6) return step_group_kernel<VT, double, false, AT, CPL, 6>;
    return step_group_kernel<VT, double, false, AT, CPL, 0>;
}

// Grouped form of a step's kernel: FP64 real (or imaginary) matrix values,
// CSR/ELL storage, no fused second product; null otherwise.
GroupFn select_group_kernel(const kmsd_ctx *c, const KernelKey &k) {
    if (k.bsr || k.cv || k.vout || k.f32 || c->fp32_math) return nullptr;
    if (k.dtype == KMSD_C128) return k.cpl == 2 ? pick_group<double, double, 2>(k.ell) : pick_group<double, double, 1>(k.ell);
    if (k.cpl == 4) return pick_group<float, double, 4>(k.ell);
    return k.cpl == 2 ? pick_group<float, double, 2>(k.ell) : pick_group<float, double, 1>(k.ell);
}

}  // namespace

int64_t launch_step(kmsd_ctx *c, int dtype, const Step &s) {
    if (s.n < 1 || s.w < 1) return 0;
    StepArgs a{};
    Shape sh;
    KernelFn fn = nullptr;
    KernelKey key;
    prepare_step(c, dtype, s, a, sh, fn, key);
    void *args[] = {&a};
    HIPCHECK(hipLaunchKernel(reinterpret_cast<const void *>(fn), dim3(unsigned(sh.gx), unsigned(sh.gy)), dim3(256),
                             args, 0, c->stream));
    return sh.gx;
}

bool group_prepare(kmsd_ctx *c, int dtype, const std::vector<GroupChain> &chains, GroupPlan &plan) {
    if (chains.empty()) return false;
    std::vector<GroupJob> jobs(chains.size());
    GroupFn fn = nullptr;
    int gy = 0;
    uint32_t wg = 0;
    for (size_t q = 0; q < chains.size(); ++q) {
        const GroupChain &ch = chains[q];
        if (ch.step.groups < 8 || ch.step.groups % 8) return false;  // the chain's own pinned workgroup count
        Shape sh;
        KernelFn single = nullptr;
        KernelKey key;
        prepare_step(c, dtype, ch.step, jobs[q].a, sh, single, key);
        const GroupFn g = select_group_kernel(c, key);
        if (!g || (fn && g != fn) || (gy && sh.gy != gy)) return false;
        fn = g;
        gy = sh.gy;
        jobs[q].buf[0] = ch.buf[0];
        jobs[q].buf[1] = ch.buf[1];
        jobs[q].part0 = ch.part0;
        jobs[q].slot_doubles = ch.slot_doubles;
        jobs[q].wg0 = wg;
        jobs[q].wgn = uint32_t(ch.step.groups);
        wg += uint32_t(ch.step.groups);
    }
    void *d = plan.table.ensure(c, sizeof(GroupJob) * jobs.size());
    copy_h2d(c, d, jobs.data(), sizeof(GroupJob) * jobs.size());
    plan.fn = reinterpret_cast<const void *>(fn);
    plan.gy = gy;
    plan.total = wg;
    plan.njobs = int32_t(jobs.size());
    return true;
}

void group_launch(kmsd_ctx *c, const GroupPlan &plan, int parity, int slot) {
    const GroupJob *jobs = static_cast<const GroupJob *>(plan.table.p);
    int32_t nj = plan.njobs, par = parity, sl = slot;
    void *args[] = {&jobs, &nj, &par, &sl};
    HIPCHECK(hipLaunchKernel(plan.fn, dim3(unsigned(plan.total), unsigned(plan.gy)), dim3(256), args, 0, c->stream));
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

const int32_t *launch_convert_check(kmsd_ctx *c, int din, const void *X, int64_t ldx, int dout, void *O, int64_t ldo,
                                    int64_t n, int32_t w) {
    const int64_t g = blocks_for(n * w);
    int32_t *flag = c->flag + 1 + c->cvt_slot, *clear = c->flag + 2 - c->cvt_slot;
    c->cvt_slot ^= 1;
    if (din == KMSD_C128 && dout == KMSD_C64)
        hipLaunchKernelGGL((convert_check_kernel<double, float>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx,
                           O, ldo, n, w, flag, clear);
    else if (din == KMSD_C64 && dout == KMSD_C128)
        hipLaunchKernelGGL((convert_check_kernel<float, double>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx,
                           O, ldo, n, w, flag, clear);
    else if (din == KMSD_C64)
        hipLaunchKernelGGL((convert_check_kernel<float, float>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx,
                           O, ldo, n, w, flag, clear);
    else
        hipLaunchKernelGGL((convert_check_kernel<double, double>), dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx,
                           O, ldo, n, w, flag, clear);
    HIPCHECK(hipGetLastError());
    return flag;
}

bool launch_nonfinite(kmsd_ctx *c, int dtype, const void *X, int64_t ldx, int64_t n, int32_t w) {
    int32_t got = 0;
    HIPCHECK(hipMemsetAsync(c->flag, 0, sizeof(int32_t), c->stream));
    const int64_t g = blocks_for(n * w);
    if (dtype == KMSD_C128)
        hipLaunchKernelGGL(nonfinite_kernel<double>, dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, n, w,
                           c->flag);
    else
        hipLaunchKernelGGL(nonfinite_kernel<float>, dim3(unsigned(g)), dim3(256), 0, c->stream, X, ldx, n, w,
                           c->flag);
    HIPCHECK(hipGetLastError());
    read_status(c, {{&got, c->flag, sizeof(got)}});
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

kmsd_matrix *matrix_congruence_copy(kmsd_ctx *c, const kmsd_matrix *S, const double *diag) {
    kmsd_matrix *W = new kmsd_matrix(*S);
    W->bytes = 0;
    W->uniform_slot = -1;  // the scaled values are no longer one value
    W->ptr = W->idx = nullptr;
    W->val64 = W->val32 = W->dval64 = W->dval32 = nullptr;
    try {
        const int comps = S->vkind == VK_COMPLEX ? 2 : 1;
        const int64_t rows = S->n;
        auto dup = [&](const void *src, size_t bytes) {
            void *d = dalloc(c, bytes);
            W->bytes += bytes;
            HIPCHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, c->stream));
            return d;
        };
        const size_t nv = size_t(std::max<int64_t>(1, S->noff)) * comps;
        if (S->fmt == KMSD_BSR4) {
            const int64_t nq = rows / 4, nnzb = S->noff / 16;
            W->ptr = static_cast<int32_t *>(dup(S->ptr, sizeof(int32_t) * size_t(nq + 1)));
            W->idx = static_cast<int32_t *>(dup(S->idx, sizeof(int32_t) * size_t(std::max<int64_t>(1, nnzb))));
            W->val64 = dup(S->val64, sizeof(double) * nv);
            W->val32 = dup(S->val32, sizeof(float) * nv);
            const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((S->noff + 255) / 256, 8192)));
            if (S->vkind == VK_COMPLEX) {
                hipLaunchKernelGGL((bsr4_congruence_kernel<double, true>), dim3(g), dim3(256), 0, c->stream, W->val64,
                                   W->ptr, W->idx, diag, nq, nnzb);
                hipLaunchKernelGGL((bsr4_congruence_kernel<float, true>), dim3(g), dim3(256), 0, c->stream, W->val32,
                                   W->ptr, W->idx, diag, nq, nnzb);
            } else {
                hipLaunchKernelGGL((bsr4_congruence_kernel<double, false>), dim3(g), dim3(256), 0, c->stream,
                                   W->val64, W->ptr, W->idx, diag, nq, nnzb);
                hipLaunchKernelGGL((bsr4_congruence_kernel<float, false>), dim3(g), dim3(256), 0, c->stream,
                                   W->val32, W->ptr, W->idx, diag, nq, nnzb);
            }
            HIPCHECK(hipGetLastError());
            W->identity = false;
            HIPCHECK(hipStreamSynchronize(c->stream));
            return W;
        }
        if (S->ptr) W->ptr = static_cast<int32_t *>(dup(S->ptr, sizeof(int32_t) * size_t(rows + 1)));
        W->idx = static_cast<int32_t *>(dup(S->idx, sizeof(int32_t) * size_t(std::max<int64_t>(1, S->noff))));
        W->val64 = dup(S->val64, sizeof(double) * nv);
        W->val32 = dup(S->val32, sizeof(float) * nv);
        W->dval64 = dup(S->dval64, sizeof(double) * size_t(rows) * comps);
        W->dval32 = dup(S->dval32, sizeof(float) * size_t(rows) * comps);
        const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((rows + 255) / 256, 4096)));
        if (S->vkind == VK_COMPLEX) {
            hipLaunchKernelGGL((congruence_scale_kernel<double, true>), dim3(g), dim3(256), 0, c->stream, W->val64,
                               W->dval64, W->ptr, W->idx, W->ell, diag, rows);
            hipLaunchKernelGGL((congruence_scale_kernel<float, true>), dim3(g), dim3(256), 0, c->stream, W->val32,
                               W->dval32, W->ptr, W->idx, W->ell, diag, rows);
        } else {
            hipLaunchKernelGGL((congruence_scale_kernel<double, false>), dim3(g), dim3(256), 0, c->stream, W->val64,
                               W->dval64, W->ptr, W->idx, W->ell, diag, rows);
            hipLaunchKernelGGL((congruence_scale_kernel<float, false>), dim3(g), dim3(256), 0, c->stream, W->val32,
                               W->dval32, W->ptr, W->idx, W->ell, diag, rows);
        }
        HIPCHECK(hipGetLastError());
        W->identity = false;
        W->has_diag = true;
        HIPCHECK(hipStreamSynchronize(c->stream));
    } catch (...) {
        matrix_release(W);
        delete W;
        throw;
    }
    return W;
}

void launch_pixel_transpose(kmsd_ctx *c, const void *in, void *out, int64_t R, int64_t C, int64_t words) {
    if (R < 1 || C < 1) return;
    const uint64_t *I = static_cast<const uint64_t *>(in);
    uint64_t *O = static_cast<uint64_t *>(out);
    if (words <= 4) {
        const int64_t tiles = ((C + 31) / 32) * ((R + 31) / 32);
        if (tiles >= (int64_t(1) << 31)) throw Error(KMSD_EINVAL, "transpose grid too large");
        const dim3 grid{unsigned(tiles)};
        switch (words) {
            case 1: hipLaunchKernelGGL(pixel_transpose_tiled<1>, grid, dim3(256), 0, c->stream, I, O, R, C); break;
            case 2: hipLaunchKernelGGL(pixel_transpose_tiled<2>, grid, dim3(256), 0, c->stream, I, O, R, C); break;
            case 3: hipLaunchKernelGGL(pixel_transpose_tiled<3>, grid, dim3(256), 0, c->stream, I, O, R, C); break;
            default: hipLaunchKernelGGL(pixel_transpose_tiled<4>, grid, dim3(256), 0, c->stream, I, O, R, C);
        }
    } else {
        const int64_t g = blocks_for(R * C * words);
        hipLaunchKernelGGL(pixel_transpose_direct, dim3(unsigned(g)), dim3(256), 0, c->stream, I, O, R, C, words);
    }
    HIPCHECK(hipGetLastError());
}

void launch_doubling_finish(kmsd_ctx *c, const double *ab, double *mu, int64_t L, int32_t w) {
    const int64_t g = blocks_for(L * w);
    hipLaunchKernelGGL(doubling_finish_kernel, dim3(unsigned(g)), dim3(256), 0, c->stream, ab, mu, L, w);
    HIPCHECK(hipGetLastError());
}

void launch_sample_finish(kmsd_ctx *c, int dtype, void *out, const void *xi, int64_t ld, int64_t n, int32_t w,
                          const double *diag, double c0) {
    const int64_t g = blocks_for(n * w);
    if (dtype == KMSD_C128)
        hipLaunchKernelGGL(sample_finish_kernel<double>, dim3(unsigned(g)), dim3(256), 0, c->stream, out, xi, ld, n,
                           w, diag, c0);
    else
        hipLaunchKernelGGL(sample_finish_kernel<float>, dim3(unsigned(g)), dim3(256), 0, c->stream, out, xi, ld, n, w,
                           diag, c0);
    HIPCHECK(hipGetLastError());
}

}  // namespace kmsd
