// Metric preconditioners beyond the diagonal and small atom blocks
// (reference Metric::precondition and its set-up, cpp/native.cpp:92-121,
// 163-205, 237-297):
//   - cluster blocks: dense Cholesky factors of arbitrary row sets;
//   - blocks of any size: blocks of more than SMALL_BLOCK rows are solved one
//     workgroup per (block, column), which also carries the direct solve with
//     one full-size factor;
//   - IC(0): the incomplete factor itself (kmsd_ic0_factor) and its two
//     triangular solves, both level-scheduled: rows whose dependencies are
//     complete run together, one launch per level.
// The IC kernels do the reference's complex arithmetic in its order with no
// contraction into FMAs (each row's sums run over its entries in the
// reference's order), so the factor and each solved entry are the
// reference's to the last bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>

#include "kmsd_internal.hpp"
#include "kmsd_pcg_kernels.hpp"

using namespace kmsd;

struct kmsd_metric::IcData {
    int64_t n = 0, nnz = 0;
    int64_t *ptr = nullptr;   // device int64[n+1]: rows of L, diagonal last
    int32_t *idx = nullptr;   // device int32[nnz]
    double *val = nullptr;    // device cx<double>[nnz]
    int64_t *uptr = nullptr;  // device int64[n+1]: column j of L as the row j of L^H ...
    int32_t *urow = nullptr;  // ... its rows i > j, in decreasing i
    double *uval = nullptr;   // ... and the entries L[i][j] (conjugated at use)
    int32_t *perm = nullptr;  // device int32[n]: original row of factor row i
    double *scale = nullptr;  // device double[n]: D^-1 per original row
    int32_t *fwd = nullptr;   // factor rows by forward level
    int32_t *bwd = nullptr;   // factor rows by backward level
    std::vector<int64_t> fwd_lev, bwd_lev;  // level boundaries into fwd / bwd
    std::vector<std::pair<void *, size_t>> allocs;
    DBuf work;  // n x w complex128 solve workspace
    // The two sweeps are one launch per level; they are captured once per
    // argument set as a HIP graph and replayed (same kernels, same
    // arguments, bitwise the same result as eager launches).
    struct Graph {
        int dtype;
        int32_t w;
        const void *R, *Z, *W, *fl;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    ~IcData() {
        for (Graph &g : graphs) {
            (void)hipGraphExecDestroy(g.exec);
            (void)hipGraphDestroy(g.graph);
        }
    }
};

namespace {

constexpr int IC_CHUNK = 8;  // factor entries whose loads an IC solve thread issues together

#pragma clang fp contract(off)

// Blocks of more than SMALL_BLOCK rows: one workgroup per (block, column).
// Forward L v = r reads row i of L (contiguous) against the solved entries
// with a workgroup sum; backward L^H z = v subtracts each solved entry times
// row i of L (contiguous) from the entries above it.  The block's column
// lives in Z between the two passes.
template <class VT>
__global__ __launch_bounds__(256) void large_block_kernel(const void *Rv, void *Zv, int64_t w, const int32_t *bptr,
                                                          const int32_t *rows, const int64_t *foff, const double *fac,
                                                          const int32_t *list, const PcgFlags *fl) {
    if (fl && fl->broke) return;
    const cx<VT> *R = static_cast<const cx<VT> *>(Rv);
    cx<VT> *Z = static_cast<cx<VT> *>(Zv);
    const int64_t q = list[blockIdx.x / w], j = blockIdx.x % w;
    const int64_t a0 = bptr[q], m = bptr[q + 1] - a0;
    const cx<double> *L = reinterpret_cast<const cx<double> *>(fac) + foff[q];
    const int32_t *rq = rows ? rows + a0 : nullptr;
    const auto at = [&](int64_t i) -> int64_t { return (rq ? int64_t(rq[i]) : a0 + i) * w + j; };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double red[2][4];
    __shared__ double piv[2];
    for (int64_t i = 0; i < m; ++i) {
        double sre = 0.0, sim = 0.0;
        const cx<double> *Li = L + i * m;
        for (int64_t k = tid; k < i; k += 256) {
            const cx<double> l = Li[k], u = to<double>(Z[at(k)]);
            sre += l.re * u.re - l.im * u.im;
            sim += l.re * u.im + l.im * u.re;
        }
        for (int off = 32; off > 0; off >>= 1) {
            sre += __shfl_xor(sre, off);
            sim += __shfl_xor(sim, off);
        }
        if (lane == 0) {
            red[0][wave] = sre;
            red[1][wave] = sim;
        }
        __syncthreads();
        if (tid == 0) {
            const cx<double> r = to<double>(R[at(i)]);
            const double s_re = r.re - (((red[0][0] + red[0][1]) + red[0][2]) + red[0][3]);
            const double s_im = r.im - (((red[1][0] + red[1][1]) + red[1][2]) + red[1][3]);
            const cx<double> d = Li[i];
            const double den = d.re * d.re + d.im * d.im;
            Z[at(i)] = to<VT>(cx<double>{(s_re * d.re + s_im * d.im) / den, (s_im * d.re - s_re * d.im) / den});
        }
        __syncthreads();
    }
    for (int64_t i = m - 1; i >= 0; --i) {
        const cx<double> *Li = L + i * m;
        if (tid == 0) {
            const cx<double> s = to<double>(Z[at(i)]), d = Li[i];  // divide by conj(d)
            const double den = d.re * d.re + d.im * d.im;
            const cx<double> v{(s.re * d.re - s.im * d.im) / den, (s.im * d.re + s.re * d.im) / den};
            Z[at(i)] = to<VT>(v);
            piv[0] = v.re;
            piv[1] = v.im;
        }
        __syncthreads();
        const double vre = piv[0], vim = piv[1];
        for (int64_t k = tid; k < i; k += 256) {
            const cx<double> l = Li[k];  // conj(L[i][k]) * v
            cx<double> z = to<double>(Z[at(k)]);
            z.re -= l.re * vre + l.im * vim;
            z.im -= l.re * vim - l.im * vre;
            Z[at(k)] = to<VT>(z);
        }
        __syncthreads();
    }
}

// IC(0) forward solve of one level: W[i] = (D^-1 r)[perm[i]] - sum_k L[i][k] W[k], / L[i][i].
template <class VT>
__global__ __launch_bounds__(256) void ic_fwd_kernel(const void *Rv, int64_t w, const int32_t *rows, int64_t count,
                                                     const int64_t *ptr, const int32_t *idx, const cx<double> *val,
                                                     const int32_t *perm, const double *scale, cx<double> *W,
                                                     const PcgFlags *fl) {
    if (fl && fl->broke) return;
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= count * w) return;
    const cx<VT> *R = static_cast<const cx<VT> *>(Rv);
    const int64_t r = rows[t / w], j = t % w;
    const int64_t o = perm[r], d = ptr[r + 1] - 1;
    const cx<double> x = to<double>(R[o * w + j]);
    const double s = scale[o];
    cx<double> v{s * x.re, s * x.im};
    // Entries in chunks of IC_CHUNK: all loads of a chunk are issued before
    // its products, which are then subtracted in entry order.
    for (int64_t k0 = ptr[r]; k0 < d; k0 += IC_CHUNK) {
        cx<double> l[IC_CHUNK], u[IC_CHUNK];
#pragma unroll
        for (int q = 0; q < IC_CHUNK; ++q) {
            const int64_t k = k0 + q < d ? k0 + q : d;  // past the row: the diagonal, unused
            l[q] = val[k];
            u[q] = W[int64_t(k0 + q < d ? idx[k] : 0) * w + j];
        }
#pragma unroll
        for (int q = 0; q < IC_CHUNK; ++q)
            if (k0 + q < d) {
                v.re -= l[q].re * u[q].re - l[q].im * u[q].im;
                v.im -= l[q].re * u[q].im + l[q].im * u[q].re;
            }
    }
    const double dd = val[d].re;  // the factor's diagonal is real
    W[r * w + j] = cx<double>{v.re / dd, v.im / dd};
}

// IC(0) backward solve of one level: W[j] -= conj(L[i][j]) W[i] over the rows
// i > j of column j in decreasing i, / L[j][j]; z[perm[j]] = D^-1 W[j].
template <class VT>
__global__ __launch_bounds__(256) void ic_bwd_kernel(void *Zv, int64_t w, const int32_t *rows, int64_t count,
                                                     const int64_t *ptr, const cx<double> *val, const int64_t *uptr,
                                                     const int32_t *urow, const cx<double> *uval, const int32_t *perm,
                                                     const double *scale, cx<double> *W, const PcgFlags *fl) {
    if (fl && fl->broke) return;
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= count * w) return;
    cx<VT> *Z = static_cast<cx<VT> *>(Zv);
    const int64_t r = rows[t / w], j = t % w;
    cx<double> v = W[r * w + j];
    const int64_t e = uptr[r + 1];
    for (int64_t k0 = uptr[r]; k0 < e; k0 += IC_CHUNK) {
        cx<double> l[IC_CHUNK], x[IC_CHUNK];
#pragma unroll
        for (int q = 0; q < IC_CHUNK; ++q) {
            const bool in = k0 + q < e;
            l[q] = in ? uval[k0 + q] : cx<double>{0.0, 0.0};
            x[q] = W[int64_t(in ? urow[k0 + q] : r) * w + j];
        }
#pragma unroll
        for (int q = 0; q < IC_CHUNK; ++q)
            if (k0 + q < e) {
                v.re -= l[q].re * x[q].re + l[q].im * x[q].im;
                v.im -= l[q].re * x[q].im - l[q].im * x[q].re;
            }
    }
    const double dd = val[ptr[r + 1] - 1].re;
    v = cx<double>{v.re / dd, v.im / dd};
    W[r * w + j] = v;
    const int64_t o = perm[r];
    const double s = scale[o];
    Z[o * w + j] = to<VT>(cx<double>{s * v.re, s * v.im});
}

// IC(0) factor rows of one level (cpp/native.cpp:92-121): entry (i,j) minus
// the products of row i's and row j's earlier entries of equal column, over
// row j's diagonal; the pivot is the diagonal plus the shift minus the row's
// squared moduli.  A failing pivot records its row (the smallest wins, which
// is the first row a sequential factorization reaches).
__global__ __launch_bounds__(256) void ic_factor_kernel(const int32_t *rows, int64_t count, const int64_t *ptr,
                                                        const int32_t *idx, const cx<double> *a, cx<double> *f,
                                                        double shift, unsigned long long *fail) {
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= count) return;
    const int64_t i = rows[t], b = ptr[i], e = ptr[i + 1];
    for (int64_t ij = b; ij + 1 < e; ++ij) {
        const int64_t j = idx[ij];
        cx<double> v = a[ij];
        int64_t p = b, q = ptr[j];
        const int64_t qd = ptr[j + 1] - 1;
        while (p < ij && q < qd) {
            const int32_t cp = idx[p], cq = idx[q];
            if (cp < cq) {
                ++p;
            } else if (cq < cp) {
                ++q;
            } else {
                const cx<double> x = f[p], y = f[q];  // x conj(y)
                v.re -= x.re * y.re + x.im * y.im;
                v.im -= x.im * y.re - x.re * y.im;
                ++p;
                ++q;
            }
        }
        const double dd = f[qd].re;
        f[ij] = cx<double>{v.re / dd, v.im / dd};
    }
    double pre = a[e - 1].re + shift;
    const double pim = a[e - 1].im;
    for (int64_t k = b; k + 1 < e; ++k) pre -= f[k].re * f[k].re + f[k].im * f[k].im;
    if (!__builtin_isfinite(pre) || !__builtin_isfinite(pim) || fabs(pim) > 1e-12 || pre <= 1e-14)
        atomicMin(fail, static_cast<unsigned long long>(i));
    f[e - 1] = cx<double>{sqrt(pre), 0.0};
}

#pragma clang fp contract(on)

// Rows [0, n) sorted by dependency level: level(i) = 1 + max level of the
// rows it reads (deps(i) lists them).  Returns the rows and level bounds.
template <class Deps>
void levels(int64_t n, Deps &&deps, std::vector<int32_t> &order, std::vector<int64_t> &bounds) {
    std::vector<int32_t> lev(size_t(n), 0);
    int32_t top = -1;
    for (int64_t i = 0; i < n; ++i) {
        int32_t l = 0;
        deps(i, [&](int64_t k) { l = std::max(l, lev[size_t(k)] + 1); });
        lev[size_t(i)] = l;
        top = std::max(top, l);
    }
    bounds.assign(size_t(top + 2), 0);
    for (int64_t i = 0; i < n; ++i) ++bounds[size_t(lev[size_t(i)]) + 1];
    for (size_t l = 1; l < bounds.size(); ++l) bounds[l] += bounds[l - 1];
    std::vector<int64_t> fill(bounds.begin(), bounds.end() - 1);
    order.assign(size_t(n), 0);
    for (int64_t i = 0; i < n; ++i) order[size_t(fill[size_t(lev[size_t(i)])]++)] = int32_t(i);
}

template <class T>
T *upload(kmsd_ctx *c, const T *host, size_t count, std::vector<std::pair<void *, size_t>> *track) {
    T *d = static_cast<T *>(dalloc(c, sizeof(T) * std::max<size_t>(count, 1)));
    if (track) track->push_back({d, sizeof(T) * std::max<size_t>(count, 1)});
    if (count) copy_h2d(c, d, host, sizeof(T) * count);
    return d;
}

unsigned grid_for(int64_t threads) { return unsigned((threads + 255) / 256); }

}  // namespace

namespace kmsd {

void metric_release_blocks(kmsd_metric *m) {
    kmsd_ctx *c = m->ctx;
    for (void *p : {static_cast<void *>(m->block_ptr), static_cast<void *>(m->factor_off),
                    static_cast<void *>(m->factors), static_cast<void *>(m->block_rows),
                    static_cast<void *>(m->large), static_cast<void *>(m->factors32)})
        if (p) (void)hipFree(p);
    c->used -= std::min<uint64_t>(c->used, m->block_bytes);
    m->block_ptr = m->block_rows = m->large = nullptr;
    m->factor_off = nullptr;
    m->factors = nullptr;
    m->factors32 = nullptr;
    m->block_bytes = 0;
    m->nblocks = m->nlarge = 0;
    m->max_block = 0;
    m->uniform4 = false;
    m->host_block_ptr.clear();
}

void metric_upload_blocks(kmsd_metric *m, int64_t nblocks, const int64_t *off, const int64_t *rows,
                          const double *factors) {
    kmsd_ctx *c = m->ctx;
    HIPCHECK(hipStreamSynchronize(c->stream));
    metric_release_blocks(m);
    std::vector<int32_t> bp(size_t(nblocks) + 1), big;
    std::vector<int64_t> fo(static_cast<size_t>(nblocks));
    size_t nfac = 0;
    bool uniform4 = !rows && nblocks > 0;
    int32_t maxb = 0;
    for (int64_t q = 0; q < nblocks; ++q) {
        const int64_t mq = off[q + 1] - off[q];
        bp[size_t(q)] = int32_t(off[q]);
        fo[size_t(q)] = int64_t(nfac);
        nfac += size_t(mq * mq);
        uniform4 = uniform4 && mq == 4;
        if (mq > SMALL_BLOCK) big.push_back(int32_t(q));
        else maxb = std::max<int32_t>(maxb, int32_t(mq));
    }
    bp[size_t(nblocks)] = int32_t(off[nblocks]);
    std::vector<std::pair<void *, size_t>> track;
    try {
        m->block_ptr = upload(c, bp.data(), bp.size(), &track);
        m->factor_off = upload(c, fo.data(), fo.size(), &track);
        m->factors = upload(c, factors, 2 * nfac, &track);
        if (rows) {
            std::vector<int32_t> r32(rows, rows + off[nblocks]);
            m->block_rows = upload(c, r32.data(), r32.size(), &track);
        }
        if (!big.empty()) m->large = upload(c, big.data(), big.size(), &track);
        if (uniform4) {
            // real factors (the usual case: real S): an FP32 copy with the
            // diagonal inverted for the fused complex64 iterations
            bool real = true;
            for (size_t e = 0; e < nfac && real; ++e) real = factors[2 * e + 1] == 0.0;
            if (real) {
                std::vector<float> f32(nfac, 0.0f);
                for (int64_t q = 0; q < nblocks; ++q)
                    for (int i = 0; i < 4; ++i)
                        for (int k = 0; k <= i; ++k) {
                            const double v = factors[2 * (size_t(q) * 16 + size_t(i * 4 + k))];
                            f32[size_t(q) * 16 + size_t(i * 4 + k)] = float(k == i ? 1.0 / v : v);
                        }
                m->factors32 = upload(c, f32.data(), f32.size(), &track);
            }
        }
        HIPCHECK(hipStreamSynchronize(c->stream));
    } catch (...) {
        for (auto &a : track) dfree(c, a.first, a.second);
        m->block_ptr = m->block_rows = m->large = nullptr;
        m->factor_off = nullptr;
        m->factors = nullptr;
        m->factors32 = nullptr;
        throw;
    }
    for (auto &a : track) m->block_bytes += a.second;
    m->nblocks = nblocks;
    m->nlarge = int64_t(big.size());
    m->max_block = maxb;
    m->uniform4 = uniform4;
    m->host_block_ptr.assign(off, off + nblocks + 1);
}

void metric_release_ic(kmsd_metric *m) {
    if (!m->ic) return;
    for (auto &a : m->ic->allocs) dfree(m->ctx, a.first, a.second);
    delete m->ic;
    m->ic = nullptr;
}

namespace {

void ic_sweeps(kmsd_ctx *c, kmsd_metric::IcData &ic, int dtype, const void *R, void *Z, int32_t w, cx<double> *W,
               const PcgFlags *fl);

}  // namespace

void ic_apply(kmsd_metric *m, int dtype, const void *R, void *Z, int32_t w, const void *flags) {
    kmsd_ctx *c = m->ctx;
    kmsd_metric::IcData &ic = *m->ic;
    const PcgFlags *fl = static_cast<const PcgFlags *>(flags);
    cx<double> *W = static_cast<cx<double> *>(ic.work.ensure(c, size_t(ic.n) * size_t(w) * 16));
    if (!c->graphs) {
        ic_sweeps(c, ic, dtype, R, Z, w, W, fl);
        return;
    }
    for (const auto &g : ic.graphs)
        if (g.dtype == dtype && g.w == w && g.R == R && g.Z == Z && g.W == W && g.fl == fl) {
            HIPCHECK(hipGraphLaunch(g.exec, c->stream));
            return;
        }
    if (!c->capture) HIPCHECK(hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking));
    hipStream_t user = c->stream;
    c->stream = c->capture;
    HIPCHECK_RESTORE(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), c, user);
    try {
        ic_sweeps(c, ic, dtype, R, Z, w, W, fl);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c->stream = user;
        throw;
    }
    kmsd_metric::IcData::Graph g{dtype, w, R, Z, W, fl, nullptr, nullptr};
    const hipError_t ec = hipStreamEndCapture(c->stream, &g.graph);
    c->stream = user;
    HIPCHECK(ec);
    const hipError_t ei = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g.graph);
        HIPCHECK(ei);
    }
    if (ic.graphs.size() >= 8) {  // a few argument sets per metric (dtype, width, workspace)
        HIPCHECK(hipStreamSynchronize(c->stream));
        (void)hipGraphExecDestroy(ic.graphs.front().exec);
        (void)hipGraphDestroy(ic.graphs.front().graph);
        ic.graphs.erase(ic.graphs.begin());
    }
    ic.graphs.push_back(g);
    HIPCHECK(hipGraphLaunch(g.exec, c->stream));
}

namespace {

void ic_sweeps(kmsd_ctx *c, kmsd_metric::IcData &ic, int dtype, const void *R, void *Z, int32_t w, cx<double> *W,
               const PcgFlags *fl) {
    const cx<double> *val = reinterpret_cast<const cx<double> *>(ic.val);
    const cx<double> *uval = reinterpret_cast<const cx<double> *>(ic.uval);
    for (size_t l = 0; l + 1 < ic.fwd_lev.size(); ++l) {
        const int64_t cnt = ic.fwd_lev[l + 1] - ic.fwd_lev[l];
        const int32_t *rows = ic.fwd + ic.fwd_lev[l];
        if (dtype == KMSD_C128)
            hipLaunchKernelGGL(ic_fwd_kernel<double>, dim3(grid_for(cnt * w)), dim3(256), 0, c->stream, R, int64_t(w),
                               rows, cnt, ic.ptr, ic.idx, val, ic.perm, ic.scale, W, fl);
        else
            hipLaunchKernelGGL(ic_fwd_kernel<float>, dim3(grid_for(cnt * w)), dim3(256), 0, c->stream, R, int64_t(w),
                               rows, cnt, ic.ptr, ic.idx, val, ic.perm, ic.scale, W, fl);
    }
    for (size_t l = 0; l + 1 < ic.bwd_lev.size(); ++l) {
        const int64_t cnt = ic.bwd_lev[l + 1] - ic.bwd_lev[l];
        const int32_t *rows = ic.bwd + ic.bwd_lev[l];
        if (dtype == KMSD_C128)
            hipLaunchKernelGGL(ic_bwd_kernel<double>, dim3(grid_for(cnt * w)), dim3(256), 0, c->stream, Z, int64_t(w),
                               rows, cnt, ic.ptr, val, ic.uptr, ic.urow, uval, ic.perm, ic.scale, W, fl);
        else
            hipLaunchKernelGGL(ic_bwd_kernel<float>, dim3(grid_for(cnt * w)), dim3(256), 0, c->stream, Z, int64_t(w),
                               rows, cnt, ic.ptr, val, ic.uptr, ic.urow, uval, ic.perm, ic.scale, W, fl);
    }
    HIPCHECK(hipGetLastError());
}

}  // namespace

void large_blocks_apply(kmsd_metric *m, int dtype, const void *R, void *Z, int32_t w, const void *flags) {
    if (!m->nlarge) return;
    kmsd_ctx *c = m->ctx;
    const PcgFlags *fl = static_cast<const PcgFlags *>(flags);
    const dim3 grid(unsigned(m->nlarge * w));
    if (dtype == KMSD_C128)
        hipLaunchKernelGGL(large_block_kernel<double>, grid, dim3(256), 0, c->stream, R, Z, int64_t(w), m->block_ptr,
                           m->block_rows, m->factor_off, m->factors, m->large, fl);
    else
        hipLaunchKernelGGL(large_block_kernel<float>, grid, dim3(256), 0, c->stream, R, Z, int64_t(w), m->block_ptr,
                           m->block_rows, m->factor_off, m->factors, m->large, fl);
    HIPCHECK(hipGetLastError());
}

}  // namespace kmsd

namespace {

template <class VT>
__global__ __launch_bounds__(256) void diag_precond_kernel(const void *Rv, void *Zv, int64_t n, int64_t w,
                                                           const double *d) {
    const cx<VT> *R = static_cast<const cx<VT> *>(Rv);
    cx<VT> *Z = static_cast<cx<VT> *>(Zv);
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < n * w; e += int64_t(gridDim.x) * 256) {
        const cx<double> r = to<double>(R[e]);
        const double di = d[e / w];
        Z[e] = to<VT>(cx<double>{r.re / di, r.im / di});
    }
}

}  // namespace

extern "C" int kmsd_metric_precondition(kmsd_metric *m, const void *R, void *Z, int64_t nb, int dtype) {
    if (!m) return KMSD_EINVAL;
    kmsd_ctx *c = m->ctx;
    return guard(c, [&] {
        check_block_args(R, Z, nb, dtype);
        if (R == Z) throw Error(KMSD_EINVAL, "preconditioner input and output must not overlap");
        const int32_t w = int32_t(nb);
        const int64_t n = m->n;
        if (m->ic) {
            ic_apply(m, dtype, R, Z, w, nullptr);
        } else if (m->nblocks == 0) {
            const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((n * w + 255) / 256, 8192)));
            if (dtype == KMSD_C128)
                hipLaunchKernelGGL(diag_precond_kernel<double>, dim3(g), dim3(256), 0, c->stream, R, Z, n, int64_t(w),
                                   m->diag);
            else
                hipLaunchKernelGGL(diag_precond_kernel<float>, dim3(g), dim3(256), 0, c->stream, R, Z, n, int64_t(w),
                                   m->diag);
        } else {
            if (m->nlarge < m->nblocks) {
                const int64_t total = m->nblocks * w;
                const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 16384)));
                if (dtype == KMSD_C128)
                    hipLaunchKernelGGL((block_precond_kernel<double, SMALL_BLOCK>), dim3(g), dim3(256), 0, c->stream,
                                       R, Z, int64_t(w), m->block_ptr, m->factor_off, m->factors, m->nblocks,
                                       static_cast<const PcgFlags *>(nullptr), m->block_rows);
                else
                    hipLaunchKernelGGL((block_precond_kernel<float, SMALL_BLOCK>), dim3(g), dim3(256), 0, c->stream,
                                       R, Z, int64_t(w), m->block_ptr, m->factor_off, m->factors, m->nblocks,
                                       static_cast<const PcgFlags *>(nullptr), m->block_rows);
            }
            large_blocks_apply(m, dtype, R, Z, w, nullptr);
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int kmsd_metric_set_index_blocks(kmsd_metric *m, int64_t nblocks, const int64_t *block_off,
                                            const int64_t *rows, const double *factors) {
    if (!m) return KMSD_EINVAL;
    return guard(m->ctx, [&] {
        const int64_t n = m->n;
        if (nblocks < 0 || (nblocks && (!block_off || !rows || !factors)) || (nblocks == 0 && n > 0))
            throw Error(KMSD_EINVAL, nblocks == 0 ? "incomplete cluster coverage" : "invalid cluster factor");
        if (block_off[0] != 0) throw Error(KMSD_EINVAL, "invalid cluster factor");
        std::vector<char> seen(size_t(n), 0);
        size_t nfac = 0;
        for (int64_t q = 0; q < nblocks; ++q) {
            const int64_t size = block_off[q + 1] - block_off[q];
            if (size < 1) throw Error(KMSD_EINVAL, "invalid cluster factor");
            for (int64_t k = block_off[q]; k < block_off[q + 1]; ++k) {
                const int64_t i = rows[k];
                if (i < 0 || i >= n || seen[size_t(i)]) throw Error(KMSD_EINVAL, "invalid cluster coverage");
                seen[size_t(i)] = 1;
            }
            nfac += size_t(size * size);
        }
        if (std::find(seen.begin(), seen.end(), 0) != seen.end())
            throw Error(KMSD_EINVAL, "incomplete cluster coverage");
        for (size_t e = 0; e < 2 * nfac; ++e)
            if (!std::isfinite(factors[e])) throw Error(KMSD_EINVAL, "invalid cluster factor");
        metric_upload_blocks(m, nblocks, block_off, rows, factors);
        m->identity = false;
    });
}

extern "C" int kmsd_metric_set_ic0(kmsd_metric *m, int64_t nnz, const int64_t *rowptr, const int64_t *colidx,
                                   const double *values, const int64_t *perm, const double *scaling) {
    if (!m) return KMSD_EINVAL;
    kmsd_ctx *c = m->ctx;
    return guard(c, [&] {
        const int64_t n = m->n;
        if (!rowptr || !colidx || !values || !perm || !scaling || nnz < n || rowptr[0] != 0 || rowptr[n] != nnz)
            throw Error(KMSD_EINVAL, "invalid IC dimensions");
        std::vector<char> seen(size_t(n), 0);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t p = perm[i];
            if (p < 0 || p >= n || seen[size_t(p)] || !(scaling[i] > 0) || !std::isfinite(scaling[i]))
                throw Error(KMSD_EINVAL, "invalid IC permutation/scaling");
            seen[size_t(p)] = 1;
            const int64_t b = rowptr[i], e = rowptr[i + 1];
            if (e <= b || e > nnz || colidx[e - 1] != i || !(values[2 * (e - 1)] > 0))
                throw Error(KMSD_EINVAL, "invalid IC factor diagonal");
            for (int64_t k = b; k + 1 < e; ++k)  // the level schedules need a strictly lower pattern
                if (colidx[k] < 0 || colidx[k] >= i) throw Error(KMSD_EINVAL, "invalid IC factor pattern");
        }
        for (int64_t k = 0; k < 2 * nnz; ++k)
            if (!std::isfinite(values[k])) throw Error(KMSD_EINVAL, "invalid IC factor values");
        // L^H by rows: column j of L in decreasing row order.
        const int64_t noff = nnz - n;
        std::vector<int64_t> uptr(size_t(n) + 1, 0);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = rowptr[i]; k + 1 < rowptr[i + 1]; ++k) ++uptr[size_t(colidx[k]) + 1];
        for (int64_t j = 0; j < n; ++j) uptr[size_t(j) + 1] += uptr[size_t(j)];
        std::vector<int64_t> fill(uptr.begin(), uptr.end() - 1);
        std::vector<int32_t> urow(static_cast<size_t>(std::max<int64_t>(noff, 0)));
        std::vector<double> uval(2 * urow.size());
        for (int64_t i = n - 1; i >= 0; --i)
            for (int64_t k = rowptr[i]; k + 1 < rowptr[i + 1]; ++k) {
                const int64_t s = fill[size_t(colidx[k])]++;
                urow[size_t(s)] = int32_t(i);
                uval[2 * size_t(s)] = values[2 * k];
                uval[2 * size_t(s) + 1] = values[2 * k + 1];
            }
        std::vector<int32_t> fwd, bwd;
        std::vector<int64_t> fl, bl;
        levels(n, [&](int64_t i, auto &&dep) {
            for (int64_t k = rowptr[i]; k + 1 < rowptr[i + 1]; ++k) dep(colidx[k]);
        }, fwd, fl);
        // backward: row j waits for the rows i > j of its column; process
        // in reversed numbering so dependencies come first.
        levels(n, [&](int64_t jr, auto &&dep) {
            const int64_t j = n - 1 - jr;
            for (int64_t s = uptr[size_t(j)]; s < uptr[size_t(j) + 1]; ++s) dep(n - 1 - urow[size_t(s)]);
        }, bwd, bl);
        for (int32_t &r : bwd) r = int32_t(n - 1 - r);
        std::vector<int32_t> idx32(colidx, colidx + nnz), perm32(perm, perm + n);
        std::vector<double> scale_orig(scaling, scaling + n);
        auto *ic = new kmsd_metric::IcData();
        try {
            ic->n = n;
            ic->nnz = nnz;
            ic->ptr = upload(c, rowptr, size_t(n) + 1, &ic->allocs);
            ic->idx = upload(c, idx32.data(), idx32.size(), &ic->allocs);
            ic->val = upload(c, values, 2 * size_t(nnz), &ic->allocs);
            ic->uptr = upload(c, uptr.data(), uptr.size(), &ic->allocs);
            ic->urow = upload(c, urow.data(), urow.size(), &ic->allocs);
            ic->uval = upload(c, uval.data(), uval.size(), &ic->allocs);
            ic->perm = upload(c, perm32.data(), perm32.size(), &ic->allocs);
            ic->scale = upload(c, scale_orig.data(), scale_orig.size(), &ic->allocs);
            ic->fwd = upload(c, fwd.data(), fwd.size(), &ic->allocs);
            ic->bwd = upload(c, bwd.data(), bwd.size(), &ic->allocs);
            ic->fwd_lev = std::move(fl);
            ic->bwd_lev = std::move(bl);
            HIPCHECK(hipStreamSynchronize(c->stream));
        } catch (...) {
            for (auto &a : ic->allocs) dfree(c, a.first, a.second);
            delete ic;
            throw;
        }
        HIPCHECK(hipStreamSynchronize(c->stream));
        metric_release_ic(m);
        m->ic = ic;
        m->identity = false;
    });
}

extern "C" int kmsd_metric_set_direct(kmsd_metric *m, int enabled, int steps) {
    if (!m) return KMSD_EINVAL;
    return guard(m->ctx, [&] {
        const bool one_block = m->nblocks == 1 && m->host_block_ptr.size() == 2 && m->host_block_ptr[0] == 0 &&
                               m->host_block_ptr[1] == m->n;
        if (steps < 0 || (enabled && !one_block)) throw Error(KMSD_EINVAL, "invalid direct metric factor");
        m->direct = enabled != 0;
        m->refine_maxiter = steps;
        if (m->direct) m->identity = false;
    });
}

extern "C" int kmsd_metric_refinement_steps(const kmsd_metric *m, int64_t *out) {
    if (!m || !out) return KMSD_EINVAL;
    *out = m->refinement_steps;
    return KMSD_OK;
}

extern "C" int kmsd_ic0_factor(kmsd_ctx *c, int64_t n, int64_t nnz, const int64_t *rowptr, const int64_t *colidx,
                               const double *values, double shift, double *out) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] {
        if (n < 0 || !rowptr || (nnz && (!colidx || !values || !out)) || rowptr[0] != 0 || rowptr[n] != nnz)
            throw Error(KMSD_EINVAL, "invalid IC pattern dimensions");
        if (!std::isfinite(shift) || shift < 0) throw Error(KMSD_EINVAL, "invalid IC shift");
        // The pattern is checked row by row as a sequential factorization
        // would meet it: rows before the first bad row are factored, and a
        // pivot failure among them is reported first.
        int64_t bad = n;
        std::string why;
        for (int64_t i = 0; i < n && bad == n; ++i) {
            const int64_t b = rowptr[i], e = rowptr[i + 1];
            if (e <= b || e > nnz || colidx[e - 1] != i) {
                bad = i;
                why = "IC pattern requires diagonal at row end";
                break;
            }
            for (int64_t k = b; k + 1 < e; ++k)
                if (colidx[k] < 0 || colidx[k] >= i || (k > b && colidx[k - 1] >= colidx[k])) {
                    bad = i;
                    why = "IC pattern must be sorted strict lower triangle";
                    break;
                }
        }
        const int64_t rows_ok = bad;
        const int64_t nz = rows_ok < n ? rowptr[rows_ok] : nnz;
        std::vector<int32_t> order;
        std::vector<int64_t> bounds;
        levels(rows_ok, [&](int64_t i, auto &&dep) {
            for (int64_t k = rowptr[i]; k + 1 < rowptr[i + 1]; ++k) dep(colidx[k]);
        }, order, bounds);
        std::vector<std::pair<void *, size_t>> track;
        unsigned long long fail = ~0ull;
        try {
            std::vector<int32_t> idx32(colidx, colidx + nz);
            int64_t *dptr = upload(c, rowptr, size_t(rows_ok) + 1, &track);
            int32_t *didx = upload(c, idx32.data(), idx32.size(), &track);
            double *da = upload(c, values, 2 * size_t(nz), &track);
            double *df = static_cast<double *>(dalloc(c, 16 * std::max<size_t>(size_t(nz), 1)));
            track.push_back({df, 16 * std::max<size_t>(size_t(nz), 1)});
            int32_t *dord = upload(c, order.data(), order.size(), &track);
            unsigned long long *dfail = upload(c, &fail, 1, &track);
            for (size_t l = 0; l + 1 < bounds.size(); ++l) {
                const int64_t cnt = bounds[l + 1] - bounds[l];
                hipLaunchKernelGGL(ic_factor_kernel, dim3(grid_for(cnt)), dim3(256), 0, c->stream, dord + bounds[l],
                                   cnt, dptr, didx, reinterpret_cast<const cx<double> *>(da),
                                   reinterpret_cast<cx<double> *>(df), shift, dfail);
            }
            HIPCHECK(hipGetLastError());
            copy_d2h(c, &fail, dfail, sizeof(fail), true);
            if (nz) copy_d2h(c, out, df, 16 * size_t(nz), false);
            HIPCHECK(hipStreamSynchronize(c->stream));
        } catch (...) {
            for (auto &a : track) dfree(c, a.first, a.second);
            throw;
        }
        for (auto &a : track) dfree(c, a.first, a.second);
        if (fail != ~0ull)
            throw Error(KMSD_ENUMERIC, "IC nonpositive/nonfinite pivot at row " + std::to_string(fail));
        if (bad < n) throw Error(KMSD_EINVAL, why);
    });
}
