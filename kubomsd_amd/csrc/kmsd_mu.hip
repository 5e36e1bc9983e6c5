// Spectral double-moment tables: table[b][j][q] = sum_i conj(L[j][i][b]) R[q][i][b]
// (reference: einsum('jnb,knb->bjk', left.conj(), cov), src/kubo_msd/solvers/
// spectral.py:37).  A batched tall-skinny complex GEMM reducing over the N
// orbitals.  KMSD_MU_FP64: vector FMA with FP64 accumulation for either
// input precision (complex64 products are exact in FP64); the N reduction is
// split over workgroups into partial tables summed in a fixed order.
#include <algorithm>
#include <cmath>

#include "kmsd_internal.hpp"
#include "kmsd_kernels.hpp"
#include "kmsd_mu_mfma.hpp"

using namespace kmsd;

namespace {

constexpr int TJ = 16;  // j rows per thread

// grid (ceil(m/TJ), pairs/256, splits); thread <-> (b, q) pair, TJ j's in registers.
template <class VT>
__global__ __launch_bounds__(256) void mu_vector_kernel(const void *Lv, const void *Rv, double *part, int64_t n,
                                                        int32_t m, int32_t k, int32_t nb, int64_t rows_per_split) {
    const cx<VT> *L = static_cast<const cx<VT> *>(Lv);
    const cx<VT> *R = static_cast<const cx<VT> *>(Rv);
    const int j0 = blockIdx.x * TJ;
    const int pair = blockIdx.y * 256 + threadIdx.x;
    const int b = pair % nb, q = pair / nb;
    const bool ok = q < k;
    const int64_t i0 = int64_t(blockIdx.z) * rows_per_split;
    const int64_t i1 = i0 + rows_per_split < n ? i0 + rows_per_split : n;
    __shared__ cx<double> lt[TJ][64];  // left tile for up to 64 batch columns
    double ar[TJ], ai[TJ];
#pragma unroll
    for (int t = 0; t < TJ; ++t) ar[t] = ai[t] = 0.0;
    const int nbc = nb < 64 ? nb : 64;  // nb <= 64 enforced by the host
    for (int64_t i = i0; i < i1; ++i) {
        __syncthreads();
        for (int e = threadIdx.x; e < TJ * nbc; e += 256) {
            const int t = e / nbc, bb = e - t * nbc;
            const int j = j0 + t;
            lt[t][bb] = j < m ? to<double>(L[(int64_t(j) * n + i) * nb + bb]) : cx<double>{0.0, 0.0};
        }
        __syncthreads();
        if (!ok) continue;
        const cx<double> r = to<double>(R[(int64_t(q) * n + i) * nb + b]);
#pragma unroll
        for (int t = 0; t < TJ; ++t) {
            const cx<double> l = lt[t][b];
            // conj(l) * r
            ar[t] = __builtin_fma(l.re, r.re, ar[t]);
            ar[t] = __builtin_fma(l.im, r.im, ar[t]);
            ai[t] = __builtin_fma(l.re, r.im, ai[t]);
            ai[t] = __builtin_fma(-l.im, r.re, ai[t]);
        }
    }
    if (!ok) return;
    cx<double> *P = reinterpret_cast<cx<double> *>(part) + int64_t(blockIdx.z) * nb * m * k;
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
        const int j = j0 + t;
        if (j < m) P[(int64_t(b) * m + j) * k + q] = cx<double>{ar[t], ai[t]};
    }
}

__global__ __launch_bounds__(256) void mu_reduce_kernel(const double *part, double *table, int64_t count,
                                                        int32_t splits, int32_t accumulate) {
    const cx<double> *P = reinterpret_cast<const cx<double> *>(part);
    cx<double> *T = reinterpret_cast<cx<double> *>(table);
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < count; e += int64_t(gridDim.x) * 256) {
        cx<double> s = accumulate ? T[e] : cx<double>{0.0, 0.0};
        for (int32_t z = 0; z < splits; ++z) {
            const cx<double> v = P[int64_t(z) * count + e];
            s.re += v.re;
            s.im += v.im;
        }
        T[e] = s;
    }
}

}  // namespace

extern "C" int kmsd_mu_table(kmsd_ctx *c, const void *left, const void *right, void *table, int64_t n, int64_t m,
                             int64_t k, int64_t nb, int dtype, int mode, int accumulate) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] {
        if (!left || !right || !table || n < 1 || m < 1 || k < 1 || nb < 1)
            throw Error(KMSD_EINVAL, "invalid mu table dimensions");
        if (dtype != KMSD_C64 && dtype != KMSD_C128) throw Error(KMSD_EINVAL, "dtype must be KMSD_C64 or KMSD_C128");
        if (mode == KMSD_MU_FP32 || mode == KMSD_MU_BF16X3) {
            if (dtype != KMSD_C64 || nb != 1)
                throw Error(KMSD_EINVAL, "MFMA mu tables take complex64 blocks with one batch column");
            mu_mfma(c, left, right, static_cast<double *>(table), n, m, k, mode, accumulate != 0);
            HIPCHECK(hipStreamSynchronize(c->stream));
            return;
        }
        if (mode != KMSD_MU_FP64) throw Error(KMSD_EINVAL, "unknown mu table mode");
        if (nb > 64) throw Error(KMSD_EINVAL, "mu tables take at most 64 batch columns per call");
        const int64_t pairs = k * nb;
        const int64_t gx = (m + TJ - 1) / TJ, gy = (pairs + 255) / 256;
        int64_t splits = std::max<int64_t>(1, std::min<int64_t>(n, (2048 + gx * gy - 1) / (gx * gy)));
        const int64_t count = nb * m * k;
        // bound the partial buffer to ~1 GiB
        splits = std::max<int64_t>(1, std::min<int64_t>(splits, (int64_t(1) << 30) / (16 * count)));
        const int64_t rows = (n + splits - 1) / splits;
        splits = (n + rows - 1) / rows;
        const size_t bytes = size_t(16) * count * splits;
        if (!c->mu_part || c->mu_part_bytes < bytes) {
            if (c->mu_part) {
                HIPCHECK(hipStreamSynchronize(c->stream));
                dfree(c, c->mu_part, c->mu_part_bytes);
                c->mu_part = nullptr;
                c->mu_part_bytes = 0;
            }
            c->mu_part = dalloc(c, bytes);
            c->mu_part_bytes = bytes;
        }
        double *part = static_cast<double *>(c->mu_part);
        const dim3 grid{unsigned(gx), unsigned(gy), unsigned(splits)};
        if (dtype == KMSD_C128)
            hipLaunchKernelGGL(mu_vector_kernel<double>, grid, dim3(256), 0, c->stream, left, right, part, n,
                               int32_t(m), int32_t(k), int32_t(nb), rows);
        else
            hipLaunchKernelGGL(mu_vector_kernel<float>, grid, dim3(256), 0, c->stream, left, right, part, n,
                               int32_t(m), int32_t(k), int32_t(nb), rows);
        HIPCHECK(hipGetLastError());
        const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)));
        hipLaunchKernelGGL(mu_reduce_kernel, dim3(g), dim3(256), 0, c->stream, part, static_cast<double *>(table),
                           count, int32_t(splits), accumulate);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}
