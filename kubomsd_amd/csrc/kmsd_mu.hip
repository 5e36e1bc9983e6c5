// Spectral double-moment tables: table[b][j][q] = sum_i conj(L[j][i][b]) R[q][i][b]
// (reference: einsum('jnb,knb->bjk', left.conj(), cov), src/kubo_msd/solvers/
// spectral.py:37).  A batched tall-skinny complex GEMM reducing over the N
// orbitals, one GEMM per batch column b.
//
// KMSD_MU_FP64 runs on the FP64 matrix cores (v_mfma_f64_16x16x4f64) for
// either input precision: complex64 entries are widened exactly, products
// and sums are FP64.  A workgroup covers up to 64 columns q of the tables of
// G = 4, 2 or 1 batch columns (4/G wavefronts per column, 32 rows j each),
// whose entries sit next to each other in the (rows, N, nb) blocks (32
// contiguous bytes per (j, i) for four complex64 columns); narrow batches
// take G = 1 or 2 so that no wavefront idles.  i is walked in slabs staged
// through LDS as real and
// imaginary planes, the next slab loaded into registers while the current
// one feeds the matrix cores.  Per wavefront 2 x 4 blocks of 16 x 16 outputs
// for the real and the imaginary part (128 accumulator registers); column
// blocks beyond k are skipped.  Workgroups that share rows and an N range but
// hold different batch columns are numbered consecutively per XCD, so the
// 32-byte pieces of one row's cache lines are used from that XCD's L2.  The N
// reduction is split across workgroups into FP64 partial tables summed in a
// fixed order.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "kmsd_internal.hpp"
#include "kmsd_kernels.hpp"
#include "kmsd_mu_mfma.hpp"

using namespace kmsd;

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int WR = 32;  // rows j per wavefront (two 16-row MFMA blocks)
constexpr int BQ = 64;  // columns q per workgroup (up to four 16-column blocks)

// G consecutive batch columns b0 .. b0+G-1 of entry (row, i): one contiguous
// run when vec (nb % G == 0 and aligned rows), else per column.
template <class VT, int G>
__device__ __forceinline__ void loadg(const cx<VT> *M, int64_t n, int nb, int row, int64_t i, int b0, bool ok,
                                      bool vec, cx<VT> (&v)[G]) {
    if (!ok) {
#pragma unroll
        for (int e = 0; e < G; ++e) v[e] = cx<VT>{VT(0), VT(0)};
        return;
    }
    const cx<VT> *p = M + (int64_t(row) * n + i) * nb + b0;
    if (vec) {
#pragma unroll
        for (int e = 0; e < G; ++e) v[e] = p[e];
    } else {
#pragma unroll
        for (int e = 0; e < G; ++e) v[e] = b0 + e < nb ? p[e] : cx<VT>{VT(0), VT(0)};
    }
}

// G batch columns per workgroup, 4/G wavefronts per batch column, each
// wavefront 32 rows j x up to 64 columns q of its column's table.  grid: one
// workgroup per (batch group, row tile, column tile, N split), the batch
// group fastest within each XCD's share.  IS: i entries per slab.  Slab
// loads walk i fastest (consecutive lanes read consecutive entries of a row).
template <class VT, int IS, int G>
__global__ __launch_bounds__(256) void mu_batched_kernel(const void *Lv, const void *Rv, double *part, int64_t n,
                                                         int m, int k, int nb, int gj, int gq, int gb, int64_t total,
                                                         int64_t rows_per_split, int vec) {
    static_assert(IS == 8 || IS == 16, "slab of 8 or 16 entries");
    static_assert(G == 1 || G == 2 || G == 4, "1, 2 or 4 batch columns per workgroup");
    constexpr int WJ = 4 / G;        // wavefronts per batch column
    constexpr int BJ = WR * WJ;      // rows j per workgroup
    constexpr int LPJ = BJ + 2, LPQ = BQ + 2;  // padded LDS rows (conflict-free 16-lane reads)
    constexpr int RP = 256 / IS;     // rows per load pass (lanes walk i fastest)
    constexpr int NL = BJ / RP, NR = BQ / RP;
    __shared__ VT lre[G][IS][LPJ], lim[G][IS][LPJ], rre[G][IS][LPQ], rim[G][IS][LPQ];
    const cx<VT> *L = static_cast<const cx<VT> *>(Lv);
    const cx<VT> *R = static_cast<const cx<VT> *>(Rv);
    int64_t w = blockIdx.x;
    if ((total & 7) == 0) w = (w & 7) * (total >> 3) + (w >> 3);
    const int bg = int(w % gb);
    int64_t rest = w / gb;
    const int jt = int(rest % gj);
    rest /= gj;
    const int qt = int(rest % gq);
    const int64_t z = rest / gq;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int bcol = wave / WJ, jw = wave % WJ;
    const int j0 = jt * BJ, q0 = qt * BQ, b0 = bg * G;
    const int64_t i0 = z * rows_per_split;
    const int64_t i1 = i0 + rows_per_split < n ? i0 + rows_per_split : n;
    const int nqb = min(4, (k - q0 + 15) / 16);  // column blocks inside the table (uniform)
    const int li = t % IS, lr = t / IS;          // load role: entry li of rows lr + RP s
    cx<VT> pl[NL][G], pr[NR][G];
    auto fetch = [&](int64_t ib) {
        const int64_t i = ib + li;
#pragma unroll
        for (int s = 0; s < NL; ++s) {
            const int j = j0 + lr + RP * s;
            loadg<VT, G>(L, n, nb, j, i, b0, j < m && i < i1, vec != 0, pl[s]);
        }
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            const int q = q0 + lr + RP * s;
            loadg<VT, G>(R, n, nb, q, i, b0, q < k && i < i1, vec != 0, pr[s]);
        }
    };
    d4 cre[2][4], cim[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) cre[a][c] = cim[a][c] = d4{0.0, 0.0, 0.0, 0.0};
    if (i0 < i1) fetch(i0);
    const int c16 = lane & 15, r4 = lane >> 4;
    for (int64_t ib = i0; ib < i1; ib += IS) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NL; ++s)
#pragma unroll
            for (int e = 0; e < G; ++e) {
                lre[e][li][lr + RP * s] = pl[s][e].re;
                lim[e][li][lr + RP * s] = pl[s][e].im;
            }
#pragma unroll
        for (int s = 0; s < NR; ++s)
#pragma unroll
            for (int e = 0; e < G; ++e) {
                rre[e][li][lr + RP * s] = pr[s][e].re;
                rim[e][li][lr + RP * s] = pr[s][e].im;
            }
        __syncthreads();
        if (ib + IS < i1) fetch(ib + IS);
#pragma unroll
        for (int kk = 0; kk < IS / 4; ++kk) {
            const int il = 4 * kk + r4;
            double ar[2], ai[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                ar[a] = double(lre[bcol][il][WR * jw + 16 * a + c16]);
                ai[a] = double(lim[bcol][il][WR * jw + 16 * a + c16]);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (c >= nqb) break;
                const double br = double(rre[bcol][il][16 * c + c16]);
                const double bi = double(rim[bcol][il][16 * c + c16]);
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    // conj(l) r: Re = lr rr + li ri, Im = lr ri - li rr
                    cre[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br, cre[a][c], 0, 0, 0);
                    cre[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], bi, cre[a][c], 0, 0, 0);
                    cim[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi, cim[a][c], 0, 0, 0);
                    cim[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], br, cim[a][c], 0, 0, 0);
                }
            }
        }
    }
    const int b = b0 + bcol;
    if (b >= nb) return;
    // output layout of v_mfma_f64_16x16x4f64: lane l, register r holds row l/16 + 4r, column l%16
    cx<double> *P = reinterpret_cast<cx<double> *>(part) + z * int64_t(nb) * m * k;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + WR * jw + 16 * a + r4 + 4 * r, q = q0 + 16 * c + c16;
                if (j < m && q < k) P[(int64_t(b) * m + j) * k + q] = cx<double>{cre[a][c][r], cim[a][c][r]};
            }
}

template <class VT, int IS, int G>
void launch_batched(kmsd_ctx *c, const void *left, const void *right, double *part, int64_t n, int64_t m,
                    int64_t k, int64_t nb, int64_t gj, int64_t gq, int64_t gb, int64_t total, int64_t rows,
                    bool vec) {
    hipLaunchKernelGGL((mu_batched_kernel<VT, IS, G>), dim3(unsigned(total)), dim3(256), 0, c->stream, left, right,
                       part, n, int(m), int(k), int(nb), int(gj), int(gq), int(gb), total, rows, int(vec));
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
        if (m >= (int64_t(1) << 30) || k >= (int64_t(1) << 30) || nb >= (int64_t(1) << 30))
            throw Error(KMSD_EINVAL, "mu table too large");
        // batch columns per workgroup: four when they divide the width (or
        // the width is large), else two, else one (no idle wavefronts)
        const int G = (nb % 4 == 0 || nb >= 12) ? 4 : (nb % 2 == 0 ? 2 : 1);
        const int64_t gj = (m + 128 / G - 1) / (128 / G), gq = (k + BQ - 1) / BQ, gb = (nb + G - 1) / G;
        const int64_t tiles = gj * gq * gb;
        const int64_t count = nb * m * k;
        const int is = dtype == KMSD_C128 ? 8 : 16;
        // N splits: about 2048 workgroups (several per CU), partials <= 1 GiB
        const int64_t slabs = (n + is - 1) / is;
        int64_t splits = std::max<int64_t>(1, std::min<int64_t>(slabs, (2048 + tiles - 1) / tiles));
        splits = std::max<int64_t>(1, std::min<int64_t>(splits, (int64_t(1) << 30) / (16 * count)));
        int64_t rows = (n + splits - 1) / splits;
        rows = (rows + is - 1) / is * is;
        splits = (n + rows - 1) / rows;
        const int64_t total = tiles * splits;
        if (total >= (int64_t(1) << 31)) throw Error(KMSD_EINVAL, "mu table too large");
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
        const size_t eb = dtype == KMSD_C128 ? 16 : 8;
        const bool vec = nb % G == 0 && reinterpret_cast<uintptr_t>(left) % (G * eb) == 0 &&
                         reinterpret_cast<uintptr_t>(right) % (G * eb) == 0;
        const auto run = [&](auto vt, auto is_c, auto g_c) {
            using VT = decltype(vt);
            launch_batched<VT, decltype(is_c)::value, decltype(g_c)::value>(c, left, right, part, n, m, k, nb, gj,
                                                                            gq, gb, total, rows, vec);
        };
        using I8 = std::integral_constant<int, 8>;
        using I16 = std::integral_constant<int, 16>;
        using G1 = std::integral_constant<int, 1>;
        using G2 = std::integral_constant<int, 2>;
        using G4 = std::integral_constant<int, 4>;
        if (dtype == KMSD_C128) {
            if (G == 4) run(double{}, I8{}, G4{});
            else if (G == 2) run(double{}, I8{}, G2{});
            else run(double{}, I8{}, G1{});
        } else {
            if (G == 4) run(float{}, I16{}, G4{});
            else if (G == 2) run(float{}, I16{}, G2{});
            else run(float{}, I16{}, G1{});
        }
        HIPCHECK(hipGetLastError());
        const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)));
        hipLaunchKernelGGL(mu_reduce_kernel, dim3(g), dim3(256), 0, c->stream, part, static_cast<double *>(table),
                           count, int32_t(splits), accumulate);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}
