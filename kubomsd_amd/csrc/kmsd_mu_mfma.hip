// MFMA kernels for square double-moment tables (see kmsd_mu_mfma.hpp).
//
// Workgroup tile 128 (j) x 128 (q) of the complex table, four wavefronts in
// a 2 x 2 arrangement, each owning 64 x 64 = 2 x 2 blocks of 32 x 32 MFMA
// outputs for the real and the imaginary part (128 accumulator registers per
// lane).  K (the N orbitals) is walked in slabs of 16 (FP32) or 32 (bf16x3)
// complex entries staged through LDS; the next slab's global loads are issued
// before the current slab's MFMAs (register prefetch).  The bf16x3 products
// take about a fifth of the matrix-core time of the FP32 ones per entry, so
// its slabs are twice as wide: half the barriers per product and twice the
// MFMA work to cover the prefetch latency.  The N range is split across workgroups
// into FP64 partial tables that a second kernel sums in a fixed order.
//
// Operand layout: the sum over K is order-free, so each lane half h takes
// its own complex entries (FP32: entries 4u+2h, 4u+2h+1 of a group of four;
// bf16: entries 8u+4h .. 8u+4h+3 of a group of eight) and reads them as one
// 16-byte LDS access.  With (Lr, Li), (Rr, Ri) the parts of one entry:
//   Re C += Lr Rr + Li Ri        Im C += Lr Ri - Li Rr
// FP32: four v_mfma_f32_32x32x2_f32 per entry pair (products of single
// parts); bf16x3: the interleaved (re, im) pairs are the K slots of
// v_mfma_f32_32x32x16_bf16, the imaginary product uses (Ri, -Rr) pairs.
//
// Workgroups are numbered so that each XCD receives a contiguous range of
// (split, tile) work items: the 64 workgroups an XCD holds at once share one
// N range, so their L and R rows come from that XCD's L2.
#include <algorithm>

#include "kmsd_mu_mfma.hpp"

namespace kmsd {

namespace {

typedef float f16x __attribute__((ext_vector_type(16)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 128;  // workgroup tile (j and q)
constexpr int KT = 16;     // complex entries per FP32 slab
constexpr int KTB = 32;    // complex entries per bf16x3 slab
constexpr int LSF = 36;    // FP32 LDS row stride (floats; 16-byte rows)
constexpr int LSB = 72;    // bf16 LDS row stride (16-bit values; 16-byte rows, 36 words as LSF)

struct Work {
    int x, y;
    int64_t z;
};

__device__ __forceinline__ Work decode(int gx, int gy, int64_t total) {
    const int64_t id = blockIdx.x;
    int64_t w = id;
    if ((total & 7) == 0) w = (id & 7) * (total >> 3) + (id >> 3);
    const int64_t tiles = int64_t(gx) * gy;
    Work r;
    r.z = w / tiles;
    const int64_t t = w - r.z * tiles;
    r.x = int(t % gx);
    r.y = int(t / gx);
    return r;
}

// Slab loads: thread t covers tile row t/2 and complex entries
// [ib + 2Q(t%2), ib + 2Q(t%2) + 2Q) as Q float4 (zeros outside the matrix).
template <int Q>
__device__ __forceinline__ void load_slab(const float *M, int64_t n, int rows, int r0, int64_t ib, int64_t iend,
                                          f4 (&v)[Q]) {
    const int t = threadIdx.x;
    const int r = r0 + (t >> 1);
    const int64_t e0 = ib + 2 * Q * (t & 1);
    const float *p = M + int64_t(r) * 2 * n + 2 * e0;
    if (r < rows && e0 + 2 * Q <= iend && (n & 1) == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = *reinterpret_cast<const f4 *>(p + 4 * q);
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int64_t e = e0 + (4 * q + s) / 2;
                v[q][s] = (r < rows && e < iend) ? p[4 * q + s] : 0.0f;
            }
    }
}

__device__ __forceinline__ void store_partials(double *part, int64_t z, int m, int k, int j0, int q0,
                                               const f16x (&cre)[2][2], const f16x (&cim)[2][2]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int wj = (wave >> 1) * 64, wq = (wave & 1) * 64;
    cx<double> *P = reinterpret_cast<cx<double> *>(part) + z * int64_t(m) * k;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int j = j0 + wj + 32 * a + 8 * (i / 4) + 4 * half + (i % 4);
                const int q = q0 + wq + 32 * b + l32;
                if (j < m && q < k) P[int64_t(j) * k + q] = cx<double>{double(cre[a][b][i]), double(cim[a][b][i])};
            }
}

__global__ __launch_bounds__(256, 2) void mu_f32_kernel(const float *L, const float *R, double *part, int64_t n,
                                                        int m, int k, int gx, int gy, int64_t total,
                                                        int64_t rows_per_split) {
    __shared__ float la[TILE][LSF], ra[TILE][LSF];
    const Work wk = decode(gx, gy, total);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int j0 = wk.x * TILE, q0 = wk.y * TILE;
    const int wj = (wave >> 1) * 64, wq = (wave & 1) * 64;
    const int64_t ib0 = wk.z * rows_per_split;
    const int64_t iend = ib0 + rows_per_split < n ? ib0 + rows_per_split : n;
    f16x cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) cre[a][b] = cim[a][b] = f16x{};
    f4 pl[4], pr[4];
    if (ib0 < iend) {
        load_slab(L, n, m, j0, ib0, iend, pl);
        load_slab(R, n, k, q0, ib0, iend, pr);
    }
    const int sr = t >> 1, sc = 16 * (t & 1);
    for (int64_t ib = ib0; ib < iend; ib += KT) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            *reinterpret_cast<f4 *>(&la[sr][sc + 4 * q]) = pl[q];
            *reinterpret_cast<f4 *>(&ra[sr][sc + 4 * q]) = pr[q];
        }
        __syncthreads();
        if (ib + KT < iend) {
            load_slab(L, n, m, j0, ib + KT, iend, pl);
            load_slab(R, n, k, q0, ib + KT, iend, pr);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int f = 8 * u + 4 * half;
            f4 A[2], B[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) A[a] = *reinterpret_cast<const f4 *>(&la[wj + 32 * a + l32][f]);
#pragma unroll
            for (int b = 0; b < 2; ++b) B[b] = *reinterpret_cast<const f4 *>(&ra[wq + 32 * b + l32][f]);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const float lr = A[a][2 * p], li = A[a][2 * p + 1];
                        const float rr = B[b][2 * p], ri = B[b][2 * p + 1];
                        cre[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(lr, rr, cre[a][b], 0, 0, 0);
                        cre[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(li, ri, cre[a][b], 0, 0, 0);
                        cim[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(lr, ri, cim[a][b], 0, 0, 0);
                        cim[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(li, -rr, cim[a][b], 0, 0, 0);
                    }
        }
    }
    store_partials(part, wk.z, m, k, j0, q0, cre, cim);
}

typedef float f2 __attribute__((ext_vector_type(2)));
typedef __bf16 b2 __attribute__((ext_vector_type(2)));

// Eight floats -> eight bf16 high parts (round to nearest even) and eight
// bf16 residuals x - hi (exact in FP32), packed two per word.  Two-value
// vector conversions map to gfx950's packed v_cvt_pk_bf16_f32.
__device__ __forceinline__ void split8(f4 a, f4 b, u4 &hi, u4 &lo) {
    const f2 x[4] = {{a[0], a[1]}, {a[2], a[3]}, {b[0], b[1]}, {b[2], b[3]}};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const b2 h = __builtin_convertvector(x[e], b2);
        const b2 l = __builtin_convertvector(x[e] - __builtin_convertvector(h, f2), b2);
        hi[e] = __builtin_bit_cast(unsigned, h);
        lo[e] = __builtin_bit_cast(unsigned, l);
    }
}

// (re, im) bf16 pairs -> (im, -re)
__device__ __forceinline__ u4 swap_neg(u4 v) {
    u4 r;
#pragma unroll
    for (int e = 0; e < 4; ++e) r[e] = ((v[e] >> 16) | (v[e] << 16)) ^ 0x80000000u;
    return r;
}

__global__ __launch_bounds__(256, 2) void mu_bf16x3_kernel(const float *L, const float *R, double *part, int64_t n,
                                                           int m, int k, int gx, int gy, int64_t total,
                                                           int64_t rows_per_split) {
    __shared__ unsigned short lh[TILE][LSB], ll[TILE][LSB], rh[TILE][LSB], rl[TILE][LSB];
    const Work wk = decode(gx, gy, total);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int j0 = wk.x * TILE, q0 = wk.y * TILE;
    const int wj = (wave >> 1) * 64, wq = (wave & 1) * 64;
    const int64_t ib0 = wk.z * rows_per_split;
    const int64_t iend = ib0 + rows_per_split < n ? ib0 + rows_per_split : n;
    f16x cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) cre[a][b] = cim[a][b] = f16x{};
    f4 pl[8], pr[8];
    if (ib0 < iend) {
        load_slab(L, n, m, j0, ib0, iend, pl);
        load_slab(R, n, k, q0, ib0, iend, pr);
    }
    const int sr = t >> 1, sc = 32 * (t & 1);
    for (int64_t ib = ib0; ib < iend; ib += KTB) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            u4 h, l;
            split8(pl[2 * q], pl[2 * q + 1], h, l);
            *reinterpret_cast<u4 *>(&lh[sr][sc + 8 * q]) = h;
            *reinterpret_cast<u4 *>(&ll[sr][sc + 8 * q]) = l;
            split8(pr[2 * q], pr[2 * q + 1], h, l);
            *reinterpret_cast<u4 *>(&rh[sr][sc + 8 * q]) = h;
            *reinterpret_cast<u4 *>(&rl[sr][sc + 8 * q]) = l;
        }
        __syncthreads();
        if (ib + KTB < iend) {
            load_slab(L, n, m, j0, ib + KTB, iend, pl);
            load_slab(R, n, k, q0, ib + KTB, iend, pr);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int f = 16 * u + 8 * half;
            u4 Ah[2], Al[2], Bh[2], Bl[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                Ah[a] = *reinterpret_cast<const u4 *>(&lh[wj + 32 * a + l32][f]);
                Al[a] = *reinterpret_cast<const u4 *>(&ll[wj + 32 * a + l32][f]);
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                Bh[b] = *reinterpret_cast<const u4 *>(&rh[wq + 32 * b + l32][f]);
                Bl[b] = *reinterpret_cast<const u4 *>(&rl[wq + 32 * b + l32][f]);
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const bf8 bh = __builtin_bit_cast(bf8, Bh[b]), bl = __builtin_bit_cast(bf8, Bl[b]);
                const bf8 bhi = __builtin_bit_cast(bf8, swap_neg(Bh[b]));
                const bf8 bli = __builtin_bit_cast(bf8, swap_neg(Bl[b]));
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const bf8 ah = __builtin_bit_cast(bf8, Ah[a]), al = __builtin_bit_cast(bf8, Al[a]);
                    cre[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, cre[a][b], 0, 0, 0);
                    cre[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, cre[a][b], 0, 0, 0);
                    cre[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, cre[a][b], 0, 0, 0);
                    cim[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bhi, cim[a][b], 0, 0, 0);
                    cim[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bli, cim[a][b], 0, 0, 0);
                    cim[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bhi, cim[a][b], 0, 0, 0);
                }
            }
        }
    }
    store_partials(part, wk.z, m, k, j0, q0, cre, cim);
}

__global__ __launch_bounds__(256) void mu_sum_kernel(const double *part, double *table, int64_t count, int32_t splits,
                                                     int32_t accumulate) {
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

void mu_mfma(kmsd_ctx *c, const void *left, const void *right, double *table, int64_t n, int64_t m, int64_t k,
             int mode, bool accumulate) {
    if (m >= (int64_t(1) << 30) || k >= (int64_t(1) << 30)) throw Error(KMSD_EINVAL, "mu table too large");
    const int64_t gx = (m + TILE - 1) / TILE, gy = (k + TILE - 1) / TILE;
    const int64_t count = m * k;
    // Splits of the N range: enough workgroups to fill the chip twice over,
    // at most 2^16 entries per FP32 accumulation run (the FP32 rounding of the
    // accumulators grows with the run length; partial sums are FP64),
    // partials <= 1 GiB.
    const int64_t kt = mode == KMSD_MU_BF16X3 ? KTB : KT;
    const int64_t slabs = (n + kt - 1) / kt;
    int64_t splits = std::max<int64_t>((2048 + gx * gy - 1) / (gx * gy), (n + (int64_t(1) << 16) - 1) >> 16);
    splits = std::min<int64_t>(splits, slabs);
    splits = std::max<int64_t>(1, std::min<int64_t>(splits, (int64_t(1) << 30) / (16 * count)));
    int64_t rows = (n + splits - 1) / splits;
    rows = ((rows + kt - 1) / kt) * kt;
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
    const int64_t total = gx * gy * splits;
    if (total >= (int64_t(1) << 31)) throw Error(KMSD_EINVAL, "mu table too large");
    const float *L = static_cast<const float *>(left);
    const float *R = static_cast<const float *>(right);
    if (mode == KMSD_MU_BF16X3)
        hipLaunchKernelGGL(mu_bf16x3_kernel, dim3(unsigned(total)), dim3(256), 0, c->stream, L, R, part, n, int(m),
                           int(k), int(gx), int(gy), total, rows);
    else
        hipLaunchKernelGGL(mu_f32_kernel, dim3(unsigned(total)), dim3(256), 0, c->stream, L, R, part, n, int(m),
                           int(k), int(gx), int(gy), total, rows);
    HIPCHECK(hipGetLastError());
    const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)));
    hipLaunchKernelGGL(mu_sum_kernel, dim3(g), dim3(256), 0, c->stream, part, table, count, int32_t(splits),
                       int32_t(accumulate));
    HIPCHECK(hipGetLastError());
}

}  // namespace kmsd
