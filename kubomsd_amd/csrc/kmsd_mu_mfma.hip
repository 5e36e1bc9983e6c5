// MFMA kernels for square double-moment tables (see kmsd_mu_mfma.hpp).
#include <algorithm>

#include "kmsd_mu_mfma.hpp"

namespace kmsd {

namespace {

typedef float f16x __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int TILE = 64;         // output tile (j and q)
constexpr int KT = 32;           // complex entries per K slab
constexpr int KF = 2 * KT;       // floats per row of a slab
constexpr int FSTRIDE = KF + 1;  // padded LDS row stride (floats)
constexpr int BSTRIDE = KF + 8;  // padded LDS row stride (bf16 elements)

__device__ __forceinline__ unsigned short bf16_bits_rne(float x) {
    const unsigned u = __builtin_bit_cast(unsigned, x);
    return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_float(unsigned short b) {
    return __builtin_bit_cast(float, static_cast<unsigned>(b) << 16);
}

// Loads the slab rows [r0, r0+64) x complex [ib, ib+KT) of a (rows x n) complex64
// matrix (as floats) into per-thread registers: thread t covers row t/4 and
// floats [16*(t%4), 16*(t%4)+16).
__device__ __forceinline__ void load_slab(const float *M, int64_t n, int rows, int r0, int64_t ib, int64_t iend,
                                          float (&v)[16]) {
    const int t = threadIdx.x;
    const int r = r0 + (t >> 2);
    const int f0 = (t & 3) * 16;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int64_t ci = ib + (f0 + e) / 2;
        v[e] = (r < rows && ci < iend) ? M[int64_t(r) * 2 * n + 2 * ci + (e & 1)] : 0.0f;
    }
}

__global__ __launch_bounds__(256) void mu_mfma_f32_kernel(const float *L, const float *R, double *part, int64_t n,
                                                          int m, int k, int64_t rows_per_split) {
    __shared__ float la[TILE][FSTRIDE], ra[TILE][FSTRIDE];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int j0 = blockIdx.x * TILE, q0 = blockIdx.y * TILE;
    const int wj = (wave >> 1) * 32, wq = (wave & 1) * 32;
    const int64_t ib0 = int64_t(blockIdx.z) * rows_per_split;
    const int64_t iend = ib0 + rows_per_split < n ? ib0 + rows_per_split : n;
    f16x cre = {}, cim = {};
    const int half = lane >> 5, l32 = lane & 31;
    for (int64_t ib = ib0; ib < iend; ib += KT) {
        float vl[16], vr[16];
        load_slab(L, n, m, j0, ib, iend, vl);
        load_slab(R, n, k, q0, ib, iend, vr);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            la[t >> 2][(t & 3) * 16 + e] = vl[e];
            ra[t >> 2][(t & 3) * 16 + e] = vr[e];
        }
        __syncthreads();
#pragma unroll 4
        for (int s = 0; s < KT; ++s) {
            const int f = 2 * s + half;
            const float a = la[wj + l32][f];
            const float bre = ra[wq + l32][f];
            const float bim = half ? -ra[wq + l32][f - 1] : ra[wq + l32][f + 1];
            cre = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bre, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bim, cim, 0, 0, 0);
        }
    }
    cx<double> *P = reinterpret_cast<cx<double> *>(part) + int64_t(blockIdx.z) * m * k;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int j = j0 + wj + 8 * (i / 4) + 4 * half + (i % 4);
        const int q = q0 + wq + l32;
        if (j < m && q < k) P[int64_t(j) * k + q] = cx<double>{double(cre[i]), double(cim[i])};
    }
}

__global__ __launch_bounds__(256) void mu_mfma_bf16x3_kernel(const float *L, const float *R, double *part, int64_t n,
                                                             int m, int k, int64_t rows_per_split) {
    __shared__ unsigned short lh[TILE][BSTRIDE], ll[TILE][BSTRIDE], rh[TILE][BSTRIDE], rl[TILE][BSTRIDE];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int j0 = blockIdx.x * TILE, q0 = blockIdx.y * TILE;
    const int wj = (wave >> 1) * 32, wq = (wave & 1) * 32;
    const int64_t ib0 = int64_t(blockIdx.z) * rows_per_split;
    const int64_t iend = ib0 + rows_per_split < n ? ib0 + rows_per_split : n;
    f16x cre = {}, cim = {};
    const int half = lane >> 5, l32 = lane & 31;
    for (int64_t ib = ib0; ib < iend; ib += KT) {
        float vl[16], vr[16];
        load_slab(L, n, m, j0, ib, iend, vl);
        load_slab(R, n, k, q0, ib, iend, vr);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int col = (t & 3) * 16 + e;
            const unsigned short h1 = bf16_bits_rne(vl[e]);
            lh[t >> 2][col] = h1;
            ll[t >> 2][col] = bf16_bits_rne(vl[e] - bf16_to_float(h1));
            const unsigned short h2 = bf16_bits_rne(vr[e]);
            rh[t >> 2][col] = h2;
            rl[t >> 2][col] = bf16_bits_rne(vr[e] - bf16_to_float(h2));
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KF / 16; ++kk) {
            const int f0 = 16 * kk + 8 * half;
            bf8 ah, al, bhr, blr, bhi, bli;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                ah[e] = __builtin_bit_cast(__bf16, lh[wj + l32][f0 + e]);
                al[e] = __builtin_bit_cast(__bf16, ll[wj + l32][f0 + e]);
                bhr[e] = __builtin_bit_cast(__bf16, rh[wq + l32][f0 + e]);
                blr[e] = __builtin_bit_cast(__bf16, rl[wq + l32][f0 + e]);
                // imaginary operand: pairs (re, im) -> (im, -re)
                const int src = (e & 1) ? e - 1 : e + 1;
                const unsigned short sgn = (e & 1) ? 0x8000u : 0u;
                bhi[e] = __builtin_bit_cast(__bf16, static_cast<unsigned short>(rh[wq + l32][f0 + src] ^ sgn));
                bli[e] = __builtin_bit_cast(__bf16, static_cast<unsigned short>(rl[wq + l32][f0 + src] ^ sgn));
            }
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bhr, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, blr, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bhr, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bhi, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bli, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bhi, cim, 0, 0, 0);
        }
    }
    cx<double> *P = reinterpret_cast<cx<double> *>(part) + int64_t(blockIdx.z) * m * k;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int j = j0 + wj + 8 * (i / 4) + 4 * half + (i % 4);
        const int q = q0 + wq + l32;
        if (j < m && q < k) P[int64_t(j) * k + q] = cx<double>{double(cre[i]), double(cim[i])};
    }
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
    int64_t splits = std::max<int64_t>(1, (2048 + gx * gy - 1) / (gx * gy));
    splits = std::min<int64_t>(splits, std::max<int64_t>(1, (n + KT - 1) / KT));
    splits = std::max<int64_t>(1, std::min<int64_t>(splits, (int64_t(1) << 31) / (16 * count)));
    int64_t rows = (n + splits - 1) / splits;
    rows = ((rows + KT - 1) / KT) * KT;
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
    const float *L = static_cast<const float *>(left);
    const float *R = static_cast<const float *>(right);
    if (mode == KMSD_MU_BF16X3)
        hipLaunchKernelGGL(mu_mfma_bf16x3_kernel, grid, dim3(256), 0, c->stream, L, R, part, n, int(m), int(k), rows);
    else
        hipLaunchKernelGGL(mu_mfma_f32_kernel, grid, dim3(256), 0, c->stream, L, R, part, n, int(m), int(k), rows);
    HIPCHECK(hipGetLastError());
    const unsigned g = unsigned(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 8192)));
    hipLaunchKernelGGL(mu_sum_kernel, dim3(g), dim3(256), 0, c->stream, part, table, count, int32_t(splits),
                       int32_t(accumulate));
    HIPCHECK(hipGetLastError());
}

}  // namespace kmsd
