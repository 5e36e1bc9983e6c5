// gfx950 kernels for the stochastic KPM path: block-vector CSR products fused
// with the Chebyshev three-term update, the moment dot and the series
// accumulation; deterministic two-stage column reductions.
//
// Vector blocks are (N, ld) row-major, complex interleaved, the layout the
// reference's Dense blocks use (cpp/native.cpp:26-41), so a row's columns are
// contiguous and one row of a block is one coalesced segment.
//
// Lane mapping (256-thread workgroups = 4 wavefronts of 64):
//   lanes per row LPR = 2^lpr_log2 (<= 64), each lane owns CPL consecutive
//   complex columns; a wavefront covers 64/LPR rows; grid.y walks column
//   chunks of width LPR*CPL.  Rows are walked in blocks of 4*64/LPR rows with
//   an XCD-aware grid stride: workgroup g runs on XCD g%8 (round-robin
//   dispatch) and walks a contiguous eighth of the row blocks, so the rows a
//   CU gathers from stay in its XCD's L2.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace kmsd {

template <class R>
struct alignas(2 * sizeof(R)) cx {
    R re, im;
};

template <class A, class B>
__device__ __forceinline__ cx<A> to(cx<B> v) {
    return {A(v.re), A(v.im)};
}

// Matrix value types: real scalar or complex pair.
template <class R, bool CPLX>
struct MatVal;
template <class R>
struct MatVal<R, false> {
    using type = R;
    template <class A>
    __device__ static __forceinline__ void fma(cx<A> &acc, R v, cx<A> x) {
        acc.re = __builtin_fma(A(v), x.re, acc.re);
        acc.im = __builtin_fma(A(v), x.im, acc.im);
    }
};
template <class R>
struct MatVal<R, true> {
    using type = cx<R>;
    template <class A>
    __device__ static __forceinline__ void fma(cx<A> &acc, cx<R> v, cx<A> x) {
        acc.re = __builtin_fma(A(v.re), x.re, acc.re);
        acc.re = __builtin_fma(-A(v.im), x.im, acc.re);
        acc.im = __builtin_fma(A(v.re), x.im, acc.im);
        acc.im = __builtin_fma(A(v.im), x.re, acc.im);
    }
};

// Epilogue / operand flags (wave-uniform runtime branches).
enum : uint32_t {
    F_SPMM = 1u << 0,   // y = sum_k A_ik X_k ; otherwise y = Z_i
    F_BETA = 1u << 1,   // t += beta * X_i
    F_P = 1u << 2,      // t += gamma * P_i
    F_STORE = 1u << 3,  // OUT_i = t
    F_MOM = 1u << 4,    // partial += conj(W_i) * t
    F_ACC = 1u << 5,    // ACC_i += coef * t
};

struct StepArgs {
    const int32_t *ptr;   // CSR row pointers (n+1)
    const int32_t *idx;   // CSR column indices
    const void *val;      // CSR values (R or cx<R>)
    const void *X;        // gathered input block
    const void *Z;        // row-local input when no SpMM
    const void *P;        // previous iterate
    void *OUT;            // output block (may alias P)
    const void *W;        // moment projection block
    void *ACC;            // series accumulator block
    double *part;         // [gridDim.x][w] complex partial moment sums
    int64_t ldx, ldz, ldp, ldo, ldw, lda;
    int64_t n;            // rows
    double alpha_re, alpha_im, beta_re, beta_im, gamma;
    double coef_re, coef_im;
    int32_t w;            // columns processed (multiple of CPL)
    int32_t lpr_log2;
    uint32_t flags;
};

template <class A>
__device__ __forceinline__ cx<A> cmul(double ar, double ai, cx<A> x) {
    return {A(ar) * x.re - A(ai) * x.im, A(ar) * x.im + A(ai) * x.re};
}

// VT: vector element real type (float -> complex64, double -> complex128)
// MR: matrix value real type; MC: complex matrix values
// AT: arithmetic type of the row sums (double for FP64 vectors; float or
//     double for FP32 vectors)
// CPL: complex columns per lane (1 or 2)
template <class VT, class MR, bool MC, class AT, int CPL>
__global__ __launch_bounds__(256) void step_kernel(const StepArgs a) {
    using V = cx<VT>;
    using MV = MatVal<MR, MC>;
    using MT = typename MV::type;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lpr = 1 << a.lpr_log2;
    const int rpw = 64 >> a.lpr_log2;          // rows per wavefront
    const int rpb = 4 * rpw;                    // rows per workgroup pass
    const int rloc = wave * rpw + (lane >> a.lpr_log2);
    const int cgrp = lane & (lpr - 1);
    const int col = (blockIdx.y * lpr + cgrp) * CPL;
    const bool colok = col < a.w;

    const V *X = static_cast<const V *>(a.X);
    const V *Z = static_cast<const V *>(a.Z);
    const V *P = static_cast<const V *>(a.P);
    V *OUT = static_cast<V *>(a.OUT);
    const V *W = static_cast<const V *>(a.W);
    V *ACC = static_cast<V *>(a.ACC);
    const MT *val = static_cast<const MT *>(a.val);
    const uint32_t fl = a.flags;

    cx<double> mom[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) mom[c] = {0.0, 0.0};

    // XCD-aware row-block walk.
    const int64_t nblk = (a.n + rpb - 1) / rpb;
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;

    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t row = blk * rpb + rloc;
        if (row >= a.n || !colok) continue;
        cx<AT> y[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) y[c] = {AT(0), AT(0)};
        if (fl & F_SPMM) {
            int32_t k = a.ptr[row];
            const int32_t e = a.ptr[row + 1];
            for (; k + 4 <= e; k += 4) {
                const int32_t j0 = a.idx[k], j1 = a.idx[k + 1], j2 = a.idx[k + 2], j3 = a.idx[k + 3];
                const MT v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
                V x0[CPL], x1[CPL], x2[CPL], x3[CPL];
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    x0[c] = X[int64_t(j0) * a.ldx + col + c];
                    x1[c] = X[int64_t(j1) * a.ldx + col + c];
                    x2[c] = X[int64_t(j2) * a.ldx + col + c];
                    x3[c] = X[int64_t(j3) * a.ldx + col + c];
                }
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    MV::fma(y[c], v0, to<AT>(x0[c]));
                    MV::fma(y[c], v1, to<AT>(x1[c]));
                    MV::fma(y[c], v2, to<AT>(x2[c]));
                    MV::fma(y[c], v3, to<AT>(x3[c]));
                }
            }
            for (; k < e; ++k) {
                const int32_t j = a.idx[k];
                const MT v = val[k];
#pragma unroll
                for (int c = 0; c < CPL; ++c) MV::fma(y[c], v, to<AT>(X[int64_t(j) * a.ldx + col + c]));
            }
        } else {
#pragma unroll
            for (int c = 0; c < CPL; ++c) y[c] = to<AT>(Z[row * a.ldz + col + c]);
        }
        // Epilogue in FP64 whatever the storage type: one rounding per stored
        // element, and alpha/beta/gamma are never rounded to FP32.
        cx<double> t[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) t[c] = cmul<double>(a.alpha_re, a.alpha_im, to<double>(y[c]));
        if (fl & F_BETA) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const cx<double> xs = cmul<double>(a.beta_re, a.beta_im, to<double>(X[row * a.ldx + col + c]));
                t[c].re += xs.re;
                t[c].im += xs.im;
            }
        }
        if (fl & F_P) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const cx<double> p = to<double>(P[row * a.ldp + col + c]);
                t[c].re = __builtin_fma(a.gamma, p.re, t[c].re);
                t[c].im = __builtin_fma(a.gamma, p.im, t[c].im);
            }
        }
        V ts[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) ts[c] = to<VT>(t[c]);
        if (fl & F_STORE) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) OUT[row * a.ldo + col + c] = ts[c];
        }
        if (fl & F_MOM) {
            // Products of the stored (rounded) values, accumulated in FP64;
            // FP32 x FP32 products are exact in FP64.
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const cx<double> wv = to<double>(W[row * a.ldw + col + c]);
                const cx<double> tv = to<double>(ts[c]);
                mom[c].re = __builtin_fma(wv.re, tv.re, mom[c].re);
                mom[c].re = __builtin_fma(wv.im, tv.im, mom[c].re);
                mom[c].im = __builtin_fma(wv.re, tv.im, mom[c].im);
                mom[c].im = __builtin_fma(-wv.im, tv.re, mom[c].im);
            }
        }
        if (fl & F_ACC) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                cx<double> acc = to<double>(ACC[row * a.lda + col + c]);
                const cx<double> add = cmul<double>(a.coef_re, a.coef_im, t[c]);
                acc.re += add.re;
                acc.im += add.im;
                ACC[row * a.lda + col + c] = to<VT>(acc);
            }
        }
    }

    if (!(fl & F_MOM)) return;
    // Rows sharing a column: lanes differing in bits >= lpr_log2.
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        for (int off = lpr; off < 64; off <<= 1) {
            mom[c].re += __shfl_xor(mom[c].re, off);
            mom[c].im += __shfl_xor(mom[c].im, off);
        }
    }
    __shared__ cx<double> red[4][64 * CPL];
    if (lane < lpr) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) red[wave][lane * CPL + c] = mom[c];
    }
    __syncthreads();
    if (wave == 0 && lane < lpr && colok) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            if (col + c >= a.w) continue;
            cx<double> s = red[0][lane * CPL + c];
            for (int q = 1; q < 4; ++q) {
                s.re += red[q][lane * CPL + c].re;
                s.im += red[q][lane * CPL + c].im;
            }
            reinterpret_cast<cx<double> *>(a.part)[int64_t(blockIdx.x) * a.w + col + c] = s;
        }
    }
}

// out[s][c] = sum_g part[s][g][c] in a fixed tree order (deterministic).
// grid (w, nsteps), 256 threads.  mode 0: real part to out_re[s*ldo + c];
// mode 1: complex to out_cx[s*ldo + c].
__global__ __launch_bounds__(256) void reduce_partials(const double *part, int64_t groups, int32_t w,
                                                       int64_t step_stride, double *out, int64_t ldo,
                                                       int32_t mode) {
    const int c = blockIdx.x, s = blockIdx.y;
    const cx<double> *p = reinterpret_cast<const cx<double> *>(part) + int64_t(s) * step_stride;
    cx<double> acc = {0.0, 0.0};
    for (int64_t g = threadIdx.x; g < groups; g += 256) {
        const cx<double> v = p[g * w + c];
        acc.re += v.re;
        acc.im += v.im;
    }
    __shared__ cx<double> red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            red[threadIdx.x].re += red[threadIdx.x + off].re;
            red[threadIdx.x].im += red[threadIdx.x + off].im;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == 0)
            out[int64_t(s) * ldo + c] = red[0].re;
        else
            reinterpret_cast<cx<double> *>(out)[int64_t(s) * ldo + c] = red[0];
    }
}

// OUT = coef * X (complex scale / copy), any column width, strided rows.
template <class VT>
__global__ __launch_bounds__(256) void scale_kernel(const void *Xv, int64_t ldx, void *Ov, int64_t ldo, int64_t n,
                                                    int32_t w, double cr, double ci) {
    using V = cx<VT>;
    const V *X = static_cast<const V *>(Xv);
    V *O = static_cast<V *>(Ov);
    const int64_t total = n * w;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t r = e / w, c = e - r * w;
        const cx<double> x = to<double>(X[r * ldx + c]);
        O[r * ldo + c] = to<VT>(cx<double>{cr * x.re - ci * x.im, cr * x.im + ci * x.re});
    }
}

// Converts between vector precisions (and strides).
template <class VA, class VB>
__global__ __launch_bounds__(256) void convert_kernel(const void *Xv, int64_t ldx, void *Ov, int64_t ldo, int64_t n,
                                                      int32_t w) {
    const cx<VA> *X = static_cast<const cx<VA> *>(Xv);
    cx<VB> *O = static_cast<cx<VB> *>(Ov);
    const int64_t total = n * w;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t r = e / w, c = e - r * w;
        O[r * ldo + c] = to<VB>(X[r * ldx + c]);
    }
}

// flag |= 1 if any entry of the block is nonfinite.
template <class VT>
__global__ __launch_bounds__(256) void nonfinite_kernel(const void *Xv, int64_t ldx, int64_t n, int32_t w,
                                                        int32_t *flag) {
    const cx<VT> *X = static_cast<const cx<VT> *>(Xv);
    const int64_t total = n * w;
    bool bad = false;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t r = e / w, c = e - r * w;
        const cx<VT> v = X[r * ldx + c];
        bad |= !(__builtin_isfinite(v.re) && __builtin_isfinite(v.im));
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

}  // namespace kmsd
