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
    F_DIAG = 1u << 6,   // the product includes the stored diagonal d_i X_i
    F_ELL = 1u << 7,    // off-diagonal part is ELL (ell entries per row, row-major)
    F_MOM2 = 1u << 8,   // partial.re += |X_i|^2, partial.im += Re conj(X_i) t (moment doubling)
    F_VOUT = 1u << 9,   // VOUT_i = valpha * sum_k V_ik X_k from the same gathers (ell_step_kernel<..., true>)
    F_NORM = 1u << 10,  // partial.re += |t|^2 of the stored t (CSR/ELL step kernels only)
    F_RHS = 1u << 11,   // t += coef * W_i, W a right-hand side (CSR/ELL step and block-Jacobi matrix-core kernels)
};

struct StepArgs {
    const int32_t *ptr;   // CSR row pointers of the off-diagonal part (n+1)
    const int32_t *idx;   // off-diagonal column indices (CSR or ELL n*ell)
    const void *val;      // off-diagonal values (R or cx<R>)
    const void *dval;     // diagonal values (R or cx<R>), n entries
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
    int32_t ell;          // ELL width (F_ELL)
    int64_t uniform_slot; // >= 0: every ELL off-diagonal value equals val[uniform_slot] (padding: zero)
    uint32_t flags;
    const double *vval;   // F_VOUT: real values of a second operator V in the ELL slots of A (n*ell)
    void *VOUT;           // F_VOUT: output block of V X
    int64_t ldv;
    double valpha_re, valpha_im;
    const int32_t *gate;  // if set and *gate == 0 the launch does nothing
    uint32_t tile_plane;  // BSR4: > 0 walks block rows plane-tiled (tile_row), block rows per plane
    uint32_t tile_rows;   // block rows per tile within a plane
    // bsr4_mfma_kernel: byte extents of X, the block values and the block
    // columns (buffer descriptors; reads past them return zero)
    uint32_t xbytes, vbytes, ibytes;
    // bsr4_mfma_kernel<true>: FP32 block factors (16 floats per block row, L
    // below the diagonal, 1/L_ii on it); the product of each block row is
    // replaced by (L L^H)^-1 times it before the epilogue
    const float *bfac32;
};

// Block row at walk position pos of the plane-tiled walk: the nq block rows
// form nq/P planes of P; the walk takes tile t (block rows [tT, tT + T) of
// every plane, the last tile narrower when T does not divide P) plane after
// plane, then the next tile.  A block row's +-P neighbours are then T walk
// positions away instead of P.  Requires nq % P == 0 and nq < 2^32.
__device__ __forceinline__ uint32_t tile_row(uint32_t pos, uint32_t nq, uint32_t P, uint32_t T) {
    const uint32_t planes = nq / P, full = P / T, span = T * planes;
    if (pos < full * span) {
        const uint32_t t = pos / span, w = pos - t * span, x = w / T;
        return x * P + t * T + (w - x * T);
    }
    const uint32_t rem = P - full * T, w = pos - full * span, x = w / rem;
    return x * P + full * T + (w - x * rem);
}

template <class A>
__device__ __forceinline__ cx<A> cmul(double ar, double ai, cx<A> x) {
    return {A(ar) * x.re - A(ai) * x.im, A(ar) * x.im + A(ai) * x.re};
}

// CPL consecutive complex entries of one row as one vector access.
template <class VT, int CPL>
struct Pack {
    typedef VT vec __attribute__((ext_vector_type(2 * CPL)));
    vec v;
    __device__ __forceinline__ void load(const cx<VT> *p) { v = *reinterpret_cast<const vec *>(p); }
    __device__ __forceinline__ void load_nt(const cx<VT> *p) {
        v = __builtin_nontemporal_load(reinterpret_cast<const vec *>(p));
    }
    __device__ __forceinline__ void store(cx<VT> *p) const { *reinterpret_cast<vec *>(p) = v; }
    __device__ __forceinline__ void store_nt(cx<VT> *p) const {
        __builtin_nontemporal_store(v, reinterpret_cast<vec *>(p));
    }
    __device__ __forceinline__ cx<double> get(int c) const { return {double(v[2 * c]), double(v[2 * c + 1])}; }
    template <class A>
    __device__ __forceinline__ cx<A> geta(int c) const { return {A(v[2 * c]), A(v[2 * c + 1])}; }
    __device__ __forceinline__ void set(int c, cx<double> z) {
        v[2 * c] = VT(z.re);
        v[2 * c + 1] = VT(z.im);
    }
};

// Epilogue of one output row: t = alpha y (+ beta X_i) (+ gamma P_i), then the
// optional store, series accumulation and moment partials.  Arithmetic in
// FP64 whatever the storage type: one rounding per stored element, and
// alpha/beta/gamma are never rounded to FP32.
// RHS: F_RHS is compiled in (FP32-valued products of the inner solves only,
// keeping the other instantiations' registers); LEAN: the block-Jacobi
// semi-iteration's epilogue, without accumulation or partial sums.
template <class VT, class AT, int CPL, bool RHS = false, bool LEAN = false>
__device__ __forceinline__ void step_epilogue(const StepArgs &a, uint32_t fl, int64_t row, int col,
                                              const cx<AT> (&y)[CPL], const Pack<VT, CPL> &xi,
                                              const Pack<VT, CPL> &pv, const Pack<VT, CPL> &wv,
                                              cx<double> (&mom)[CPL]) {
    using V = cx<VT>;
    Pack<VT, CPL> ts;
    V *ACC = static_cast<V *>(a.ACC);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        cx<double> t = cmul<double>(a.alpha_re, a.alpha_im, to<double>(y[c]));
        if (fl & F_BETA) {
            const cx<double> xs = cmul<double>(a.beta_re, a.beta_im, xi.get(c));
            t.re += xs.re;
            t.im += xs.im;
        }
        if (fl & F_P) {
            const cx<double> p = pv.get(c);
            t.re = __builtin_fma(a.gamma, p.re, t.re);
            t.im = __builtin_fma(a.gamma, p.im, t.im);
        }
        if constexpr (RHS) {
            if (fl & F_RHS) {
                const cx<double> b = cmul<double>(a.coef_re, a.coef_im, wv.get(c));
                t.re += b.re;
                t.im += b.im;
            }
        }
        ts.set(c, t);
        if (LEAN) continue;  // the block-Jacobi semi-iteration: no accumulation, no partials
        if (fl & F_ACC) {
            cx<double> acc = to<double>(ACC[row * a.lda + col + c]);
            const cx<double> add = cmul<double>(a.coef_re, a.coef_im, t);
            acc.re += add.re;
            acc.im += add.im;
            ACC[row * a.lda + col + c] = to<VT>(acc);
        }
    }
    if (fl & F_STORE) ts.store_nt(static_cast<V *>(a.OUT) + row * a.ldo + col);
    if constexpr (LEAN) return;
    if (fl & F_MOM2) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const cx<double> xv = xi.get(c), tv = ts.get(c);
            mom[c].re = __builtin_fma(xv.re, xv.re, __builtin_fma(xv.im, xv.im, mom[c].re));
            mom[c].im = __builtin_fma(xv.re, tv.re, __builtin_fma(xv.im, tv.im, mom[c].im));
        }
    }
    if (fl & F_NORM) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const cx<double> tv = ts.get(c);
            mom[c].re = __builtin_fma(tv.re, tv.re, __builtin_fma(tv.im, tv.im, mom[c].re));
        }
    }
    if (fl & F_MOM) {
        // Products of the stored (rounded) values, accumulated in FP64;
        // FP32 x FP32 products are exact in FP64.
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const cx<double> wd = wv.get(c), tv = ts.get(c);
            mom[c].re = __builtin_fma(wd.re, tv.re, mom[c].re);
            mom[c].re = __builtin_fma(wd.im, tv.im, mom[c].re);
            mom[c].im = __builtin_fma(wd.re, tv.im, mom[c].im);
            mom[c].im = __builtin_fma(-wd.im, tv.re, mom[c].im);
        }
    }
}

// Workgroup partial moments: lanes sharing a column (bits >= lpr_log2 of the
// lane index) are summed by shuffles, the four wavefronts through LDS, and
// one partial per (workgroup, column) is written.  Fixed order.
template <int CPL>
__device__ __forceinline__ void step_partials(const StepArgs &a, cx<double> (&mom)[CPL], int lane, int wave, int lpr,
                                              int col, bool colok, uint32_t bx) {
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
            reinterpret_cast<cx<double> *>(a.part)[int64_t(bx) * a.w + col + c] = s;
        }
    }
}

// VT: vector element real type (float -> complex64, double -> complex128)
// MR: matrix value real type; MC: complex matrix values
// AT: arithmetic type of the row sums (double for FP64 vectors; float or
//     double for FP32 vectors)
// CPL: complex columns per lane (1, 2 or 4)
//
// Per row the row-local streams (X_i, P_i, W_i) are issued first, then the
// off-diagonal gathers, so their latencies overlap; P and W are read once per
// step and use nontemporal loads, OUT a nontemporal store.
// The body takes its workgroup index bx among gx workgroups and its column
// chunk by (blockIdx.x, gridDim.x, blockIdx.y for one operator per launch;
// the job's own range in a launch shared by several, see step_group_kernel).
template <class VT, class MR, bool MC, class AT, int CPL>
__device__ __forceinline__ void step_body(const StepArgs &a, uint32_t bx, uint32_t gx, uint32_t by) {
    using V = cx<VT>;
    using MV = MatVal<MR, MC>;
    using MT = typename MV::type;
    using PK = Pack<VT, CPL>;
    if (a.gate && *a.gate == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lpr = 1 << a.lpr_log2;
    const int rpw = 64 >> a.lpr_log2;          // rows per wavefront
    const int rpb = 4 * rpw;                    // rows per workgroup pass
    const int rloc = wave * rpw + (lane >> a.lpr_log2);
    const int cgrp = lane & (lpr - 1);
    const int col = (int(by) * lpr + cgrp) * CPL;
    const bool colok = col < a.w;

    const V *X = static_cast<const V *>(a.X);
    const V *Z = static_cast<const V *>(a.Z);
    const V *P = static_cast<const V *>(a.P);
    const V *W = static_cast<const V *>(a.W);
    const MT *val = static_cast<const MT *>(a.val);
    const MT *dval = static_cast<const MT *>(a.dval);
    const uint32_t fl = a.flags;
    const bool need_xi = (fl & F_BETA) || (fl & F_MOM2) || ((fl & F_SPMM) && (fl & F_DIAG));

    cx<double> mom[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) mom[c] = {0.0, 0.0};

    // XCD-aware row-block walk.
    const int64_t nblk = (a.n + rpb - 1) / rpb;
    const uint32_t xcd = bx & 7u, slot = bx >> 3, per = gx >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;

    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t row = blk * rpb + rloc;
        if (row >= a.n || !colok) continue;
        PK xi, pv, wv;
        if (need_xi) xi.load(X + row * a.ldx + col);
        if (fl & F_P) pv.load_nt(P + row * a.ldp + col);
        if (fl & (std::is_same<MR, float>::value ? (F_MOM | F_RHS) : F_MOM)) wv.load_nt(W + row * a.ldw + col);
        cx<AT> y[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) y[c] = {AT(0), AT(0)};
        if (fl & F_SPMM) {
            int64_t k, e;
            if (fl & F_ELL) {
                k = row * a.ell;
                e = k + a.ell;
            } else {
                k = a.ptr[row];
                e = a.ptr[row + 1];
            }
            for (; k + 4 <= e; k += 4) {
                const int32_t j0 = a.idx[k], j1 = a.idx[k + 1], j2 = a.idx[k + 2], j3 = a.idx[k + 3];
                const MT v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
                PK x0, x1, x2, x3;
                x0.load(X + int64_t(j0) * a.ldx + col);
                x1.load(X + int64_t(j1) * a.ldx + col);
                x2.load(X + int64_t(j2) * a.ldx + col);
                x3.load(X + int64_t(j3) * a.ldx + col);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    MV::fma(y[c], v0, x0.template geta<AT>(c));
                    MV::fma(y[c], v1, x1.template geta<AT>(c));
                    MV::fma(y[c], v2, x2.template geta<AT>(c));
                    MV::fma(y[c], v3, x3.template geta<AT>(c));
                }
            }
            if (e - k >= 2) {
                const int32_t j0 = a.idx[k], j1 = a.idx[k + 1];
                const MT v0 = val[k], v1 = val[k + 1];
                PK x0, x1;
                x0.load(X + int64_t(j0) * a.ldx + col);
                x1.load(X + int64_t(j1) * a.ldx + col);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    MV::fma(y[c], v0, x0.template geta<AT>(c));
                    MV::fma(y[c], v1, x1.template geta<AT>(c));
                }
                k += 2;
            }
            if (k < e) {
                const int32_t j0 = a.idx[k];
                const MT v0 = val[k];
                PK x0;
                x0.load(X + int64_t(j0) * a.ldx + col);
#pragma unroll
                for (int c = 0; c < CPL; ++c) MV::fma(y[c], v0, x0.template geta<AT>(c));
            }
            if (fl & F_DIAG) {
                const MT d = dval[row];
#pragma unroll
                for (int c = 0; c < CPL; ++c) MV::fma(y[c], d, xi.template geta<AT>(c));
            }
        } else {
            PK zv;
            zv.load(Z + row * a.ldz + col);
#pragma unroll
            for (int c = 0; c < CPL; ++c) y[c] = zv.template geta<AT>(c);
        }
        step_epilogue<VT, AT, CPL, std::is_same<MR, float>::value>(a, fl, row, col, y, xi, pv, wv, mom);
    }

    if (!(fl & (F_MOM | F_MOM2 | F_NORM))) return;
    step_partials<CPL>(a, mom, lane, wave, lpr, col, colok, bx);
}

template <class VT, class MR, bool MC, class AT, int CPL>
__global__ __launch_bounds__(256) void step_kernel(const StepArgs a) {
    step_body<VT, MR, MC, AT, CPL>(a, blockIdx.x, gridDim.x, blockIdx.y);
}

// step_kernel for an ELL off-diagonal part of compile-time width EW
// (a.ell == EW, F_SPMM | F_ELL set).  The column indices, values and
// diagonal of the NEXT row a lane walks are loaded one iteration ahead, so
// each iteration issues its row-local streams and all EW gathers at once and
// waits for memory once, instead of an index load -> gather -> index load
// chain per row.  Products are summed in the same order as step_kernel
// (entries 0..EW-1, then the diagonal), so both give identical results.
// VO: the launch also forms V X for a second operator V whose entries lie in
// A's slots (real values a.vval, zero where V has none, no diagonal) from the
// gathers it already makes, summed in slot order and scaled as the product
// epilogue scales (so equal to a separate product of V when V's entries come
// in the same column order), and stores it to VOUT.
template <class VT, class MR, bool MC, class AT, int CPL, int EW, bool VO = false>
__device__ __forceinline__ void ell_step_body(const StepArgs &a, uint32_t bx, uint32_t gx, uint32_t by) {
    using V = cx<VT>;
    using MV = MatVal<MR, MC>;
    using MT = typename MV::type;
    using PK = Pack<VT, CPL>;
    if (a.gate && *a.gate == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lpr = 1 << a.lpr_log2;
    const int rpw = 64 >> a.lpr_log2;
    const int rpb = 4 * rpw;
    const int rloc = wave * rpw + (lane >> a.lpr_log2);
    const int cgrp = lane & (lpr - 1);
    const int col = (int(by) * lpr + cgrp) * CPL;
    const bool colok = col < a.w;

    const V *X = static_cast<const V *>(a.X);
    const V *P = static_cast<const V *>(a.P);
    const V *W = static_cast<const V *>(a.W);
    const MT *val = static_cast<const MT *>(a.val);
    const MT *dval = static_cast<const MT *>(a.dval);
    const int32_t *idx = a.idx;
    const uint32_t fl = a.flags;
    const bool diag = fl & F_DIAG;
    const bool need_xi = (fl & F_BETA) || (fl & F_MOM2) || diag;

    cx<double> mom[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) mom[c] = {0.0, 0.0};

    const int64_t nblk = (a.n + rpb - 1) / rpb;
    const uint32_t xcd = bx & 7u, slot = bx >> 3, per = gx >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;

    int32_t jj[EW];
    MT vv[EW];
    MT dd{};
    const bool uni = a.uniform_slot >= 0;
    const MT uval = uni ? val[a.uniform_slot] : MT{};
    auto fetch = [&](int64_t blk, int32_t (&j)[EW], MT (&v)[EW], MT &d) {
        const int64_t r = blk * rpb + rloc;
        if (blk < b1 && r < a.n && colok) {
#pragma unroll
            for (int e = 0; e < EW; ++e) j[e] = idx[r * EW + e];
            if (uni) {
#pragma unroll
                for (int e = 0; e < EW; ++e) v[e] = int64_t(j[e]) == r ? MT{} : uval;  // padding slots hold zero
            } else {
#pragma unroll
                for (int e = 0; e < EW; ++e) v[e] = val[r * EW + e];
            }
            if (diag) d = dval[r];
        }
    };
    fetch(b0 + slot, jj, vv, dd);

    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t row = blk * rpb + rloc;
        const bool active = row < a.n && colok;
        PK xi, pv, wv, xg[EW];
        if (active) {
            if (need_xi) xi.load(X + row * a.ldx + col);
            if (fl & F_P) pv.load_nt(P + row * a.ldp + col);
            if (fl & (std::is_same<MR, float>::value ? (F_MOM | F_RHS) : F_MOM)) wv.load_nt(W + row * a.ldw + col);
#pragma unroll
            for (int e = 0; e < EW; ++e) xg[e].load(X + int64_t(jj[e]) * a.ldx + col);
        }
        double vw[VO ? EW : 1] = {};
        if constexpr (VO) {
            if (active) {
#pragma unroll
                for (int e = 0; e < EW; ++e) vw[e] = a.vval[row * EW + e];
            }
        }
        int32_t jn[EW] = {};
        MT vn[EW] = {};
        MT dn{};
        fetch(blk + per, jn, vn, dn);
        if (active) {
            cx<AT> y[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                y[c] = {AT(0), AT(0)};
#pragma unroll
                for (int e = 0; e < EW; ++e) MV::fma(y[c], vv[e], xg[e].template geta<AT>(c));
                if (diag) MV::fma(y[c], dd, xi.template geta<AT>(c));
            }
            step_epilogue<VT, AT, CPL, std::is_same<MR, float>::value && !VO>(a, fl, row, col, y, xi, pv, wv, mom);
            if constexpr (VO) {
                Pack<VT, CPL> vs;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    cx<AT> yv = {AT(0), AT(0)};
#pragma unroll
                    for (int e = 0; e < EW; ++e) MatVal<double, false>::fma(yv, vw[e], xg[e].template geta<AT>(c));
                    vs.set(c, cmul<double>(a.valpha_re, a.valpha_im, to<double>(yv)));
                }
                vs.store_nt(static_cast<V *>(a.VOUT) + row * a.ldv + col);
            }
        }
#pragma unroll
        for (int e = 0; e < EW; ++e) {
            jj[e] = jn[e];
            vv[e] = vn[e];
        }
        dd = dn;
    }

    if (!(fl & (F_MOM | F_MOM2 | F_NORM))) return;
    step_partials<CPL>(a, mom, lane, wave, lpr, col, colok, bx);
}

template <class VT, class MR, bool MC, class AT, int CPL, int EW, bool VO = false>
__global__ __launch_bounds__(256) void ell_step_kernel(const StepArgs a) {
    ell_step_body<VT, MR, MC, AT, CPL, EW, VO>(a, blockIdx.x, gridDim.x, blockIdx.y);
}

// One chain of a launch shared by several (step_group_kernel): its
// steady-state Chebyshev step, whose gathered input alternates between two
// blocks while the other holds T_(k-1) and receives T_(k+1) in place, its
// partial sums and its workgroups [wg0, wg0 + wgn) of the launch (wg0 a
// multiple of 8, so each chain's XCD-aware row walk is its own launch's).
struct GroupJob {
    StepArgs a;  // X, P, OUT and part are set per launch from the fields below
    void *buf[2];
    double *part0;
    int64_t slot_doubles;
    uint32_t wg0, wgn;
};

// Chain owning workgroup b (chains sorted by wg0).
__device__ __forceinline__ int group_find(const GroupJob *jobs, int njobs, uint32_t b) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].wg0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Steady-state steps of several chains (different matrices, sizes and
// scalings, one column width) in one launch: each workgroup runs its chain's
// step with its index inside the chain's own range, so every chain's row
// partition, partial sums and results are those of a launch of its own.
// parity: buf[parity] is T_k; slot: the partial-sum slot of this order.
// EW: ELL width of the prefetching body, 0 for the generic body.
template <class VT, class MR, bool MC, class AT, int CPL, int EW>
__global__ __launch_bounds__(256) void step_group_kernel(const GroupJob *jobs, int32_t njobs, int32_t parity,
                                                         int32_t slot) {
    const int j = group_find(jobs, njobs, blockIdx.x);
    const GroupJob &g = jobs[j];
    StepArgs a = g.a;
    a.X = g.buf[parity];
    a.P = g.buf[parity ^ 1];
    a.OUT = g.buf[parity ^ 1];
    a.part = g.part0 + int64_t(slot) * g.slot_doubles;
    const uint32_t bx = blockIdx.x - g.wg0;
    if constexpr (EW > 0)
        ell_step_body<VT, MR, MC, AT, CPL, EW>(a, bx, g.wgn, blockIdx.y);
    else
        step_body<VT, MR, MC, AT, CPL>(a, bx, g.wgn, blockIdx.y);
}

// Broadcast of lane E of every 16-lane row to the whole row (DPP
// row_newbcast, one VALU move per 32-bit word).
template <int E, class T>
__device__ __forceinline__ T row_bcast(T v) {
    constexpr int NW = sizeof(T) / 4;
    struct Words {
        int w[NW];
    };
    const Words in = __builtin_bit_cast(Words, v);
    Words out;
#pragma unroll
    for (int i = 0; i < NW; ++i) out.w[i] = __builtin_amdgcn_update_dpp(0, in.w[i], 0x150 + E, 0xF, 0xF, false);
    return __builtin_bit_cast(T, out);
}

template <class T>
__device__ __forceinline__ void row_bcast16(T mine, T (&v)[16]) {
    v[0] = row_bcast<0>(mine);
    v[1] = row_bcast<1>(mine);
    v[2] = row_bcast<2>(mine);
    v[3] = row_bcast<3>(mine);
    v[4] = row_bcast<4>(mine);
    v[5] = row_bcast<5>(mine);
    v[6] = row_bcast<6>(mine);
    v[7] = row_bcast<7>(mine);
    v[8] = row_bcast<8>(mine);
    v[9] = row_bcast<9>(mine);
    v[10] = row_bcast<10>(mine);
    v[11] = row_bcast<11>(mine);
    v[12] = row_bcast<12>(mine);
    v[13] = row_bcast<13>(mine);
    v[14] = row_bcast<14>(mine);
    v[15] = row_bcast<15>(mine);
}

// 4x4 block-sparse (BSR4) variant of step_kernel.  The matrix is stored as
// block rows of four consecutive rows: ptr[q]..ptr[q+1] index the nonzero 4x4
// blocks of block row q, idx their block columns, val 16 row-major values per
// block.  Each lane owns CPL columns of all four rows of a block row, so one
// gather of the four contiguous rows X[4j..4j+3] feeds 16 products per
// column.  When a block row's lane group spans whole 16-lane rows
// (lanes per block row >= 16), lane l loads only value l%16 of the block and
// the row shares them by DPP broadcasts, so the 64 values cost one narrow
// load per lane instead of 16 values per lane through the vector cache.
// The column index of the next block is loaded while the current block's
// gathers are in flight (measured 2.6 % faster on the simple-cubic chain).
// The epilogue is the CSR kernel's, row by row.  n must be a multiple of 4.
template <class MT>
struct alignas(16) Block16 {
    MT v[16];
};

template <class VT, class MR, bool MC, class AT, int CPL, bool DPP>
__global__ __launch_bounds__(256) void bsr4_step_kernel(const StepArgs a) {
    using V = cx<VT>;
    using MV = MatVal<MR, MC>;
    using MT = typename MV::type;
    using PK = Pack<VT, CPL>;
    using BK = Block16<MT>;
    if (a.gate && *a.gate == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lpr = 1 << a.lpr_log2;
    const int qpw = 64 >> a.lpr_log2;           // block rows per wavefront
    const int qpb = 4 * qpw;                    // block rows per workgroup pass
    const int qloc = wave * qpw + (lane >> a.lpr_log2);
    const int cgrp = lane & (lpr - 1);
    const int col = (blockIdx.y * lpr + cgrp) * CPL;
    const bool colok = col < a.w;
    constexpr bool dpp = DPP;  // requires lanes per block row >= 16

    const V *X = static_cast<const V *>(a.X);
    const V *Z = static_cast<const V *>(a.Z);
    const V *P = static_cast<const V *>(a.P);
    const V *W = static_cast<const V *>(a.W);
    const BK *val = static_cast<const BK *>(a.val);
    const MT *vals = static_cast<const MT *>(a.val);
    const uint32_t fl = a.flags;
    const bool need_xi = (fl & F_BETA) || (fl & F_MOM2);

    cx<double> mom[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) mom[c] = {0.0, 0.0};

    const int64_t nq = a.n >> 2;
    const int64_t nblk = (nq + qpb - 1) / qpb;
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;

    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t pos = blk * qpb + qloc;
        // The whole lane group of a block row leaves together; on the DPP path
        // lanes past the last column stay to share their block values.
        if (pos >= nq || (!dpp && !colok)) continue;
        const int64_t q = a.tile_plane ? int64_t(tile_row(uint32_t(pos), uint32_t(nq), a.tile_plane, a.tile_rows)) : pos;
        const int64_t row0 = q * 4;
        cx<AT> y[4][CPL];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) y[r][c] = {AT(0), AT(0)};
        if (fl & F_SPMM) {
            const int64_t k0 = a.ptr[q], e = a.ptr[q + 1];
            if constexpr (dpp) {
                int32_t jn = k0 < e ? a.idx[k0] : 0;  // block column, one block ahead
                for (int64_t k = k0; k < e; ++k) {
                    const int64_t j = int64_t(jn) * 4;
                    if (k + 1 < e) jn = a.idx[k + 1];
                    PK x0[4];
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc) {
                        if (colok)
                            x0[cc].load(X + (j + cc) * a.ldx + col);
                        else
                            x0[cc].v = {};
                    }
                    MT v[16];
                    row_bcast16(vals[k * 16 + (lane & 15)], v);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                            for (int c = 0; c < CPL; ++c) MV::fma(y[r][c], v[r * 4 + cc], x0[cc].template geta<AT>(c));
                }
            } else {
                int32_t jn = k0 < e ? a.idx[k0] : 0;  // block column, one block ahead
                for (int64_t k = k0; k < e; ++k) {
                    const int64_t j = int64_t(jn) * 4;
                    if (k + 1 < e) jn = a.idx[k + 1];
                    PK x0[4];
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc) x0[cc].load(X + (j + cc) * a.ldx + col);
                    const BK v0 = val[k];
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                            for (int c = 0; c < CPL; ++c)
                                MV::fma(y[r][c], v0.v[r * 4 + cc], x0[cc].template geta<AT>(c));
                }
            }
        } else if (colok) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                PK zv;
                zv.load(Z + (row0 + r) * a.ldz + col);
#pragma unroll
                for (int c = 0; c < CPL; ++c) y[r][c] = zv.template geta<AT>(c);
            }
        }
        if (!colok) continue;
        // Row-local streams after the gathers: fewer live registers across
        // the block loop (higher occupancy hides the gather latency).
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            PK xi, pv, wv;
            if (need_xi) xi.load(X + (row0 + r) * a.ldx + col);
            if (fl & F_P) pv.load_nt(P + (row0 + r) * a.ldp + col);
            if (fl & F_MOM) wv.load_nt(W + (row0 + r) * a.ldw + col);
            step_epilogue<VT, AT, CPL>(a, fl, row0 + r, col, y[r], xi, pv, wv, mom);
        }
    }

    if (!(fl & (F_MOM | F_MOM2))) return;
    step_partials<CPL>(a, mom, lane, wave, lpr, col, colok, blockIdx.x);
}

// FP32-valued 4x4-block product on complex64 blocks with the block products
// on the matrix cores (the PCG metric products of complex64 inner solves).
// Flags: F_SPMM with F_STORE, F_MOM and F_P only; real (or imaginary-only)
// values; w a multiple of 4.
//
// Lane layout: lanes per block row lpr = 2^lpr_log2 in {4, 8, 16}, lane g of
// a block row owns complex columns 4g..4g+3 of its four rows (grid.y walks
// column chunks of 4*lpr).  Each lane quad (one block row, 16 columns) is
// one block of v_mfma_f32_4x4x1_16b: for block column cc, lane 4b+i supplies
// S[i][cc] as the A operand and its own X[4j+cc][col].re/.im as the B
// operand (one instruction per real column t of the lane), and D[i][n]
// lands in lane 4b+n, register i: the lane's own output rows.  Sums are
// FP32 over the whole block row.
//
// A block row's column indices and values are staged in LDS per chunk of
// KC = 10*lpr/4 blocks by buffer-to-LDS loads issued together (one memory
// latency per chunk), so a block waits only for its X rows, which are read
// by buffer loads with 32-bit offsets.  LDS per wavefront: 10 x 1 KB of
// values (instruction i holds blocks kc + i*lpr/4 .. of every block row,
// 16 bytes per lane) and 3 x 256 B of indices.
// The same XCD-aware, optionally plane-tiled, block row walk as
// bsr4_step_kernel; partials [gridDim.x][w] in the same format.
// z = (L L^H)^-1 v for one 4x4 block with a real factor in FP32: L below the
// diagonal, 1/L_ii on it (products instead of divisions), FP32 arithmetic
// (complex64 blocks), in place on the real and imaginary parts.
__device__ __forceinline__ void blk4_solve_f32(const float *L, float (&re)[4], float (&im)[4]) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 l0 = *reinterpret_cast<const f4 *>(L), l1 = *reinterpret_cast<const f4 *>(L + 4),
             l2 = *reinterpret_cast<const f4 *>(L + 8), l3 = *reinterpret_cast<const f4 *>(L + 12);
    const float m[16] = {l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3],
                         l2[0], l2[1], l2[2], l2[3], l3[0], l3[1], l3[2], l3[3]};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float sr = re[i], si = im[i];
#pragma unroll
        for (int k = 0; k < i; ++k) {
            sr = __builtin_fmaf(-m[i * 4 + k], re[k], sr);
            si = __builtin_fmaf(-m[i * 4 + k], im[k], si);
        }
        re[i] = sr * m[i * 5];
        im[i] = si * m[i * 5];
    }
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        float sr = re[i], si = im[i];
#pragma unroll
        for (int k = i + 1; k < 4; ++k) {
            sr = __builtin_fmaf(-m[k * 4 + i], re[k], sr);
            si = __builtin_fmaf(-m[k * 4 + i], im[k], si);
        }
        re[i] = sr * m[i * 5];
        im[i] = si * m[i * 5];
    }
}

// The same on complex values held in FP64 (rounded to FP32 first).
__device__ __forceinline__ void blk4_solve_r32(const float *L, cx<double> (&v)[4]) {
    float re[4], im[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        re[i] = float(v[i].re);
        im[i] = float(v[i].im);
    }
    blk4_solve_f32(L, re, im);
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = {double(re[i]), double(im[i])};
}

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, int(bytes), 0x00020000);
}

template <bool BLK>
__device__ __forceinline__ void bsr4_mfma_body(const StepArgs &a) {
    using V = cx<float>;
    constexpr int CPL = 4, NV = 10, NI = 3;
    using PK = Pack<float, CPL>;
    typedef float f4 __attribute__((ext_vector_type(4)));
    if (a.gate && *a.gate == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lg = a.lpr_log2, lpr = 1 << lg;
    const int qpw = 64 >> lg, qpb = 4 * qpw;
    const int qw = lane >> lg;
    const int cgrp = lane & (lpr - 1);
    const int col = (int(blockIdx.y) * lpr + cgrp) * CPL;
    const bool colok = col < a.w;
    const int ri = lane & 3;
    const int bpi = lpr >> 2, kc_max = NV * bpi;  // blocks per value instruction and row; chunk length
    const V *P = static_cast<const V *>(a.P);
    const V *W = static_cast<const V *>(a.W);
    const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(a.X, a.xbytes), vr = buffer_rsrc(a.val, a.vbytes),
                                 ir = buffer_rsrc(a.idx, a.ibytes);
    const uint32_t fl = a.flags;
    __shared__ f4 svals[4][NV][64];
    __shared__ int32_t sidx[4][NI][64];
    cx<double> mom[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) mom[c] = {0.0, 0.0};
    const int64_t nq = a.n >> 2;
    const int64_t nblk = (nq + qpb - 1) / qpb;
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;
    const uint32_t rowb = uint32_t(a.ldx) * 8u, colb = uint32_t(col) * 8u;
    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t pos = blk * qpb + wave * qpw + qw;
        const bool valid = pos < nq;
        const int64_t q =
            !valid ? 0 : (a.tile_plane ? int64_t(tile_row(uint32_t(pos), uint32_t(nq), a.tile_plane, a.tile_rows)) : pos);
        int32_t k0 = 0, cnt = 0;
        if (valid) {
            k0 = a.ptr[q];
            cnt = a.ptr[q + 1] - k0;
        }
        int kmax = cnt;  // longest block row of the wavefront: the MFMAs need every lane
        for (int off = lpr; off < 64; off <<= 1) kmax = max(kmax, __shfl_xor(kmax, off));
        f4 d[2 * CPL];
#pragma unroll
        for (int t = 0; t < 2 * CPL; ++t) d[t] = f4{0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < kmax; kc += kc_max) {
            // The previous chunk's LDS reads are done before it is overwritten.
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < NI; ++i)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ir, (lds_void *)&sidx[wave][i][0], 4,
                                                         uint32_t(k0 + kc + i * lpr + cgrp) * 4u, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < NV; ++i)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(vr, (lds_void *)&svals[wave][i][0], 16,
                                                         uint32_t(k0 + kc + i * bpi) * 64u + uint32_t(cgrp) * 16u, 0,
                                                         0, 0);
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            const int kend = kmax - kc < kc_max ? kmax - kc : kc_max;
            for (int kq = 0; kq < kend; ++kq) {
                const bool on = kc + kq < cnt;
                const int32_t j = sidx[wave][kq >> lg][qw * lpr + (kq & (lpr - 1))];
                const uint32_t xo = uint32_t(j) * 4u * rowb + colb;
                PK x[4];
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    x[cc].v = {};
                    if (on && colok) {
                        const f4 lo = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(xr, xo + cc * rowb, 0, 0));
                        const f4 hi =
                            __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(xr, xo + cc * rowb + 16, 0, 0));
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            x[cc].v[e] = lo[e];
                            x[cc].v[4 + e] = hi[e];
                        }
                    }
                }
                const int bi = kq / bpi, bb = kq - bi * bpi;
                f4 s = svals[wave][bi][qw * lpr + bb * 4 + ri];
                if (!on) s = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                    for (int t = 0; t < 2 * CPL; ++t)
                        d[t] = __builtin_amdgcn_mfma_f32_4x4x1f32(s[cc], x[cc].v[t], d[t], 0, 0, 0);
            }
        }
        if (!valid || !colok) continue;
        const int64_t row0 = q * 4;
        if constexpr (BLK) {
            // block-Jacobi: each column's four rows of this block row times
            // (L L^H)^-1 of the block row's factor
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                float re[4], im[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    re[r] = d[2 * c][r];
                    im[r] = d[2 * c + 1][r];
                }
                blk4_solve_f32(a.bfac32 + q * 16, re, im);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    d[2 * c][r] = re[r];
                    d[2 * c + 1][r] = im[r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            PK xi, pv, wv;
            cx<double> y[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) y[c] = {double(d[2 * c][r]), double(d[2 * c + 1][r])};
            if constexpr (BLK) {
                if (fl & F_BETA) xi.load(static_cast<const V *>(a.X) + (row0 + r) * a.ldx + col);
            }
            if (fl & F_P) pv.load_nt(P + (row0 + r) * a.ldp + col);
            if (fl & (BLK ? F_RHS : F_MOM)) wv.load_nt(W + (row0 + r) * a.ldw + col);
            step_epilogue<float, double, CPL, BLK, BLK>(a, fl, row0 + r, col, y, xi, pv, wv, mom);
        }
    }
    if constexpr (BLK) return;
    if (!(fl & F_MOM)) return;
    // Lanes of one column (lane bits >= lpr_log2) by shuffles, the four
    // wavefronts through LDS, one partial per (workgroup, column).
#pragma unroll
    for (int c = 0; c < CPL; ++c)
        for (int off = lpr; off < 64; off <<= 1) {
            mom[c].re += __shfl_xor(mom[c].re, off);
            mom[c].im += __shfl_xor(mom[c].im, off);
        }
    __shared__ cx<double> red[4][16 * CPL];
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
            for (int w2 = 1; w2 < 4; ++w2) {
                s.re += red[w2][lane * CPL + c].re;
                s.im += red[w2][lane * CPL + c].im;
            }
            reinterpret_cast<cx<double> *>(a.part)[int64_t(blockIdx.x) * a.w + col + c] = s;
        }
    }
}

static __global__ __launch_bounds__(256) void bsr4_mfma_kernel(const StepArgs a) { bsr4_mfma_body<false>(a); }

// The same product with a 4x4 block-Jacobi solve of each block row's result
// before the epilogue (a.bfac32), which adds the own row (F_BETA), the
// previous iterate (F_P) and the preconditioned right-hand side (F_RHS): one
// step of the block-Jacobi Chebyshev semi-iteration in one launch.
// (three wavefronts per SIMD, as the plain product, for its gather latency)
static __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void bsr4_mfma_blk_kernel(
    const StepArgs a) {
    bsr4_mfma_body<true>(a);
}

// FP64-valued 4x4-block product (real or imaginary-only values) with the
// block columns and values of each chunk of blocks staged in LDS, as in
// bsr4_mfma_kernel, and FP64 row sums on the VALU in bsr4_step_kernel's
// order (so its products are bitwise those of bsr4_step_kernel): the
// chain's H products of complex64 or complex128 blocks and the FP64
// residuals of the metric refinement.  Flags: F_SPMM with F_BETA, F_P,
// F_STORE and F_MOM.  Lane layout of bsr4_step_kernel (lpr = 2^lpr_log2 >= 8
// lanes per block row, CPL columns each); a wavefront stages KC = lpr blocks
// of each of its block rows per chunk: 8 buffer-to-LDS loads of 16 bytes per
// lane (instruction i holds blocks kc + i*lpr/8 .. of every block row, 1 KB)
// and one of the block columns.  DPP: lanes per block row >= 16, each lane
// reads one of the block's 16 values and the row shares them by DPP
// broadcasts; otherwise each lane reads all 16.
template <class VT, int CPL, bool DPP>
__global__ __launch_bounds__(256) void bsr4_lds_kernel(const StepArgs a) {
    using V = cx<VT>;
    using MV = MatVal<double, false>;
    using PK = Pack<VT, CPL>;
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef float f4 __attribute__((ext_vector_type(4)));
    if (a.gate && *a.gate == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lg = a.lpr_log2, lpr = 1 << lg;
    const int qpw = 64 >> lg, qpb = 4 * qpw;
    const int qw = lane >> lg;
    const int cgrp = lane & (lpr - 1);
    const int col = (int(blockIdx.y) * lpr + cgrp) * CPL;
    const bool colok = col < a.w;
    const int bpi = lpr >> 3, kc_max = lpr;  // blocks per value load and row; chunk length
    const V *P = static_cast<const V *>(a.P);
    const V *W = static_cast<const V *>(a.W);
    const V *X = static_cast<const V *>(a.X);
    const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(a.X, a.xbytes), vr = buffer_rsrc(a.val, a.vbytes),
                                 ir = buffer_rsrc(a.idx, a.ibytes);
    const uint32_t fl = a.flags;
    const bool need_xi = (fl & F_BETA) != 0;
    __shared__ d2 svals[4][8][64];
    __shared__ int32_t sidx[4][64];
    cx<double> mom[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) mom[c] = {0.0, 0.0};
    const int64_t nq = a.n >> 2;
    const int64_t nblk = (nq + qpb - 1) / qpb;
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;
    const uint32_t rowb = uint32_t(a.ldx) * uint32_t(sizeof(V)), colb = uint32_t(col) * uint32_t(sizeof(V));
    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t pos = blk * qpb + wave * qpw + qw;
        const bool valid = pos < nq;
        const int64_t q =
            !valid ? 0 : (a.tile_plane ? int64_t(tile_row(uint32_t(pos), uint32_t(nq), a.tile_plane, a.tile_rows)) : pos);
        int32_t k0 = 0, cnt = 0;
        if (valid) {
            k0 = a.ptr[q];
            cnt = a.ptr[q + 1] - k0;
        }
        int kmax = cnt;  // the wavefront stages chunks together
        for (int off = lpr; off < 64; off <<= 1) kmax = max(kmax, __shfl_xor(kmax, off));
        cx<double> y[4][CPL];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < CPL; ++c) y[r][c] = {0.0, 0.0};
        for (int kc = 0; kc < kmax; kc += kc_max) {
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ir, (lds_void *)&sidx[wave][0], 4, uint32_t(k0 + kc + cgrp) * 4u, 0,
                                                     0, 0);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(vr, (lds_void *)&svals[wave][i][0], 16,
                                                         uint32_t(k0 + kc + i * bpi) * 128u + uint32_t(cgrp) * 16u, 0,
                                                         0, 0);
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            const int kend = min(kmax - kc, kc_max);
            for (int kq = 0; kq < kend; ++kq) {
                if (kc + kq >= cnt) continue;  // this block row is done (its lanes idle to the chunk's end)
                const int32_t j = sidx[wave][qw * lpr + kq];
                const uint32_t xo = uint32_t(j) * 4u * rowb + colb;
                PK x0[4];
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    x0[cc].v = {};
                    if (colok) {
                        if constexpr (sizeof(VT) * 2 * CPL == 32) {
                            const f4 lo = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(xr, xo + cc * rowb, 0, 0));
                            const f4 hi =
                                __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(xr, xo + cc * rowb + 16, 0, 0));
                            typedef VT vx __attribute__((ext_vector_type(16 / sizeof(VT))));
                            const vx l2 = __builtin_bit_cast(vx, lo), h2 = __builtin_bit_cast(vx, hi);
#pragma unroll
                            for (int e = 0; e < int(16 / sizeof(VT)); ++e) {
                                x0[cc].v[e] = l2[e];
                                x0[cc].v[int(16 / sizeof(VT)) + e] = h2[e];
                            }
                        } else {
                            x0[cc].load(X + (int64_t(j) * 4 + cc) * a.ldx + col);
                        }
                    }
                }
                const int bi = kq / bpi, bb = kq - bi * bpi;
                const double *vb = reinterpret_cast<const double *>(&svals[wave][bi][qw * lpr + bb * 8]);
                double v[16];
                if constexpr (DPP) {
                    row_bcast16(vb[lane & 15], v);
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const d2 t = reinterpret_cast<const d2 *>(vb)[e];
                        v[2 * e] = t[0];
                        v[2 * e + 1] = t[1];
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                        for (int c = 0; c < CPL; ++c) MV::fma(y[r][c], v[r * 4 + cc], x0[cc].template geta<double>(c));
            }
        }
        if (!valid || !colok) continue;
        const int64_t row0 = q * 4;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            PK xi, pv, wv;
            if (need_xi) xi.load(X + (row0 + r) * a.ldx + col);
            if (fl & F_P) pv.load_nt(P + (row0 + r) * a.ldp + col);
            if (fl & F_MOM) wv.load_nt(W + (row0 + r) * a.ldw + col);
            step_epilogue<VT, double, CPL>(a, fl, row0 + r, col, y[r], xi, pv, wv, mom);
        }
    }
    if (!(fl & F_MOM)) return;
#pragma unroll
    for (int c = 0; c < CPL; ++c)
        for (int off = lpr; off < 64; off <<= 1) {
            mom[c].re += __shfl_xor(mom[c].re, off);
            mom[c].im += __shfl_xor(mom[c].im, off);
        }
    __shared__ cx<double> red[4][32 * CPL];
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
            for (int w2 = 1; w2 < 4; ++w2) {
                s.re += red[w2][lane * CPL + c].re;
                s.im += red[w2][lane * CPL + c].im;
            }
            reinterpret_cast<cx<double> *>(a.part)[int64_t(blockIdx.x) * a.w + col + c] = s;
        }
    }
}

// out[s][c] = sum_g part[s][g][c] in a fixed tree order (deterministic).
// grid (w, nsteps), 256 threads.  mode 0: real part to out_re[s*ldo + c];
// mode 1: complex to out_cx[s*ldo + c].
static __global__ __launch_bounds__(256) void reduce_partials(const double *part, int64_t groups, int32_t w,
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

// convert_kernel with flag |= 1 if any converted entry is nonfinite;
// thread 0 of workgroup 0 zeroes clear (another flag, for the next launch).
template <class VA, class VB>
__global__ __launch_bounds__(256) void convert_check_kernel(const void *Xv, int64_t ldx, void *Ov, int64_t ldo,
                                                            int64_t n, int32_t w, int32_t *flag, int32_t *clear) {
    const cx<VA> *X = static_cast<const cx<VA> *>(Xv);
    cx<VB> *O = static_cast<cx<VB> *>(Ov);
    const int64_t total = n * w;
    bool bad = false;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t r = e / w, c = e - r * w;
        const cx<VB> v = to<VB>(X[r * ldx + c]);
        O[r * ldo + c] = v;
        bad |= !(__builtin_isfinite(v.re) && __builtin_isfinite(v.im));
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear = 0;
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

namespace kmsd {

// W = Z S Z with Z = diag(d)^-1/2: scales the stored values of a copy of S.
// One thread per row; ELL rows are [row*ell, row*ell+ell), CSR rows ptr[row]..
template <class R, bool MC>
__global__ __launch_bounds__(256) void congruence_scale_kernel(void *offv, void *diagv, const int32_t *ptr,
                                                               const int32_t *idx, int32_t ell, const double *d,
                                                               int64_t n) {
    using MT = typename MatVal<R, MC>::type;
    MT *off = static_cast<MT *>(offv);
    MT *dg = static_cast<MT *>(diagv);
    for (int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x; r < n; r += int64_t(gridDim.x) * 256) {
        const double zr = 1.0 / sqrt(d[r]);
        const int64_t k0 = ell ? r * ell : ptr[r], k1 = ell ? k0 + ell : ptr[r + 1];
        for (int64_t k = k0; k < k1; ++k) {
            const double s = zr / sqrt(d[idx[k]]);
            if constexpr (MC) {
                off[k].re = R(double(off[k].re) * s);
                off[k].im = R(double(off[k].im) * s);
            } else {
                off[k] = R(double(off[k]) * s);
            }
        }
        const double s = zr * zr;
        if constexpr (MC) {
            dg[r].re = R(double(dg[r].re) * s);
            dg[r].im = R(double(dg[r].im) * s);
        } else {
            dg[r] = R(double(dg[r]) * s);
        }
    }
}

// BSR4 form of the congruence scaling: one thread per stored block entry.
template <class R, bool MC>
__global__ __launch_bounds__(256) void bsr4_congruence_kernel(void *valv, const int32_t *ptr, const int32_t *idx,
                                                              const double *d, int64_t nq, int64_t nnzb) {
    using MT = typename MatVal<R, MC>::type;
    MT *val = static_cast<MT *>(valv);
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < nnzb * 16; e += int64_t(gridDim.x) * 256) {
        const int64_t k = e >> 4;
        const int r = int(e & 15) >> 2, cc = int(e & 3);
        // block row of entry k: largest q with ptr[q] <= k (binary search)
        int64_t lo = 0, hi = nq;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (ptr[mid] <= k) lo = mid; else hi = mid;
        }
        const double s = 1.0 / sqrt(d[lo * 4 + r] * d[int64_t(idx[k]) * 4 + cc]);
        if constexpr (MC) {
            val[e].re = R(double(val[e].re) * s);
            val[e].im = R(double(val[e].im) * s);
        } else {
            val[e] = R(double(val[e]) * s);
        }
    }
}

// (R, C) grid of pixels of W 8-byte words -> (C, R) grid.  W <= 4: 32 x 32
// pixel tiles through LDS (coalesced reads and writes); one pixel per thread
// per pass, block (32, 8).
template <int W>
__global__ __launch_bounds__(256) void pixel_transpose_tiled(const uint64_t *in, uint64_t *out, int64_t R,
                                                             int64_t C) {
    __shared__ uint64_t tile[32][33 * W];
    const int64_t tiles_c = (C + 31) / 32;
    const int64_t c0 = (int64_t(blockIdx.x) % tiles_c) * 32, r0 = (int64_t(blockIdx.x) / tiles_c) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int j = 0; j < 32; j += 8) {
        const int64_t r = r0 + ty + j, c = c0 + tx;
        if (r < R && c < C) {
#pragma unroll
            for (int q = 0; q < W; ++q) tile[ty + j][tx * W + q] = in[(r * C + c) * W + q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 32; j += 8) {
        const int64_t c = c0 + ty + j, r = r0 + tx;
        if (r < R && c < C) {
#pragma unroll
            for (int q = 0; q < W; ++q) out[(c * R + r) * W + q] = tile[tx][(ty + j) * W + q];
        }
    }
}

// Same for pixels of any width (W words, W > 4: every pixel is a contiguous
// segment of >= 40 bytes, so direct copies stay coalesced enough).
static __global__ __launch_bounds__(256) void pixel_transpose_direct(const uint64_t *in, uint64_t *out, int64_t R,
                                                                     int64_t C, int64_t W) {
    const int64_t total = R * C * W;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t pix = e / W, q = e - pix * W;
        const int64_t r = pix / C, c = pix - r * C;
        out[(c * R + r) * W + q] = in[e];
    }
}

// out_i = z_i (out_i + c0 xi_i), z = diag^-1/2 (final Clenshaw step of the sampler)
template <class VT>
__global__ __launch_bounds__(256) void sample_finish_kernel(void *Ov, const void *Xv, int64_t ld, int64_t n, int32_t w,
                                                            const double *d, double c0) {
    cx<VT> *O = static_cast<cx<VT> *>(Ov);
    const cx<VT> *X = static_cast<const cx<VT> *>(Xv);
    const int64_t total = n * w;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t r = e / w, c = e - r * w;
        const double z = 1.0 / sqrt(d[r]);
        const cx<double> o = to<double>(O[r * ld + c]), x = to<double>(X[r * ld + c]);
        O[r * ld + c] = to<VT>(cx<double>{z * (o.re + c0 * x.re), z * (o.im + c0 * x.im)});
    }
}

}  // namespace kmsd

namespace kmsd {

// Moment doubling: ab[k][j] = (<T_k,T_k>, Re<T_k,T_{k+1}>) -> mu[2k] = 2a_k - mu_0,
// mu[2k+1] = 2b_k - mu_1 with mu_0 = a_0, mu_1 = b_0 (T_{2k} = 2T_k^2 - 1,
// T_{2k+1} = 2T_kT_{k+1} - T_1).
static __global__ __launch_bounds__(256) void doubling_finish_kernel(const double *ab, double *mu, int64_t L, int32_t w) {
    const cx<double> *AB = reinterpret_cast<const cx<double> *>(ab);
    const int64_t total = L * w;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t l = e / w, j = e - l * w, k = l / 2;
        const double mu0 = AB[j].re, mu1 = AB[j].im;
        if (l == 0)
            mu[e] = mu0;
        else if (l == 1)
            mu[e] = mu1;
        else
            mu[e] = (l & 1) ? 2.0 * AB[k * w + j].im - mu1 : 2.0 * AB[k * w + j].re - mu0;
    }
}

}  // namespace kmsd
