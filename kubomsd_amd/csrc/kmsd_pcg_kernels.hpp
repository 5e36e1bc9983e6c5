// Device kernels of the batched PCG metric solve.  One independent system per
// column; every per-column scalar (scale, threshold, alpha, beta, <r,z>,
// activity, breakdown, iteration count) lives in device memory, so the
// iteration needs no host round trip; the host only polls a small flag block
// every few iterations.  Vector passes use the same row/lane mapping as the
// fused step kernel; column reductions go through [groups][w] partials that
// the per-column kernels reduce in a fixed order (deterministic).
#pragma once
#include <type_traits>
#include "kmsd_kernels.hpp"

namespace kmsd {

struct ColState {
    double safe, thr, qn, rz, alpha, beta, rn, tn, gap;
    int32_t active, restart, breakdown, fin;
    int64_t iters;
};

struct PcgFlags {
    // Activity counts by iteration parity s = k & 1: any_active[s] counts the
    // columns active at the start of iteration k (the set-up writes slot 0;
    // iteration k's PC_BETA / PC_RN_BETA count into slot s ^ 1); updated[s]
    // counts the columns active when iteration k's PC_RN_BETA began, i.e.
    // whose alpha the iteration's P update still has to apply to Y.  Iteration
    // k's PC_ALPHA launch clears any_active[s ^ 1] and updated[s] (their last
    // readers ran in iteration k - 1 or k - 2), so the loop needs no memset.
    int32_t any_active[2], updated[2], broke, need_refresh, nonfinite;
    int64_t refreshes, restarts, checks;
};

enum : int32_t {
    PS_ABSMAX = 0,   // partial: max(|re|,|im|) of B          (part re)
    PS_SCALE = 1,    // Q = B/safe, R = Q, Y = 0 ; partial |Q|^2
    PS_PINIT = 2,    // P = active ? Z : 0
    PS_UPDATE = 3,   // Y += alpha P, R -= alpha SP ; partial |R|^2
    PS_RESTART = 4,  // R = restart ? T : R
    PS_DIAG = 5,     // Z = R / d ; partial Re<R,Z>
    PS_DOT = 6,      // partial Re<R,Z> (block preconditioner applied separately)
    PS_PUPDATE = 7,  // P = active ? Z + (restart ? 0 : beta P) : 0
    PS_TRUEGAP = 8,  // partial |T|^2 (re) and |T-R|^2 (im)
    PS_RESCALE = 9,  // OUT = Y * safe ; flags nonfinite columns
    // Fused iteration (fixed residual policy, no refresh this iteration);
    // z = M^-1 R is recomputed where needed instead of stored.
    // Y += alpha P waits for the P update, which reads P anyway: the two
    // passes move 3 + 5 block streams instead of 6 + 3, and Y gets the same
    // values (same operands, same expression).
    PS_UPD_PRE = 10,   // R -= alpha SP ; partial |R|^2 (re), Re<R,z> (im)
    PS_PUPD_PRE = 11,  // Y += alpha P ; P = active ? z + (restart ? 0 : beta P) : 0
    PS_PRE_DOT = 12,   // uniform 4x4 blocks: Z = M^-1 R stored ; partial Re<R,Z> (one pass for z and the dot)
    // Set-up with the diagonal preconditioner: PS_SCALE and PS_DIAG in one
    // pass (Q = R = B/safe, Y = 0, Z = R/d; partials |Q|^2 (re), Re<R,Z> (im)).
    PS_SCALE_DIAG = 13,
};

struct PcgArgs {
    const void *B;  // input RHS (ldb)
    void *Q, *Y, *R, *Z, *P;
    const void *SP, *T;
    void *OUT;
    const double *diag;
    const double *fac;  // uniform 4x4 block factors, 16 complex entries per block (quad kernel)
    const float *fac32; // the same factors when real, FP32 with 1/L_ii on the diagonal (complex64 quad kernel)
    ColState *cs;
    PcgFlags *fl;
    double *part;
    int64_t n, ldb, ldo;  // internal blocks have ld = w
    int32_t w, lpr_log2, stage, require_active;  // require_active: skip after breakdown or with no active column
    int32_t slot;         // flag slot read by require_active (any_active[slot], or updated[slot] for PS_PUPD_PRE)
    const int32_t *gate;  // if set and *gate == 0 the launch does nothing
};

template <class VT, int CPL>
__global__ __launch_bounds__(256) void pcg_row_kernel(const PcgArgs a) {
    using V = cx<VT>;
    using PK = Pack<VT, CPL>;
    // In-loop stages stop after a curvature breakdown (the reference leaves
    // the loop); set-up and final-acceptance stages always run.
    // (the fused P update also finishes the iteration's Y update, so it runs
    // when any column was active at the iteration's start)
    const int32_t live = a.stage == PS_PUPD_PRE ? a.fl->updated[a.slot] : a.fl->any_active[a.slot];
    if (a.require_active && (a.fl->broke || !live)) return;
    if (a.gate && *a.gate == 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lpr = 1 << a.lpr_log2;
    const int rpw = 64 >> a.lpr_log2, rpb = 4 * rpw;
    const int rloc = wave * rpw + (lane >> a.lpr_log2);
    const int cgrp = lane & (lpr - 1);
    const int col = (blockIdx.y * lpr + cgrp) * CPL;
    const bool colok = col < a.w;
    const int64_t w = a.w;

    double acc_re[CPL], acc_im[CPL];
    double cf[CPL], cf2[CPL];
    int32_t act[CPL], rst[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        acc_re[c] = acc_im[c] = 0.0;
        cf[c] = cf2[c] = 0.0;
        act[c] = rst[c] = 0;
        if (colok) {
            const ColState &s = a.cs[col + c];
            act[c] = s.active;
            rst[c] = s.restart;
            cf[c] = (a.stage == PS_SCALE || a.stage == PS_SCALE_DIAG) ? 1.0 / s.safe
                    : (a.stage == PS_UPDATE || a.stage == PS_UPD_PRE) ? s.alpha
                    : (a.stage == PS_PUPDATE || a.stage == PS_PUPD_PRE) ? s.beta
                    : (a.stage == PS_RESCALE) ? s.safe
                                              : 0.0;
            if (a.stage == PS_PUPD_PRE) cf2[c] = s.alpha;
        }
    }
    const int64_t nblk = (a.n + rpb - 1) / rpb;
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;
    V *Q = static_cast<V *>(a.Q);
    V *Y = static_cast<V *>(a.Y);
    V *R = static_cast<V *>(a.R);
    V *Z = static_cast<V *>(a.Z);
    V *P = static_cast<V *>(a.P);
    const V *SP = static_cast<const V *>(a.SP);
    const V *T = static_cast<const V *>(a.T);
    const V *B = static_cast<const V *>(a.B);
    V *OUT = static_cast<V *>(a.OUT);
    bool bad[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) bad[c] = false;

    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t row = blk * rpb + rloc;
        if (row >= a.n || !colok) continue;
        const int64_t o = row * w + col;
        switch (a.stage) {
            case PS_ABSMAX: {
                PK b;
                b.load(B + row * a.ldb + col);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> v = b.get(c);
                    acc_re[c] = fmax(acc_re[c], fmax(fabs(v.re), fabs(v.im)));
                }
                break;
            }
            case PS_SCALE: {
                PK b, q;
                b.load(B + row * a.ldb + col);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    cx<double> v = b.get(c);
                    v.re *= cf[c];
                    v.im *= cf[c];
                    q.set(c, v);
                    const cx<double> qv = q.get(c);
                    acc_re[c] += qv.re * qv.re + qv.im * qv.im;
                }
                q.store(Q + o);
                q.store(R + o);
                PK zero;
#pragma unroll
                for (int c = 0; c < CPL; ++c) zero.set(c, cx<double>{0.0, 0.0});
                zero.store(Y + o);
                break;
            }
            case PS_SCALE_DIAG: {
                PK b, q, z;
                b.load(B + row * a.ldb + col);
                const double d = a.diag[row];
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    cx<double> v = b.get(c);
                    v.re *= cf[c];
                    v.im *= cf[c];
                    q.set(c, v);
                    const cx<double> qv = q.get(c);
                    acc_re[c] += qv.re * qv.re + qv.im * qv.im;
                    z.set(c, cx<double>{qv.re / d, qv.im / d});
                    const cx<double> zv = z.get(c);
                    acc_im[c] += qv.re * zv.re + qv.im * zv.im;
                }
                q.store(Q + o);
                q.store(R + o);
                z.store(Z + o);
                PK zero;
#pragma unroll
                for (int c = 0; c < CPL; ++c) zero.set(c, cx<double>{0.0, 0.0});
                zero.store(Y + o);
                break;
            }
            case PS_PINIT: {
                PK z;
                z.load(Z + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (!act[c]) z.set(c, cx<double>{0.0, 0.0});
                z.store(P + o);
                break;
            }
            case PS_UPDATE: {
                PK p, sp, y, r;
                p.load(P + o);
                sp.load(SP + o);
                y.load(Y + o);
                r.load(R + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> pv = p.get(c), sv = sp.get(c);
                    cx<double> yv = y.get(c), rv = r.get(c);
                    yv.re += pv.re * cf[c];
                    yv.im += pv.im * cf[c];
                    rv.re -= sv.re * cf[c];
                    rv.im -= sv.im * cf[c];
                    y.set(c, yv);
                    r.set(c, rv);
                    const cx<double> rs = r.get(c);
                    acc_re[c] += rs.re * rs.re + rs.im * rs.im;
                }
                y.store(Y + o);
                r.store(R + o);
                break;
            }
            case PS_UPD_PRE: {
                PK sp, r;
                sp.load(SP + o);
                r.load(R + o);
                const double d = a.diag[row];
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> sv = sp.get(c);
                    cx<double> rv = r.get(c);
                    rv.re -= sv.re * cf[c];
                    rv.im -= sv.im * cf[c];
                    r.set(c, rv);
                    const cx<double> rs = r.get(c);
                    const double rr = rs.re * rs.re + rs.im * rs.im;
                    acc_re[c] += rr;
                    acc_im[c] += rr / d;
                }
                r.store(R + o);
                break;
            }
            case PS_PUPD_PRE: {
                PK r, p, y;
                r.load(R + o);
                p.load(P + o);
                y.load(Y + o);
                const double d = a.diag[row];
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> pv = p.get(c);
                    cx<double> yv = y.get(c);
                    yv.re += pv.re * cf2[c];
                    yv.im += pv.im * cf2[c];
                    y.set(c, yv);
                }
                y.store(Y + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    cx<double> v{0.0, 0.0};
                    if (act[c]) {
                        const cx<double> rv = r.get(c);
                        v = cx<double>{rv.re / d, rv.im / d};
                        if (!rst[c]) {
                            const cx<double> pv = p.get(c);
                            v.re += pv.re * cf[c];
                            v.im += pv.im * cf[c];
                        }
                    }
                    p.set(c, v);
                }
                p.store(P + o);
                break;
            }
            case PS_RESTART: {
                bool any = false;
#pragma unroll
                for (int c = 0; c < CPL; ++c) any = any || rst[c];
                if (!any) break;
                PK t, r;
                t.load(T + o);
                r.load(R + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (rst[c]) r.set(c, t.get(c));
                r.store(R + o);
                break;
            }
            case PS_DIAG: {
                PK r, z;
                r.load(R + o);
                const double d = a.diag[row];
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> rv = r.get(c);
                    z.set(c, cx<double>{rv.re / d, rv.im / d});
                    const cx<double> zv = z.get(c);
                    acc_re[c] += rv.re * zv.re + rv.im * zv.im;
                }
                z.store(Z + o);
                break;
            }
            case PS_DOT: {
                PK r, z;
                r.load(R + o);
                z.load(Z + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> rv = r.get(c), zv = z.get(c);
                    acc_re[c] += rv.re * zv.re + rv.im * zv.im;
                }
                break;
            }
            case PS_PUPDATE: {
                PK z, p;
                z.load(Z + o);
                p.load(P + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    cx<double> v{0.0, 0.0};
                    if (act[c]) {
                        v = z.get(c);
                        if (!rst[c]) {
                            const cx<double> pv = p.get(c);
                            v.re += pv.re * cf[c];
                            v.im += pv.im * cf[c];
                        }
                    }
                    p.set(c, v);
                }
                p.store(P + o);
                break;
            }
            case PS_TRUEGAP: {
                PK t, r;
                t.load(T + o);
                r.load(R + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> tv = t.get(c), rv = r.get(c);
                    acc_re[c] += tv.re * tv.re + tv.im * tv.im;
                    const double dr = tv.re - rv.re, di = tv.im - rv.im;
                    acc_im[c] += dr * dr + di * di;
                }
                break;
            }
            case PS_RESCALE: {
                PK y, out;
                y.load(Y + o);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> yv = y.get(c);
                    const bool fin0 = __builtin_isfinite(yv.re) && __builtin_isfinite(yv.im);
                    cx<double> v{yv.re * cf[c], yv.im * cf[c]};
                    out.set(c, v);
                    const cx<double> ov = out.get(c);
                    const bool fin1 = __builtin_isfinite(ov.re) && __builtin_isfinite(ov.im);
                    bad[c] = bad[c] || !fin0 || !fin1;
                }
                out.store(OUT + row * a.ldo + col);
                break;
            }
        }
    }

    const bool reduces = a.stage == PS_ABSMAX || a.stage == PS_SCALE || a.stage == PS_UPDATE ||
                         a.stage == PS_DIAG || a.stage == PS_DOT || a.stage == PS_TRUEGAP || a.stage == PS_UPD_PRE ||
                         a.stage == PS_SCALE_DIAG;
    if (a.stage == PS_RESCALE) {
        // nonfinite columns: flag per column (vector atomics on global memory)
#pragma unroll
        for (int c = 0; c < CPL; ++c)
            if (colok && bad[c]) atomicAnd(&a.cs[col + c].fin, 0);
        return;
    }
    if (!reduces) return;
    const bool is_max = a.stage == PS_ABSMAX;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        for (int off = lpr; off < 64; off <<= 1) {
            const double ore = __shfl_xor(acc_re[c], off), oim = __shfl_xor(acc_im[c], off);
            acc_re[c] = is_max ? fmax(acc_re[c], ore) : acc_re[c] + ore;
            acc_im[c] += oim;
        }
    }
    __shared__ cx<double> red[4][64 * CPL];
    if (lane < lpr) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) red[wave][lane * CPL + c] = cx<double>{acc_re[c], acc_im[c]};
    }
    __syncthreads();
    if (wave == 0 && lane < lpr && colok) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            if (col + c >= a.w) continue;
            cx<double> s = red[0][lane * CPL + c];
            for (int q = 1; q < 4; ++q) {
                const cx<double> v = red[q][lane * CPL + c];
                s.re = is_max ? fmax(s.re, v.re) : s.re + v.re;
                s.im += v.im;
            }
            reinterpret_cast<cx<double> *>(a.part)[int64_t(blockIdx.x) * a.w + col + c] = s;
        }
    }
}

// Per-column logic stages (one workgroup per column reduces that column's
// partials in a fixed tree order, then thread 0 applies the PCG rules of
// the reference's Metric::solve, cpp/native.cpp:313-452).
enum : int32_t {
    PC_SAFE = 0,    // safe = absmax (0 -> 1)
    PC_INIT = 1,    // qn, thr, active ; (rz from a later PC_RZ0)
    PC_RZ0 = 2,     // rz = Re<r,z>
    PC_ALPHA = 3,   // curvature -> breakdown / alpha
    PC_RN = 4,      // rn, iterations, candidate, fixed-policy activity
    PC_TRUE = 5,    // tn (+gap) -> restart / activity
    PC_BETA = 6,    // beta, rz, activity count
    PC_FINAL = 7,   // final true residual -> report fields
    PC_CLEAR = 8,   // zero flag block + restart marks
    PC_RN_BETA = 9, // fused iteration: rn, iterations, activity (re = |r|^2), then beta, rz (im = Re<r,z>)
    PC_INIT_RZ0 = 10, // PC_INIT from re, then PC_RZ0 from im (after PS_SCALE_DIAG)
};

struct ColArgs {
    const double *part;  // [groups][w] complex
    int64_t groups;
    ColState *cs;
    PcgFlags *fl;
    double rtol, atol, gap_ratio;
    double imag_tol;      // breakdown: |Im <p,Sp>| > imag_tol |Re <p,Sp>|
    int32_t w, stage, adaptive, refresh;
    int32_t slot;         // iteration parity k & 1 (flag slots of PcgFlags)
    const int32_t *gate;  // if set and *gate == 0 the launch does nothing
    // if set: alpha (re) and beta (im) of each iteration, [hist_len][w]
    // complex, entry (k, j) for iteration k of column j (the CG coefficients
    // that give the Lanczos tridiagonal of the preconditioned operator)
    double *hist;
    int32_t hist_len;
};

static __global__ __launch_bounds__(256) void pcg_col_kernel(const ColArgs a) {
    const int j = blockIdx.x;
    if (j >= a.w) return;
    if (a.stage == PC_ALPHA && j == 0 && threadIdx.x == 0) {
        a.fl->any_active[a.slot ^ 1] = 0;
        a.fl->updated[a.slot] = 0;
    }
    if (a.gate && *a.gate == 0) return;
    if (a.stage != PC_SAFE && a.stage != PC_FINAL && a.stage != PC_INIT && a.stage != PC_RZ0 &&
        a.stage != PC_INIT_RZ0 && a.fl->broke)
        return;
    __shared__ double sre[256], sim[256];
    const bool is_max = a.stage == PC_SAFE;
    double re = 0.0, im = 0.0;
    if (a.part) {
        // groups t, t + 256, ... in ascending order; eight loads in flight
        const cx<double> *p = reinterpret_cast<const cx<double> *>(a.part);
        int64_t g = threadIdx.x;
        for (; g + 7 * 256 < a.groups; g += 8 * 256) {
            cx<double> v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = p[(g + u * 256) * a.w + j];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                re = is_max ? fmax(re, v[u].re) : re + v[u].re;
                im += v[u].im;
            }
        }
        for (; g < a.groups; g += 256) {
            const cx<double> v = p[g * a.w + j];
            re = is_max ? fmax(re, v.re) : re + v.re;
            im += v.im;
        }
    }
    sre[threadIdx.x] = re;
    sim[threadIdx.x] = im;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sre[threadIdx.x] = is_max ? fmax(sre[threadIdx.x], sre[threadIdx.x + off])
                                      : sre[threadIdx.x] + sre[threadIdx.x + off];
            sim[threadIdx.x] += sim[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double tre = sre[0], tim = sim[0];
    ColState &s = a.cs[j];
    PcgFlags *f = a.fl;
    switch (a.stage) {
        case PC_SAFE:
            s.safe = tre == 0.0 ? 1.0 : tre;
            s.iters = 0;
            s.breakdown = 0;
            s.restart = 0;
            s.fin = 1;
            break;
        case PC_INIT: {
            s.qn = sqrt(tre);
            s.thr = fmax(a.rtol * s.qn, a.atol / s.safe);
            s.active = s.qn > s.thr;
            if (s.active) atomicAdd(&f->any_active[0], 1);
            break;
        }
        case PC_RZ0:
            s.rz = tre;
            break;
        case PC_INIT_RZ0:
            s.qn = sqrt(tre);
            s.thr = fmax(a.rtol * s.qn, a.atol / s.safe);
            s.active = s.qn > s.thr;
            if (s.active) atomicAdd(&f->any_active[0], 1);
            s.rz = tim;
            break;
        case PC_ALPHA: {
            // curvature = <p, S p>: re = tre, im = tim
            if (s.active && (!(tre > 0) || !__builtin_isfinite(tre) || !__builtin_isfinite(tim) ||
                             fabs(tim) > a.imag_tol * fmax(fabs(tre), 2.2250738585072014e-308))) {
                s.breakdown = 1;
                atomicOr(&f->broke, 1);
            }
            s.alpha = s.active ? s.rz / tre : 0.0;
            if (a.hist && s.active && s.iters < a.hist_len) a.hist[(s.iters * a.w + j) * 2] = s.alpha;
            break;
        }
        case PC_RN: {
            s.rn = sqrt(tre);
            s.iters += s.active;
            s.restart = 0;
            if (s.active && s.rn <= s.thr) atomicOr(&f->need_refresh, 1);
            if (!a.adaptive && !a.refresh) s.active = s.active && s.rn > s.thr;
            break;
        }
        case PC_TRUE: {
            if (j == 0 && f->any_active[a.slot] > 0) {
                atomicAdd(reinterpret_cast<unsigned long long *>(&f->refreshes), 1ull);
                atomicAdd(reinterpret_cast<unsigned long long *>(&f->checks), 1ull);
            }
            s.tn = sqrt(tre);
            s.gap = sqrt(tim);
            if (s.active) {
                if (a.adaptive) {
                    s.restart = s.tn > s.thr && (s.gap > a.gap_ratio * fmax(s.tn, s.thr) || s.rn <= s.thr);
                } else {
                    s.restart = 1;
                }
                s.active = s.tn > s.thr;
                if (s.restart) atomicAdd(reinterpret_cast<unsigned long long *>(&f->restarts), 1ull);
            }
            break;
        }
        case PC_BETA: {
            s.beta = s.active ? tre / (s.rz != 0.0 ? s.rz : 1.0) : 0.0;
            if (a.hist && s.active && s.iters >= 1 && s.iters <= a.hist_len)
                a.hist[((s.iters - 1) * a.w + j) * 2 + 1] = s.beta;
            s.rz = tre;
            if (s.active) atomicAdd(&f->any_active[a.slot ^ 1], 1);
            break;
        }
        case PC_RN_BETA: {
            // PC_RN (fixed policy, no refresh) followed by PC_BETA; runs on
            // columns that are inactive too, where it only rewrites values
            // that inactive columns never read again.
            s.rn = sqrt(tre);
            s.iters += s.active;
            s.restart = 0;
            if (s.active) atomicAdd(&f->updated[a.slot], 1);
            s.active = s.active && s.rn > s.thr;
            s.beta = s.active ? tim / (s.rz != 0.0 ? s.rz : 1.0) : 0.0;
            if (a.hist && s.active && s.iters >= 1 && s.iters <= a.hist_len)
                a.hist[((s.iters - 1) * a.w + j) * 2 + 1] = s.beta;
            s.rz = tim;
            if (s.active) atomicAdd(&f->any_active[a.slot ^ 1], 1);
            break;
        }
        case PC_FINAL: {
            s.tn = sqrt(tre);
            break;
        }
        case PC_CLEAR:
            break;
    }
}

// Z = M^-1 R for uniform 4x4 block-Jacobi factors in FP32 (complex64
// blocks): one thread per (block row, column).
static __global__ __launch_bounds__(256) void block_apply32_kernel(const void *Rv, void *Zv, const float *fac32,
                                                                   int64_t nq, int32_t w) {
    const cx<float> *R = static_cast<const cx<float> *>(Rv);
    cx<float> *Z = static_cast<cx<float> *>(Zv);
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < nq * w; e += int64_t(gridDim.x) * 256) {
        const int64_t q = e / w, j = e - q * w;
        float re[4], im[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const cx<float> v = R[(q * 4 + r) * w + j];
            re[r] = v.re;
            im[r] = v.im;
        }
        blk4_solve_f32(fac32 + q * 16, re, im);
#pragma unroll
        for (int r = 0; r < 4; ++r) Z[(q * 4 + r) * w + j] = cx<float>{re[r], im[r]};
    }
}

// Gershgorin interval of a CSR/ELL matrix with a split diagonal: per
// workgroup, min over rows of Re(S_ii) - r_i and max of Re(S_ii) + r_i with
// r_i = sum over j != i of |S_ij| (ELL padding slots hold zero); out[2g],
// out[2g + 1].
template <class R, bool MC>
__global__ __launch_bounds__(256) void gershgorin_kernel(const void *offv, const void *diagv, const int32_t *ptr,
                                                         int32_t ell, int64_t n, double *out) {
    using MT = typename MatVal<R, MC>::type;
    const MT *off = static_cast<const MT *>(offv);
    const MT *dg = static_cast<const MT *>(diagv);
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x; r < n; r += int64_t(gridDim.x) * 256) {
        const int64_t k0 = ell ? r * ell : ptr[r], k1 = ell ? k0 + ell : ptr[r + 1];
        double rad = 0.0, cen = 0.0;
        for (int64_t k = k0; k < k1; ++k) {
            if constexpr (MC)
                rad += sqrt(double(off[k].re) * double(off[k].re) + double(off[k].im) * double(off[k].im));
            else
                rad += fabs(double(off[k]));
        }
        if (dg) {
            if constexpr (MC)
                cen = double(dg[r].re);
            else
                cen = double(dg[r]);
        }
        lo = fmin(lo, cen - rad);
        hi = fmax(hi, cen + rad);
    }
    __shared__ double slo[256], shi[256];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int off2 = 128; off2 > 0; off2 >>= 1) {
        if (int(threadIdx.x) < off2) {
            slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + off2]);
            shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + off2]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = slo[0];
        out[2 * blockIdx.x + 1] = shi[0];
    }
}

// out[j] = sum of the real parts of partials [groups][w] (complex) in column
// j: one workgroup per column, each lane adding groups t, t + 256, ... in
// ascending order, then a fixed LDS tree (the reduction of pcg_col_kernel),
// so the sums do not depend on timing.
// out[j] = sum over the partial rows of column j; with cs, also the column's
// refinement scale cs[j].safe = sqrt(out[j]) / sqrt_n (1 where out[j] is zero
// or not finite).
static __global__ __launch_bounds__(256) void colsum_kernel(const double *part, int64_t groups, int32_t w,
                                                            double *out, ColState *cs, double sqrt_n) {
    __shared__ double red[256];
    const int64_t j = blockIdx.x;
    double acc = 0.0;
    int64_t q = threadIdx.x;
    for (; q + 7 * 256 < groups; q += 8 * 256) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[((q + u * 256) * w + j) * 2];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; q < groups; q += 256) acc += part[(q * w + j) * 2];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (int(threadIdx.x) < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[j] = red[0];
        if (cs) cs[j].safe = red[0] > 0 ? sqrt(red[0]) / sqrt_n : 1.0;
    }
}

// y128 += scale_j * d64 (refinement update), and r64 = r128 / scale_j.
// With D (mode 0 only): also D = dscale * O from the stored O, as
// scale_kernel computes it (the first term of a Chebyshev series).
template <class VA, class VB>
__global__ __launch_bounds__(256) void scaled_convert_kernel(const void *Xv, int64_t ldx, void *Ov, int64_t ldo,
                                                             int64_t n, int32_t w, const ColState *cs, int32_t mode,
                                                             void *Dv = nullptr, double dscale = 0.0) {
    const cx<VA> *X = static_cast<const cx<VA> *>(Xv);
    cx<VB> *O = static_cast<cx<VB> *>(Ov);
    cx<VB> *D = static_cast<cx<VB> *>(Dv);
    const int64_t total = n * w;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t r = e / w, c = e - r * w;
        const cx<double> x = to<double>(X[r * ldx + c]);
        if (mode == 0) {  // O = X / safe
            const double inv = 1.0 / cs[c].safe;
            const cx<VB> o = to<VB>(cx<double>{x.re * inv, x.im * inv});
            O[r * ldo + c] = o;
            if (D) {
                const cx<double> y = to<double>(o);
                D[r * ldo + c] = to<VB>(cx<double>{dscale * y.re - 0.0 * y.im, dscale * y.im + 0.0 * y.re});
            }
        } else if (mode == 2) {  // O = X * safe
            O[r * ldo + c] = to<VB>(cx<double>{x.re * cs[c].safe, x.im * cs[c].safe});
        } else {  // O += X * safe
            cx<double> o = to<double>(O[r * ldo + c]);
            o.re += x.re * cs[c].safe;
            o.im += x.im * cs[c].safe;
            O[r * ldo + c] = to<VB>(o);
        }
    }
}

// z = L^-H L^-1 r for one 4x4 block (L lower triangular, complex, row-major),
// the arithmetic of block_precond_kernel for m = 4.
__device__ __forceinline__ void blk4_solve(const cx<double> *L, cx<double> (&v)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        cx<double> s = v[i];
#pragma unroll
        for (int k = 0; k < i; ++k) {
            const cx<double> l = L[i * 4 + k];
            s.re -= l.re * v[k].re - l.im * v[k].im;
            s.im -= l.re * v[k].im + l.im * v[k].re;
        }
        const cx<double> d = L[i * 4 + i];
        const double den = d.re * d.re + d.im * d.im;
        v[i] = {(s.re * d.re + s.im * d.im) / den, (s.im * d.re - s.re * d.im) / den};
    }
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        cx<double> s = v[i];
#pragma unroll
        for (int k = i + 1; k < 4; ++k) {
            const cx<double> l = L[k * 4 + i];
            s.re -= l.re * v[k].re + l.im * v[k].im;
            s.im -= l.re * v[k].im - l.im * v[k].re;
        }
        const cx<double> d = L[i * 4 + i];
        const double den = d.re * d.re + d.im * d.im;
        v[i] = {(s.re * d.re - s.im * d.im) / den, (s.im * d.re + s.re * d.im) / den};
    }
}

// Fused iteration stages PS_UPD_PRE / PS_PUPD_PRE for a uniform 4x4
// block-Jacobi preconditioner: each lane owns CPL columns of the four rows of
// one block, so z = L^-H L^-1 r is formed in registers.  Partials as in
// pcg_row_kernel ([gridDim.x][w], re = |R|^2, im = Re<R,z>).
template <class VT, int CPL>
__global__ __launch_bounds__(256) void pcg_quad_kernel(const PcgArgs a) {
    using V = cx<VT>;
    using PK = Pack<VT, CPL>;
    const int32_t live = a.stage == PS_PUPD_PRE ? a.fl->updated[a.slot] : a.fl->any_active[a.slot];
    if (a.require_active && (a.fl->broke || !live)) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lpr = 1 << a.lpr_log2;
    const int qpw = 64 >> a.lpr_log2, qpb = 4 * qpw;
    const int qloc = wave * qpw + (lane >> a.lpr_log2);
    const int cgrp = lane & (lpr - 1);
    const int col = (blockIdx.y * lpr + cgrp) * CPL;
    const bool colok = col < a.w;
    const int64_t w = a.w;
    const bool upd = a.stage == PS_UPD_PRE, pre = a.stage == PS_PRE_DOT;
    if (a.gate && *a.gate == 0) return;
    double acc_re[CPL], acc_im[CPL], cf[CPL], al[CPL];
    int32_t act[CPL], rst[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        acc_re[c] = acc_im[c] = cf[c] = al[c] = 0.0;
        act[c] = rst[c] = 0;
        if (colok) {
            const ColState &s = a.cs[col + c];
            act[c] = s.active;
            rst[c] = s.restart;
            cf[c] = upd ? s.alpha : s.beta;
            al[c] = s.alpha;
        }
    }
    const int64_t nq = a.n >> 2;
    const int64_t nblk = (nq + qpb - 1) / qpb;
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
    const int64_t span = (nblk + 7) / 8;
    const int64_t b0 = int64_t(xcd) * span;
    const int64_t b1 = b0 + span < nblk ? b0 + span : nblk;
    V *Y = static_cast<V *>(a.Y);
    V *R = static_cast<V *>(a.R);
    V *P = static_cast<V *>(a.P);
    V *Z = static_cast<V *>(a.Z);
    const V *SP = static_cast<const V *>(a.SP);
    const cx<double> *F = reinterpret_cast<const cx<double> *>(a.fac);

    for (int64_t blk = b0 + slot; blk < b1; blk += per) {
        const int64_t q = blk * qpb + qloc;
        if (q >= nq || !colok) continue;
        const int64_t o0 = q * 4 * w + col;
        PK r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i].load(R + o0 + i * w);
        if (upd) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                PK sp;
                sp.load(SP + o0 + i * w);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> sv = sp.get(c);
                    cx<double> rv = r[i].get(c);
                    rv.re -= sv.re * cf[c];
                    rv.im -= sv.im * cf[c];
                    r[i].set(c, rv);
                }
                r[i].store(R + o0 + i * w);
            }
        }
        const cx<double> *L = F + q * 16;
        const bool r32 = std::is_same<VT, float>::value && a.fac32 != nullptr;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            cx<double> v[4], rv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = rv[i] = r[i].get(c);
            if (r32)
                blk4_solve_r32(a.fac32 + q * 16, v);
            else
                blk4_solve(L, v);
            if (upd) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    acc_re[c] += rv[i].re * rv[i].re + rv[i].im * rv[i].im;
                    acc_im[c] += rv[i].re * v[i].re + rv[i].im * v[i].im;
                }
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) r[i].set(c, v[i]);  // r now holds z
            }
        }
        if (pre) {
            // Z is stored rounded, and <R, Z> is taken of the stored values
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                r[i].store(Z + o0 + i * w);
                PK rr;
                rr.load(R + o0 + i * w);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> rv = rr.get(c), zv = r[i].get(c);
                    acc_re[c] += rv.re * zv.re + rv.im * zv.im;
                }
            }
            continue;
        }
        if (!upd) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                PK p, y;
                p.load(P + o0 + i * w);
                y.load(Y + o0 + i * w);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const cx<double> pv = p.get(c);
                    cx<double> yv = y.get(c);
                    yv.re += pv.re * al[c];
                    yv.im += pv.im * al[c];
                    y.set(c, yv);
                }
                y.store(Y + o0 + i * w);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    cx<double> v{0.0, 0.0};
                    if (act[c]) {
                        v = r[i].get(c);
                        if (!rst[c]) {
                            const cx<double> pv = p.get(c);
                            v.re += pv.re * cf[c];
                            v.im += pv.im * cf[c];
                        }
                    }
                    p.set(c, v);
                }
                p.store(P + o0 + i * w);
            }
        }
    }
    if (!upd && !pre) return;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        for (int off = lpr; off < 64; off <<= 1) {
            acc_re[c] += __shfl_xor(acc_re[c], off);
            acc_im[c] += __shfl_xor(acc_im[c], off);
        }
    }
    __shared__ cx<double> red[4][64 * CPL];
    if (lane < lpr) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) red[wave][lane * CPL + c] = cx<double>{acc_re[c], acc_im[c]};
    }
    __syncthreads();
    if (wave == 0 && lane < lpr && colok) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            if (col + c >= a.w) continue;
            cx<double> s = red[0][lane * CPL + c];
            for (int qq = 1; qq < 4; ++qq) {
                s.re += red[qq][lane * CPL + c].re;
                s.im += red[qq][lane * CPL + c].im;
            }
            reinterpret_cast<cx<double> *>(a.part)[int64_t(blockIdx.x) * a.w + col + c] = s;
        }
    }
}

// Atom-block Jacobi: z_blk = L^-H L^-1 r_blk for contiguous blocks, one
// thread per (block, column).  Factors are complex128 row-major.
template <class VT, int MAXB>
__global__ __launch_bounds__(256) void block_precond_kernel(const void *Rv, void *Zv, int64_t w, const int32_t *bptr,
                                                            const int64_t *foff, const double *fac, int64_t nblocks,
                                                            const PcgFlags *fl, const int32_t *rows) {
    if (fl && fl->broke) return;
    const cx<VT> *R = static_cast<const cx<VT> *>(Rv);
    cx<VT> *Z = static_cast<cx<VT> *>(Zv);
    const cx<double> *L = reinterpret_cast<const cx<double> *>(fac);
    const int64_t total = nblocks * w;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t q = e / w, j = e - q * w;
        const int32_t a0 = bptr[q], m = bptr[q + 1] - a0;
        if (m > MAXB) continue;  // large blocks: one workgroup each (large_block_kernel)
        const int32_t *rq = rows ? rows + a0 : nullptr;
        const auto row = [&](int i) -> int64_t { return rq ? int64_t(rq[i]) : int64_t(a0 + i); };
        const cx<double> *Lq = L + foff[q];
        cx<double> v[MAXB];
        for (int i = 0; i < m; ++i) {
            cx<double> s = to<double>(R[row(i) * w + j]);
            for (int k = 0; k < i; ++k) {
                const cx<double> l = Lq[i * m + k];
                s.re -= l.re * v[k].re - l.im * v[k].im;
                s.im -= l.re * v[k].im + l.im * v[k].re;
            }
            const cx<double> d = Lq[i * m + i];
            const double den = d.re * d.re + d.im * d.im;
            v[i] = {(s.re * d.re + s.im * d.im) / den, (s.im * d.re - s.re * d.im) / den};
        }
        for (int i = m - 1; i >= 0; --i) {
            cx<double> s = v[i];
            for (int k = i + 1; k < m; ++k) {
                const cx<double> l = Lq[k * m + i];  // conj(L[k][i])
                s.re -= l.re * v[k].re + l.im * v[k].im;
                s.im -= l.re * v[k].im - l.im * v[k].re;
            }
            const cx<double> d = Lq[i * m + i];  // divide by conj(d)
            const double den = d.re * d.re + d.im * d.im;
            v[i] = {(s.re * d.re - s.im * d.im) / den, (s.im * d.re + s.re * d.im) / den};
        }
        for (int i = 0; i < m; ++i) Z[row(i) * w + j] = to<VT>(v[i]);
    }
}

}  // namespace kmsd
