// Batched PCG metric solve on the device (reference Metric::solve,
// cpp/native.cpp:313-452): per-column RHS scaling by the largest component,
// threshold max(rtol |q|, atol/safe), curvature breakdown test, true-residual
// refresh every `check_interval` iterations (fixed) or on candidate
// convergence (adaptive), final original-equation acceptance, rescale with
// representability checks.  Column state stays on the device; the host polls
// a flag block every `pcg_poll` iterations.
//
// complex64 right-hand sides run FP32 PCG inner solves inside an FP64
// iterative refinement: the residual b - S y and the solution accumulator y
// are complex128, and acceptance is decided on that FP64 residual.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kmsd_pcg.hpp"
#include "kmsd_pcg_kernels.hpp"

using namespace kmsd;

struct kmsd_metric::Work {
    DBuf q, y, r, z, p, sp, t, cs, fl, part, part2;
    DBuf b128, y128, r128, r64, d64, cs2, sums, rsums, hist;
};

namespace kmsd {

namespace {

struct RowShape {
    int cpl, lpr_log2, gy;
    int64_t gx;
};

int ilog2c(int64_t v) {
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

RowShape row_shape(kmsd_ctx *c, int64_t n, int32_t w, int dtype, int cap) {
    RowShape s;
    s.cpl = std::min(cap, dtype == KMSD_C64 ? 2 : 1);
    while (s.cpl > 1 && w % s.cpl) s.cpl >>= 1;
    const int64_t lanes = (w + s.cpl - 1) / s.cpl;
    s.lpr_log2 = std::min(6, ilog2c(lanes));
    const int64_t lpr = int64_t(1) << s.lpr_log2;
    s.gy = int((lanes + lpr - 1) / lpr);
    const int64_t rpb = 4 * (64 / lpr);
    const int64_t nblk = (n + rpb - 1) / rpb;
    // at least pcg_min_blocks row blocks per workgroup, but no fewer than
    // 512 workgroups (two per CU) where there are that many row blocks
    const int64_t floor_g = std::min<int64_t>(512, groups_for_blocks(c, nblk, 1));
    s.gx = std::max<int64_t>(8, std::min<int64_t>(c->groups, std::max(floor_g, groups_for_blocks(c, nblk, c->pcg_min_blocks))));
    return s;
}

struct Solver {
    kmsd_metric *m;
    kmsd_ctx *c;
    int dtype;
    int64_t n;
    int32_t w;
    size_t eb;
    RowShape rs;
    void *Q, *Y, *R, *Z, *P, *SP, *T;
    ColState *cs;
    PcgFlags *fl;
    double *part;   // row-kernel partials [rs.gx][w]
    double *part2;  // step-kernel partials
    int64_t g2 = 0;
    int32_t slot = 0;  // flag slot of the current iteration (k & 1): gates and the liveness reads
    double *hist = nullptr;  // CG coefficient history [hist_len][w] (alpha, beta), or null
    int32_t hist_len = 0;

    void row(int stage, const void *B = nullptr, int64_t ldb = 0, void *OUT = nullptr, int64_t ldo = 0,
             bool require_active = false, const int32_t *gate = nullptr) {
        PcgArgs a{};
        a.B = B;
        a.Q = Q;
        a.Y = Y;
        a.R = R;
        a.Z = Z;
        a.P = P;
        a.SP = SP;
        a.T = T;
        a.OUT = OUT;
        a.diag = m->diag;
        a.cs = cs;
        a.fl = fl;
        a.part = part;
        a.n = n;
        a.ldb = ldb;
        a.ldo = ldo;
        a.w = w;
        a.lpr_log2 = rs.lpr_log2;
        a.stage = stage;
        a.require_active = require_active;
        a.slot = slot;
        a.gate = gate;
        const dim3 grid(unsigned(rs.gx), unsigned(rs.gy));
        if (dtype == KMSD_C128)
            hipLaunchKernelGGL((pcg_row_kernel<double, 1>), grid, dim3(256), 0, c->stream, a);
        else if (rs.cpl == 2)
            hipLaunchKernelGGL((pcg_row_kernel<float, 2>), grid, dim3(256), 0, c->stream, a);
        else
            hipLaunchKernelGGL((pcg_row_kernel<float, 1>), grid, dim3(256), 0, c->stream, a);
        HIPCHECK(hipGetLastError());
    }

    // Fused iteration stages: diagonal preconditioner in the row kernel,
    // uniform 4x4 blocks in the quad kernel.
    void fused(int stage) {
        if (m->nblocks == 0) {
            row(stage, nullptr, 0, nullptr, 0, true);
            return;
        }
        PcgArgs a{};
        a.Y = Y;
        a.R = R;
        a.P = P;
        a.SP = SP;
        a.fac = m->factors;
        a.fac32 = dtype == KMSD_C64 ? m->factors32 : nullptr;
        a.cs = cs;
        a.fl = fl;
        a.part = part;
        a.n = n;
        a.w = w;
        a.lpr_log2 = rs.lpr_log2;
        a.stage = stage;
        a.require_active = 1;
        a.slot = slot;
        const dim3 grid(unsigned(rs.gx), unsigned(rs.gy));
        if (dtype == KMSD_C128)
            hipLaunchKernelGGL((pcg_quad_kernel<double, 1>), grid, dim3(256), 0, c->stream, a);
        else if (rs.cpl == 2)
            hipLaunchKernelGGL((pcg_quad_kernel<float, 2>), grid, dim3(256), 0, c->stream, a);
        else
            hipLaunchKernelGGL((pcg_quad_kernel<float, 1>), grid, dim3(256), 0, c->stream, a);
        HIPCHECK(hipGetLastError());
    }

    void col(int stage, const double *p, int64_t groups, double rtol, double atol, bool refresh = false,
             const int32_t *gate = nullptr) {
        ColArgs a{};
        a.part = p;
        a.groups = groups;
        a.cs = cs;
        a.fl = fl;
        a.rtol = rtol;
        a.atol = atol;
        a.gap_ratio = m->gap_ratio;
        // The reference's 1e-10 curvature test is an FP64 criterion; complex64
        // inner solves round every stored vector at ~6e-8, so their test
        // tolerates imaginary parts up to 1e-5 of the real part.
        a.imag_tol = dtype == KMSD_C128 ? 1e-10 : 1e-5;
        a.w = w;
        a.stage = stage;
        a.adaptive = m->adaptive;
        a.refresh = refresh;
        a.slot = slot;
        a.gate = gate;
        a.hist = hist;
        a.hist_len = hist_len;
        hipLaunchKernelGGL(pcg_col_kernel, dim3(unsigned(w)), dim3(256), 0, c->stream, a);
        HIPCHECK(hipGetLastError());
    }

    // OUT = alpha * S X + gamma * Pt with partial <W, OUT>
    void spmv(const void *X, void *OUT, const void *Pt, double alpha, double gamma, const void *Wp,
              const int32_t *gate) {
        Step s;
        s.A = m->S;
        s.X = X;
        s.ldx = w;
        s.OUT = OUT;
        s.ldo = w;
        s.alpha = Cplx{alpha, 0};
        if (Pt) {
            s.P = Pt;
            s.ldp = w;
            s.gamma = gamma;
            s.use_p = true;
        }
        if (Wp) {
            s.W = Wp;
            s.ldw = w;
            s.part = part2;
        }
        s.n = n;
        s.w = w;
        s.gate = gate;
        // small systems: no more workgroups than the vector passes use, so
        // the curvature partials stay as few as the passes' partials
        s.max_groups = rs.gx;
        // complex64 solves are the inner solves of an FP64 refinement (or a
        // caller's FP32 request): FP32 matrix values halve the S stream.
        s.f32_matrix = dtype == KMSD_C64;
        g2 = launch_step(c, dtype, s);
    }

    // z = M^-1 r with the reference's precedence (cpp/native.cpp:237-297):
    // IC(0), else the diagonal when there are no blocks, else the blocks
    // (small ones in registers, large ones one workgroup each).
    void precondition(bool require_active, const int32_t *gate = nullptr) {
        if (m->ic) {
            ic_apply(m, dtype, R, Z, w, fl);
            row(PS_DOT, nullptr, 0, nullptr, 0, require_active, gate);
            return;
        }
        if (m->nblocks == 0) {
            row(PS_DIAG, nullptr, 0, nullptr, 0, require_active, gate);
            return;
        }
        if (m->uniform4 && m->factors32 && dtype == KMSD_C64 && m->nlarge == 0) {
            // z and <r, z> in one pass of the 4x4-block kernel (FP32 factors)
            PcgArgs a{};
            a.R = R;
            a.Z = Z;
            a.fac = m->factors;
            a.fac32 = m->factors32;
            a.cs = cs;
            a.fl = fl;
            a.part = part;
            a.n = n;
            a.w = w;
            a.lpr_log2 = rs.lpr_log2;
            a.stage = PS_PRE_DOT;
            a.require_active = require_active ? 1 : 0;
            a.slot = slot;
            a.gate = gate;
            const dim3 grid(unsigned(rs.gx), unsigned(rs.gy));
            if (rs.cpl == 2)
                hipLaunchKernelGGL((pcg_quad_kernel<float, 2>), grid, dim3(256), 0, c->stream, a);
            else
                hipLaunchKernelGGL((pcg_quad_kernel<float, 1>), grid, dim3(256), 0, c->stream, a);
            HIPCHECK(hipGetLastError());
            return;
        }
        if (m->nlarge < m->nblocks) {
            const int64_t total = m->nblocks * w;
            const int64_t g = std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 16384));
            const int32_t *rows = m->block_rows;
            if (dtype == KMSD_C128) {
                if (m->max_block <= 4)
                    hipLaunchKernelGGL((block_precond_kernel<double, 4>), dim3(unsigned(g)), dim3(256), 0, c->stream,
                                       R, Z, int64_t(w), m->block_ptr, m->factor_off, m->factors, m->nblocks, fl, rows);
                else
                    hipLaunchKernelGGL((block_precond_kernel<double, SMALL_BLOCK>), dim3(unsigned(g)), dim3(256), 0,
                                       c->stream, R, Z, int64_t(w), m->block_ptr, m->factor_off, m->factors,
                                       m->nblocks, fl, rows);
            } else {
                if (m->max_block <= 4)
                    hipLaunchKernelGGL((block_precond_kernel<float, 4>), dim3(unsigned(g)), dim3(256), 0, c->stream,
                                       R, Z, int64_t(w), m->block_ptr, m->factor_off, m->factors, m->nblocks, fl, rows);
                else
                    hipLaunchKernelGGL((block_precond_kernel<float, SMALL_BLOCK>), dim3(unsigned(g)), dim3(256), 0,
                                       c->stream, R, Z, int64_t(w), m->block_ptr, m->factor_off, m->factors,
                                       m->nblocks, fl, rows);
            }
            HIPCHECK(hipGetLastError());
        }
        large_blocks_apply(m, dtype, R, Z, w, fl);
        row(PS_DOT, nullptr, 0, nullptr, 0, require_active, gate);
    }

    void clear_flag(size_t offset, size_t bytes) {
        HIPCHECK(hipMemsetAsync(reinterpret_cast<char *>(fl) + offset, 0, bytes, c->stream));
    }

    PcgFlags read_flags() {
        PcgFlags h{};
        read_status(c, {{&h, fl, sizeof(h)}});
        return h;
    }
};

Solver make_solver(kmsd_metric *m, int dtype, int32_t nb, bool second_state = false) {
    kmsd_ctx *c = m->ctx;
    if (!m->work) m->work = new kmsd_metric::Work();
    kmsd_metric::Work &wk = *m->work;
    Solver s;
    s.m = m;
    s.c = c;
    s.dtype = dtype;
    s.n = m->n;
    s.w = nb;
    s.eb = elem_bytes(dtype);
    const size_t bytes = size_t(s.n) * size_t(nb) * s.eb;
    s.Q = wk.q.ensure(c, bytes);
    s.Y = wk.y.ensure(c, bytes);
    s.R = wk.r.ensure(c, bytes);
    s.Z = wk.z.ensure(c, bytes);
    s.P = wk.p.ensure(c, bytes);
    s.SP = wk.sp.ensure(c, bytes);
    s.T = wk.t.ensure(c, bytes);
    s.cs = static_cast<ColState *>((second_state ? wk.cs2 : wk.cs).ensure(c, sizeof(ColState) * size_t(nb)));
    s.fl = static_cast<PcgFlags *>(wk.fl.ensure(c, 256));
    s.rs = row_shape(c, s.n, nb, dtype, 4);
    const size_t pg = size_t(std::max<int64_t>(s.rs.gx, c->groups)) * size_t(nb) * 2 * sizeof(double);
    s.part = static_cast<double *>(wk.part.ensure(c, pg));
    s.part2 = static_cast<double *>(wk.part2.ensure(c, pg));
    return s;
}

struct Outcome {
    std::vector<ColState> cs;
    PcgFlags fl;
};

// PCG in the solver's own precision.  B (ldb) is any block of that dtype;
// the solution (times the column scale) lands in OUT (ldo).  Never throws on
// numerical failure: the caller inspects the outcome.  final_check: the
// closing true-residual acceptance (report fields tn, convergence); the
// inner solves of the FP64 refinement skip it, their acceptance is the
// refinement's FP64 residual.
Outcome pcg_run(Solver &S, const void *B, int64_t ldb, void *OUT, int64_t ldo, double rtol, double atol,
                int maxiter, bool final_check = true) {
    kmsd_metric *m = S.m;
    kmsd_ctx *c = S.c;
    const int32_t w = S.w;
    S.clear_flag(0, sizeof(PcgFlags));
    // 1. per-column scale
    S.row(PS_ABSMAX, B, ldb);
    S.col(PC_SAFE, S.part, S.rs.gx, rtol, atol);
    // 2. q = b/safe, r = q, y = 0 ; |q| -> thr, activity
    // 3. z = M^-1 r, rz (with the diagonal preconditioner in the same pass)
    if (!m->ic && m->nblocks == 0) {
        S.row(PS_SCALE_DIAG, B, ldb);
        S.col(PC_INIT_RZ0, S.part, S.rs.gx, rtol, atol);
    } else {
        S.row(PS_SCALE, B, ldb);
        S.col(PC_INIT, S.part, S.rs.gx, rtol, atol);
        S.precondition(false);
        S.col(PC_RZ0, S.part, S.rs.gx, rtol, atol);
    }
    // 4. p = active ? z : 0
    S.row(PS_PINIT);
    const int32_t *cand = &S.fl->need_refresh;
    const int poll = std::max(1, c->pcg_poll);
    // Fixed residual policy with a diagonal or uniform 4x4 preconditioner:
    // iterations without a true-residual refresh run fused (Y, R update with
    // z = M^-1 R and both reductions in one pass, P update recomputing z).
    const bool fusable = !m->adaptive && !m->ic && (m->nblocks == 0 || m->uniform4);
    for (int k = 0; k < maxiter; ++k) {
        S.slot = k & 1;
        const int32_t *any = &S.fl->any_active[S.slot];
        // (iterations of columns that are done, or of a broken solve, are
        // gated off on the device; the poll only ends the loop)
        if (k % poll == 0 && (k > 0 || poll == 1)) {
            const PcgFlags h = S.read_flags();
            if (h.broke || h.any_active[S.slot] == 0) break;
        }
        const bool refresh = (k + 1) % m->check_interval == 0;
        // sp = S p ; curvature <p, sp>
        S.spmv(S.P, S.SP, nullptr, 1.0, 0.0, S.P, any);
        S.col(PC_ALPHA, S.part2, S.g2, rtol, atol, false, any);
        if (fusable && !refresh) {
            S.fused(PS_UPD_PRE);
            S.col(PC_RN_BETA, S.part, S.rs.gx, rtol, atol);
            S.fused(PS_PUPD_PRE);
            continue;
        }
        // y += alpha p ; r -= alpha sp ; |r|
        S.row(PS_UPDATE, nullptr, 0, nullptr, 0, true);
        S.clear_flag(offsetof(PcgFlags, need_refresh), sizeof(int32_t));
        S.col(PC_RN, S.part, S.rs.gx, rtol, atol, refresh, any);
        if (refresh || m->adaptive) {
            const int32_t *gate = refresh ? any : cand;
            // t = q - S y ; |t|, |t - r| ; restart / activity
            S.spmv(S.Y, S.T, S.Q, -1.0, 1.0, nullptr, gate);
            S.row(PS_TRUEGAP, nullptr, 0, nullptr, 0, true, gate);
            S.col(PC_TRUE, S.part, S.rs.gx, rtol, atol, refresh, gate);
            S.row(PS_RESTART, nullptr, 0, nullptr, 0, true, gate);
        }
        // z = M^-1 r ; beta (counts the next iteration's active columns into
        // slot ^ 1, which the P update then reads)
        S.precondition(true);
        S.col(PC_BETA, S.part, S.rs.gx, rtol, atol);
        S.slot ^= 1;
        S.row(PS_PUPDATE, nullptr, 0, nullptr, 0, true);
        S.slot ^= 1;
    }
    S.slot = 0;
    // final acceptance on the original (scaled) equation
    if (final_check) {
        S.spmv(S.Y, S.T, S.Q, -1.0, 1.0, nullptr, nullptr);
        S.row(PS_TRUEGAP);
        S.col(PC_FINAL, S.part, S.rs.gx, rtol, atol);
    }
    S.row(PS_RESCALE, nullptr, 0, OUT, ldo);
    Outcome o;
    o.cs.resize(w);
    read_status(c, {{o.cs.data(), S.cs, sizeof(ColState) * w}, {&o.fl, S.fl, sizeof(PcgFlags)}});
    return o;
}

void store_report(kmsd_metric *m, const std::vector<ColState> &cs, const std::vector<double> *tn_override = nullptr,
                  const std::vector<double> *qn_override = nullptr, const std::vector<double> *thr_override = nullptr) {
    const size_t w = cs.size();
    m->rep_conv.assign(w, 0);
    m->rep_brk.assign(w, 0);
    m->rep_its.assign(w, 0);
    m->rep_res.assign(w, 0.0);
    m->rep_rel.assign(w, 0.0);
    for (size_t j = 0; j < w; ++j) {
        const ColState &s = cs[j];
        const double tn = tn_override ? (*tn_override)[j] : s.tn;
        const double qn = qn_override ? (*qn_override)[j] : s.qn;
        const double thr = thr_override ? (*thr_override)[j] : s.thr;
        const double scale = tn_override ? 1.0 : s.safe;
        m->rep_brk[j] = uint8_t(s.breakdown != 0);
        m->rep_conv[j] = uint8_t(tn <= thr && !s.breakdown && s.fin);
        m->rep_its[j] = s.iters;
        m->rep_rel[j] = qn > 0 ? tn / qn : 0.0;
        m->rep_res[j] = tn * scale;
        m->max_relative = std::max(m->max_relative, m->rep_rel[j]);
    }
}

void account(kmsd_metric *m, const Outcome &o, int64_t extra_checks) {
    ++m->calls;
    int64_t mx = 0;
    for (const ColState &s : o.cs) mx = std::max<int64_t>(mx, s.iters);
    m->iterations += mx;
    m->refreshes += o.fl.refreshes;
    m->residual_checks += o.fl.checks + extra_checks;
    m->restarts += o.fl.restarts;
}

// Per-column sqrt(sum over groups of the partials' real parts) in a fixed
// order (colsum_kernel, one workgroup per column), read as w doubles.
std::vector<double> column_norms(kmsd_metric *m, const double *part, int64_t groups, int32_t w) {
    kmsd_ctx *c = m->ctx;
    double *sums = static_cast<double *>(m->work->sums.ensure(c, sizeof(double) * size_t(w)));
    hipLaunchKernelGGL(colsum_kernel, dim3(unsigned(w)), dim3(256), 0, c->stream, part, groups, w, sums,
                       static_cast<ColState *>(nullptr), 1.0);
    HIPCHECK(hipGetLastError());
    std::vector<double> out(static_cast<size_t>(w));
    read_status(c, {{out.data(), sums, sizeof(double) * size_t(w)}});
    for (double &v : out) v = std::sqrt(v);
    return out;
}

// Per-column |T| from PS_TRUEGAP partials, summed in a fixed order.
std::vector<double> true_norms(Solver &S) {
    S.row(PS_TRUEGAP);
    return column_norms(S.m, S.part, S.rs.gx, S.w);
}

// Direct solve with the one full-size factor (reference Metric::solve with
// direct set, cpp/native.cpp:331-347): y = M^-1 q, then up to refine_maxiter
// corrections M^-1 (q - S y) on the columns whose true residual is above
// threshold (the others get a zero right-hand side), then the same final
// acceptance as the PCG.  complex128 only; iterations stay 0.
Outcome direct_run(Solver &S, const void *B, int64_t ldb, void *OUT, int64_t ldo, double rtol, double atol,
                   int64_t &checks) {
    kmsd_metric *m = S.m;
    kmsd_ctx *c = S.c;
    const int32_t w = S.w;
    const int64_t n = S.n;
    S.clear_flag(0, sizeof(PcgFlags));
    S.row(PS_ABSMAX, B, ldb);
    S.col(PC_SAFE, S.part, S.rs.gx, rtol, atol);
    S.row(PS_SCALE, B, ldb);  // Q = B/safe, R = Q, Y = 0
    S.col(PC_INIT, S.part, S.rs.gx, rtol, atol);
    auto copy_into_y = [&](bool add) {  // Y = Z or Y += Z
        Step s;
        s.Z = S.Z;
        s.ldz = w;
        if (add) {
            s.P = S.Y;
            s.ldp = w;
            s.gamma = 1.0;
            s.use_p = true;
        }
        s.OUT = S.Y;
        s.ldo = w;
        s.n = n;
        s.w = w;
        launch_step(c, KMSD_C128, s);
    };
    S.precondition(false);  // Z = M^-1 q
    copy_into_y(false);
    std::vector<ColState> cs(static_cast<size_t>(w));
    read_status(c, {{cs.data(), S.cs, sizeof(ColState) * w}});
    for (int step = 0; step < m->refine_maxiter; ++step) {
        S.spmv(S.Y, S.T, S.Q, -1.0, 1.0, nullptr, nullptr);  // T = q - S y
        ++checks;
        const std::vector<double> tn = true_norms(S);
        bool needed = false;
        for (int32_t j = 0; j < w; ++j) {
            cs[size_t(j)].restart = tn[size_t(j)] > cs[size_t(j)].thr;
            needed = needed || cs[size_t(j)].restart;
        }
        if (!needed) break;
        // R = T on the columns still above threshold, 0 elsewhere
        copy_h2d(c, S.cs, cs.data(), sizeof(ColState) * w);
        launch_scale(c, KMSD_C128, S.R, w, S.R, w, n, w, Cplx{0, 0});
        S.row(PS_RESTART);
        S.precondition(false);
        copy_into_y(true);
        ++m->refinement_steps;
    }
    for (ColState &x : cs) x.restart = 0;
    copy_h2d(c, S.cs, cs.data(), sizeof(ColState) * w);
    S.spmv(S.Y, S.T, S.Q, -1.0, 1.0, nullptr, nullptr);
    S.row(PS_TRUEGAP);
    S.col(PC_FINAL, S.part, S.rs.gx, rtol, atol);
    S.row(PS_RESCALE, nullptr, 0, OUT, ldo);
    Outcome o;
    o.cs.resize(w);
    read_status(c, {{o.cs.data(), S.cs, sizeof(ColState) * w}, {&o.fl, S.fl, sizeof(PcgFlags)}});
    return o;
}

void direct_solve(kmsd_metric *m, int dtype, const void *B, int64_t ldb, void *Y, int64_t ldy, int32_t nb,
                  double rtol, double atol) {
    kmsd_ctx *c = m->ctx;
    const int64_t n = m->n;
    int64_t checks = 0;
    Outcome o;
    bool fin_all = true;
    if (dtype == KMSD_C128) {
        Solver S = make_solver(m, KMSD_C128, nb);
        o = direct_run(S, B, ldb, Y, ldy, rtol, atol, checks);
    } else {
        // complex64 blocks: the direct solve runs in complex128 throughout.
        if (!m->work) m->work = new kmsd_metric::Work();
        kmsd_metric::Work &wk = *m->work;
        void *bb = wk.b128.ensure(c, size_t(n) * nb * 16);
        void *yy = wk.y128.ensure(c, size_t(n) * nb * 16);
        launch_convert(c, KMSD_C64, B, ldb, KMSD_C128, bb, nb, n, nb);
        Solver S = make_solver(m, KMSD_C128, nb);
        o = direct_run(S, bb, nb, yy, nb, rtol, atol, checks);
        launch_convert(c, KMSD_C128, yy, nb, KMSD_C64, Y, ldy, n, nb);
        fin_all = !launch_nonfinite(c, KMSD_C64, Y, ldy, n, nb);
        for (ColState &s : o.cs) s.fin = s.fin && fin_all;
    }
    store_report(m, o.cs);
    ++m->calls;
    m->refreshes += checks;
    m->residual_checks += checks + 1;
    bool ok = true, fin = true;
    for (const ColState &s : o.cs) fin = fin && s.fin;
    for (size_t j = 0; j < m->rep_conv.size(); ++j) ok = ok && m->rep_conv[j];
    if (!ok)
        throw Error(KMSD_ENUMERIC, !fin ? "S solve solution is nonfinite or not representable in complex128 after rescaling"
                                        : "S solve failed true-residual acceptance");
}

// Whether the complex64 refinement may take Chebyshev inner solves: S in
// CSR/ELL form with real or complex values, no block, IC(0) or direct
// preconditioner, and a Gershgorin interval [lo, hi] with lo > 0 and
// hi <= kChebKappa lo (then a series of about a dozen terms reduces the
// residual by 1e-6).  The interval is computed once per metric.
constexpr double kChebKappa = 4.0;

bool cheb_usable(kmsd_metric *m) {
    kmsd_ctx *c = m->ctx;
    if (!c->cheb_inner || m->ic || m->nblocks || m->direct || m->identity) return false;
    if (m->cheb_state) return m->cheb_state == 1;
    m->cheb_state = 2;
    const kmsd_matrix *S = m->S;
    if (!S || S->fmt == KMSD_BSR4 || S->vkind == VK_IMAG || !S->val64) return false;
    const int64_t n = S->n;
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 256));
    double *out = static_cast<double *>(m->work->sums.ensure(c, sizeof(double) * 2 * size_t(g)));
    const void *dg = S->has_diag ? S->dval64 : nullptr;
    if (S->vkind == VK_COMPLEX)
        hipLaunchKernelGGL((gershgorin_kernel<double, true>), dim3(unsigned(g)), dim3(256), 0, c->stream, S->val64, dg,
                           S->ptr, S->ell, n, out);
    else
        hipLaunchKernelGGL((gershgorin_kernel<double, false>), dim3(unsigned(g)), dim3(256), 0, c->stream, S->val64,
                           dg, S->ptr, S->ell, n, out);
    HIPCHECK(hipGetLastError());
    std::vector<double> h(2 * size_t(g));
    read_status(c, {{h.data(), out, h.size() * sizeof(double)}});
    double lo = h[0], hi = h[1];
    for (int64_t q = 1; q < g; ++q) {
        lo = std::min(lo, h[2 * q]);
        hi = std::max(hi, h[2 * q + 1]);
    }
    if (std::isfinite(lo) && std::isfinite(hi) && lo > 0 && hi <= kChebKappa * lo) {
        m->cheb_state = 1;
        m->cheb_lo = lo;
        m->cheb_hi = hi;
    }
    return m->cheb_state == 1;
}

// Point Chebyshev semi-iteration for S d = r on S's Gershgorin interval
// [cheb_lo, cheb_hi] (theta, delta its center and half width, sigma =
// theta / delta): d_1 = r / theta, d_(k+1) = omega_(k+1) (d_k + (r - S d_k) /
// theta) + (1 - omega_(k+1)) d_(k-1) with omega_2 = 1 / (1 - 1 / (2 sigma^2)),
// omega_(k+1) = 1 / (1 - omega_k / (4 sigma^2)); the residual after K steps is
// T_K((theta - S) / delta) / T_K(sigma) r, so K is the least count with
// 1 / T_K(sigma) <= reduction (1 where the interval is one point: d_1 is
// then the solution).
int cheb_steps(const kmsd_metric *m, double reduction) {
    const double theta = 0.5 * (m->cheb_hi + m->cheb_lo), delta = 0.5 * (m->cheb_hi - m->cheb_lo);
    if (!(delta > 1e-14 * theta)) return 1;
    return std::max(1, std::min(400, int(std::ceil(std::acosh(1.0 / reduction) / std::acosh(theta / delta)))));
}

// The first step's coefficient: d_1 = r / theta.
double cheb_first(const kmsd_metric *m) { return 1.0 / (0.5 * (m->cheb_hi + m->cheb_lo)); }

// Steps 2..K of the point semi-iteration, d_1 already stored in d when K is
// odd and in t when K is even, so that d_K lands in d; each step is one
// fused product (S d_k, the own row d_k, d_(k-1) and r in the epilogue),
// d_(k+1) replacing d_(k-1) row by row.  r is kept.  Returns the number of
// products (K - 1).
int cheb_run(kmsd_metric *m, const void *r, void *d, void *t, int32_t nb, int K) {
    kmsd_ctx *c = m->ctx;
    const double theta = 0.5 * (m->cheb_hi + m->cheb_lo), delta = 0.5 * (m->cheb_hi - m->cheb_lo);
    const double sigma = theta / delta;
    void *cur = (K % 2) ? d : t, *prev = (K % 2) ? t : d;
    Step s;
    s.A = m->S;
    s.ldx = nb;
    s.use_beta = true;
    s.W = r;
    s.ldw = nb;
    s.rhs = true;
    s.ldp = nb;
    s.ldo = nb;
    s.n = m->n;
    s.w = nb;
    s.f32_matrix = true;
    double omega = 1.0;
    for (int k = 1; k < K; ++k) {
        omega = k == 1 ? 1.0 / (1.0 - 1.0 / (2.0 * sigma * sigma)) : 1.0 / (1.0 - omega / (4.0 * sigma * sigma));
        Step u = s;
        u.X = cur;
        u.alpha = Cplx{-omega / theta, 0.0};
        u.beta = Cplx{omega, 0.0};
        u.coef = Cplx{omega / theta, 0.0};
        if (k > 1) {  // d_0 = 0
            u.P = prev;
            u.use_p = true;
            u.gamma = 1.0 - omega;
        }
        u.OUT = prev;
        launch_step(c, KMSD_C64, u);
        std::swap(prev, cur);
    }
    return K - 1;
}

// Whether the complex64 refinement may take block-Jacobi Chebyshev inner
// solves: uniform 4x4 atom blocks with real FP32 factors, a real-valued
// 4x4-block S, widths and byte extents the matrix-core kernel accepts, and an
// interval not yet found unusable.
bool bcheb_candidate(kmsd_metric *m, int32_t nb) {
    kmsd_ctx *c = m->ctx;
    const kmsd_matrix *S = m->S;
    if (!c->cheb_inner || m->ic || m->direct || m->identity || !m->uniform4 || !m->factors32 || m->nlarge)
        return false;
    if (!S || S->fmt != KMSD_BSR4 || S->vkind != VK_REAL || !S->val32 || nb % 4) return false;
    const int64_t lim = (int64_t(1) << 31) - (int64_t(1) << 16);
    if (S->n * int64_t(nb) * 8 > lim || (S->noff / 16) * 64 > lim) return false;
    return m->bcheb_state != 2;
}

// Number of eigenvalues below x of the symmetric tridiagonal matrix with
// diagonal a and squared off-diagonals b2 (Sturm count of the LDL^T pivots).
int tri_count_below(const std::vector<double> &a, const std::vector<double> &b2, double x) {
    int cnt = 0;
    double q = a[0] - x;
    if (q < 0) ++cnt;
    for (size_t i = 1; i < a.size(); ++i) {
        if (q == 0) q = 1e-300;
        q = a[i] - x - b2[i - 1] / q;
        if (q < 0) ++cnt;
    }
    return cnt;
}

// Smallest and largest eigenvalues of that matrix by bisection.
void tri_extremes(const std::vector<double> &a, const std::vector<double> &b2, double &lo, double &hi) {
    const size_t m = a.size();
    double glo = a[0], ghi = a[0];
    for (size_t i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::sqrt(b2[i - 1]) : 0.0) + (i + 1 < m ? std::sqrt(b2[i]) : 0.0);
        glo = std::min(glo, a[i] - r);
        ghi = std::max(ghi, a[i] + r);
    }
    double l = glo, u = ghi;  // smallest: count_below(l) = 0 <= ... < 1 = count_below(u)
    for (int it = 0; it < 100; ++it) {
        const double x = 0.5 * (l + u);
        (tri_count_below(a, b2, x) >= 1 ? u : l) = x;
    }
    lo = l;
    l = glo;
    u = ghi;  // largest: count_below(l) < m <= count_below(u)
    for (int it = 0; it < 100; ++it) {
        const double x = 0.5 * (l + u);
        (tri_count_below(a, b2, x) >= int(m) ? u : l) = x;
    }
    hi = u;
}

// Interval for the block-Jacobi Chebyshev semi-iteration from the Ritz values of a PCG inner
// solve: the CG coefficients (alpha_k, beta_k) of each column give the
// Lanczos tridiagonal of the preconditioned operator, T_kk = 1/alpha_k +
// beta_(k-1)/alpha_(k-1), T_k,k+1 = sqrt(beta_k)/alpha_k; the extreme Ritz
// values over the columns with at least 6 iterations; the lower end is
// bcheb_lo_pct percent of the smallest (96 by default), the upper end 2 %
// above the largest.  Ritz values lie inside the spectrum; an iteration whose
// interval misses part of it still reduces every eigencomponent in
// (0, lo + hi), only more slowly below lo, the refinement's FP64 residual
// decides acceptance, and a refinement round that reduces the residual by
// less than the square root of what the interval promised makes the metric
// return to PCG inner solves.
void bcheb_measure(kmsd_metric *m, const double *hist, int32_t nb, int32_t hist_len,
                   const std::vector<ColState> &cs) {
    kmsd_ctx *c = m->ctx;
    m->bcheb_state = 2;
    std::vector<double> h(size_t(hist_len) * nb * 2);
    read_status(c, {{h.data(), hist, h.size() * sizeof(double)}});
    double lo = __builtin_inf(), hi = -__builtin_inf();
    int used = 0;
    for (int32_t j = 0; j < nb; ++j) {
        const int64_t len = std::min<int64_t>(cs[size_t(j)].iters, hist_len);
        if (len < 6) continue;
        std::vector<double> a(static_cast<size_t>(len)), b2(static_cast<size_t>(len - 1));
        bool ok = true;
        for (int64_t k = 0; k < len; ++k) {
            const double al = h[(size_t(k) * nb + j) * 2];
            const double be_prev = k > 0 ? h[(size_t(k - 1) * nb + j) * 2 + 1] : 0.0;
            const double al_prev = k > 0 ? h[(size_t(k - 1) * nb + j) * 2] : 1.0;
            if (!(al > 0) || !(be_prev >= 0)) ok = false;
            a[size_t(k)] = 1.0 / al + (k > 0 ? be_prev / al_prev : 0.0);
            if (k + 1 < len) {
                const double be = h[(size_t(k) * nb + j) * 2 + 1];
                if (!(be >= 0)) ok = false;
                b2[size_t(k)] = be / (al * al);
            }
        }
        if (!ok) continue;
        double l, u;
        tri_extremes(a, b2, l, u);
        lo = std::min(lo, l);
        hi = std::max(hi, u);
        ++used;
    }
    if (!used || !std::isfinite(lo) || !std::isfinite(hi) || !(lo > 0)) return;
    lo *= 0.01 * c->bcheb_lo_pct;
    hi *= 1.02;
    if (!(hi > lo) || hi > 20.0 * lo) return;
    m->bcheb_lo = lo;
    m->bcheb_hi = hi;
    m->bcheb_state = 1;
}

// Block-Jacobi Chebyshev semi-iteration for S d = r with M^-1 S's spectrum
// taken as [bcheb_lo, bcheb_hi]: with b = M^-1 r, theta and delta the
// interval's center and half width and sigma = theta / delta,
//   d_1 = b / theta,
//   d_(k+1) = omega_(k+1) (d_k + (b - M^-1 S d_k) / theta) + (1 - omega_(k+1)) d_(k-1),
//   omega_2 = 1 / (1 - 1 / (2 sigma^2)), omega_(k+1) = 1 / (1 - omega_k / (4 sigma^2)),
// so that b - M^-1 S d_K = T_K((theta - M^-1 S) / delta) / T_K(sigma) b, the
// residual polynomial of least maximum on the interval; K is the least step
// count with 1 / T_K(sigma) <= reduction.  Each step after the first is one
// matrix-core product with the 4x4 block solve of its result and the
// combination with d_k, d_(k-1) and b in the epilogue (bsr4_mfma_blk_kernel),
// d_(k+1) replacing d_(k-1) row by row.  r is overwritten.  Returns the
// number of products (K - 1).
int cheb_block_run(kmsd_metric *m, void *r, void *d, int32_t nb, double reduction) {
    kmsd_ctx *c = m->ctx;
    const int64_t n = m->n;
    const double lo = m->bcheb_lo, hi = m->bcheb_hi, theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    const double sigma = theta / delta;
    const int K = std::max(1, std::min(400, int(std::ceil(std::acosh(1.0 / reduction) / std::acosh(sigma)))));
    void *b = m->work->z.ensure(c, size_t(n) * size_t(nb) * 8);
    const int64_t nq = n / 4;
    hipLaunchKernelGGL(block_apply32_kernel, dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>((nq * nb + 255) / 256, 16384)))),
                       dim3(256), 0, c->stream, r, b, m->factors32, nq, nb);
    HIPCHECK(hipGetLastError());
    // d_K lands in d: d_1 goes to d when K is odd, to r when K is even
    void *cur = (K % 2) ? d : r, *prev = (K % 2) ? r : d;
    launch_scale(c, KMSD_C64, b, nb, cur, nb, n, nb, Cplx{1.0 / theta, 0.0});
    Step s;
    s.A = m->S;
    s.bfac32 = m->factors32;
    s.ldx = nb;
    s.use_beta = true;
    s.W = b;
    s.ldw = nb;
    s.rhs = true;
    s.ldp = nb;
    s.ldo = nb;
    s.n = n;
    s.w = nb;
    s.f32_matrix = true;
    double omega = 1.0;
    for (int k = 1; k < K; ++k) {
        omega = k == 1 ? 1.0 / (1.0 - 1.0 / (2.0 * sigma * sigma)) : 1.0 / (1.0 - omega / (4.0 * sigma * sigma));
        Step u = s;
        u.X = cur;
        u.alpha = Cplx{-omega / theta, 0.0};
        u.beta = Cplx{omega, 0.0};
        u.coef = Cplx{omega / theta, 0.0};
        if (k > 1) {  // d_0 = 0
            u.P = prev;
            u.use_p = true;
            u.gamma = 1.0 - omega;
        }
        u.OUT = prev;
        launch_step(c, KMSD_C64, u);
        std::swap(prev, cur);
    }
    return K - 1;
}

}  // namespace

void pcg_solve(kmsd_metric *m, int dtype, const void *B, int64_t ldb, void *Y, int64_t ldy, int32_t nb, double rtol,
               double atol, int maxiter) {
    kmsd_ctx *c = m->ctx;
    if (m->direct) {
        direct_solve(m, dtype, B, ldb, Y, ldy, nb, rtol, atol);
        return;
    }
    if (dtype == KMSD_C128 && !m->mixed_c128) {
        Solver S = make_solver(m, dtype, nb);
        const Outcome o = pcg_run(S, B, ldb, Y, ldy, rtol, atol, maxiter);
        store_report(m, o.cs);
        account(m, o, 1);
        bool ok = true, fin = true, brk = false;
        for (const ColState &s : o.cs) {
            fin = fin && s.fin;
            brk = brk || s.breakdown;
        }
        for (size_t j = 0; j < m->rep_conv.size(); ++j) ok = ok && m->rep_conv[j];
        if (!ok)
            throw Error(KMSD_ENUMERIC,
                        !fin ? "S solve solution is nonfinite or not representable in complex128 after rescaling"
                        : brk ? "PCG non-positive/complex curvature breakdown"
                              : "S solve failed true-residual acceptance");
        return;
    }
    // complex64 (or complex128 with mixed_c128): FP32 inner solves inside
    // FP64 iterative refinement.
    if (!m->work) m->work = new kmsd_metric::Work();
    kmsd_metric::Work &wk = *m->work;
    const int64_t n = m->n;
    const size_t b128 = size_t(n) * nb * 16, b64 = size_t(n) * nb * 8;
    void *bb = wk.b128.ensure(c, b128);
    void *yy = wk.y128.ensure(c, b128);
    void *rr = wk.r128.ensure(c, b128);
    void *r64 = wk.r64.ensure(c, b64);
    void *d64 = wk.d64.ensure(c, b64);
    ColState *cs2 = static_cast<ColState *>(wk.cs2.ensure(c, sizeof(ColState) * size_t(nb)));
    double *nrm = static_cast<double *>(wk.part2.ensure(c, size_t(std::max<int64_t>(c->groups, 8)) * nb * 16));
    launch_convert(c, dtype, B, ldb, KMSD_C128, bb, nb, n, nb);
    // y = 0 and r = b are implicit in the first round: it reads b as its
    // residual and assigns y (zeroed here only if no round runs).
    // Column norms (FP64) are summed on the device into rsums[0, nb) for b
    // and rsums[nb, 2 nb) for the latest residual; each colsum also stores
    // the column scales of the next round's FP32 residual in cs2[j].safe
    // (rn / sqrt(n), 1 for a zero column), so rounds need no host copy.
    double *rsums = static_cast<double *>(wk.rsums.ensure(c, sizeof(double) * 2 * size_t(nb)));
    const double sqrt_n = std::sqrt(double(n));
    auto colnorm_launch = [&](const void *X, double *dst) {
        Step s;
        s.Z = X;
        s.ldz = nb;
        s.W = X;
        s.ldw = nb;
        s.part = nrm;
        s.n = n;
        s.w = nb;
        const int64_t g = launch_step(c, KMSD_C128, s);
        hipLaunchKernelGGL(colsum_kernel, dim3(unsigned(nb)), dim3(256), 0, c->stream, nrm, g, nb, dst, cs2, sqrt_n);
        HIPCHECK(hipGetLastError());
    };
    // CSR/ELL S: the residual product writes the norm partials itself
    const bool fuse_norm = m->S && m->S->fmt != KMSD_BSR4;
    std::vector<ColState> agg(nb);
    for (auto &s : agg) {
        s = ColState{};
        s.fin = 1;
    }
    int rounds = 0;
    // One refinement round: the FP32 residual r64 = r / scale, the inner
    // solve to inner_rtol, y += scale d, r = b - S y in FP64 and its column
    // norms on the device.  Returns 1 or 2 for a point series or block semi-iteration inner
    // solve, 0 for PCG.
    auto round = [&](double inner_rtol) {
        const int64_t g = std::max<int64_t>(1, std::min<int64_t>((n * nb + 255) / 256, 8192));
        // point semi-iteration: the conversion also writes its first step
        const bool point = cheb_usable(m);
        const int Kp = point ? cheb_steps(m, inner_rtol) : 0;
        void *t64 = point ? wk.t.ensure(c, b64) : nullptr;
        hipLaunchKernelGGL((scaled_convert_kernel<double, float>), dim3(unsigned(g)), dim3(256), 0, c->stream,
                           rounds ? rr : bb, int64_t(nb), r64, int64_t(nb), n, nb, cs2, 0,
                           point ? ((Kp % 2) ? d64 : t64) : nullptr, point ? cheb_first(m) : 0.0);
        HIPCHECK(hipGetLastError());
        Outcome o;
        int series = 0;
        if (point) {
            series = 1;
            const int K = cheb_run(m, r64, d64, t64, nb, Kp);
            o.cs.assign(size_t(nb), ColState{});
            for (ColState &x : o.cs) {
                x.iters = K;
                x.fin = 1;
            }
            o.fl = PcgFlags{};
        } else if (bcheb_candidate(m, nb) && m->bcheb_state == 1) {
            series = 2;
            const int K = cheb_block_run(m, r64, d64, nb, inner_rtol);
            o.cs.assign(size_t(nb), ColState{});
            for (ColState &x : o.cs) {
                x.iters = K;
                x.fin = 1;
            }
            o.fl = PcgFlags{};
        } else {
            Solver S = make_solver(m, KMSD_C64, nb);
            const bool record = bcheb_candidate(m, nb) && m->bcheb_state == 0;
            if (record) {
                S.hist_len = 64;
                S.hist = static_cast<double *>(wk.hist.ensure(c, sizeof(double) * 2 * 64 * size_t(nb)));
            }
            o = pcg_run(S, r64, nb, d64, nb, inner_rtol, 0.0, maxiter, false);
            if (record) bcheb_measure(m, S.hist, nb, S.hist_len, o.cs);
        }
        for (int32_t j = 0; j < nb; ++j) {
            agg[j].iters += o.cs[j].iters;
            agg[j].breakdown |= o.cs[j].breakdown;
            agg[j].fin &= o.cs[j].fin;
        }
        account(m, o, 1);
        --m->calls;
        hipLaunchKernelGGL((scaled_convert_kernel<float, double>), dim3(unsigned(g)), dim3(256), 0, c->stream, d64,
                           int64_t(nb), yy, int64_t(nb), n, nb, cs2, rounds ? 1 : 2);
        ++rounds;
        HIPCHECK(hipGetLastError());
        // r = b - S y in FP64
        Step s;
        s.A = m->S;
        s.X = yy;
        s.ldx = nb;
        s.P = bb;
        s.ldp = nb;
        s.gamma = 1.0;
        s.use_p = true;
        s.alpha = Cplx{-1, 0};
        s.OUT = rr;
        s.ldo = nb;
        s.n = n;
        s.w = nb;
        if (fuse_norm) {
            s.norm = true;
            s.part = nrm;
            const int64_t gr = launch_step(c, KMSD_C128, s);
            hipLaunchKernelGGL(colsum_kernel, dim3(unsigned(nb)), dim3(256), 0, c->stream, nrm, gr, nb, rsums + nb,
                               cs2, sqrt_n);
            HIPCHECK(hipGetLastError());
        } else {
            launch_step(c, KMSD_C128, s);
            colnorm_launch(rr, rsums + nb);
        }
        return series;
    };
    auto sqrt_all = [](std::vector<double> v) {
        for (double &x : v) x = std::sqrt(x);
        return v;
    };
    // Where the inner solves are series, rtol < 1e-7 (two rounds of FP32
    // inner solves are always needed) and the block is small (at most 2^22
    // elements: a series term takes about as long as a host round trip),
    // the first two rounds run back to back with the reductions fixed
    // beforehand: 1e-6 (what the first round would take, since every column
    // then needs rtol) and a quarter of rtol / 1e-6, which may cost a term
    // more than sizing the second round from the first one's residual;
    // ||b||, the second residual's norms and the finiteness of the converted
    // solution are read once after both (the conversion and its check are
    // repeated only if a further round runs).
    colnorm_launch(bb, rsums);
    std::vector<double> bn, rn, thr(nb);
    bool out_ready = false;
    int32_t out_nonfinite = 0;
    const double ir1 = std::max(rtol, 1e-6), ir2 = std::min(0.1, std::max(ir1, 0.25 * rtol / ir1));
    const bool ahead = atol == 0 && rtol < 1e-7 && n * nb <= (int64_t(1) << 22) &&
                       (cheb_usable(m) || (bcheb_candidate(m, nb) && m->bcheb_state == 1));
    if (ahead) {
        const int s1 = round(ir1), s2 = round(ir2);
        const int32_t *fl = launch_convert_check(c, KMSD_C128, yy, nb, dtype, Y, ldy, n, nb);
        std::vector<double> both(2 * size_t(nb));
        read_status(c, {{both.data(), rsums, sizeof(double) * 2 * size_t(nb)},
                        {&out_nonfinite, fl, sizeof(out_nonfinite)}});
        out_ready = true;
        bn = sqrt_all(std::vector<double>(both.begin(), both.begin() + nb));
        rn = sqrt_all(std::vector<double>(both.begin() + nb, both.end()));
        for (int32_t j = 0; j < nb; ++j) thr[j] = std::max(rtol * bn[j], atol);
        // the two series rounds together: the stall rule below with their
        // combined promise
        bool stalled = false;
        for (int32_t j = 0; j < nb; ++j)
            if (bn[j] > thr[j] && !(rn[j] <= std::sqrt(ir1 * ir2) * bn[j])) stalled = true;
        if (stalled) {
            if (s1 == 1 || s2 == 1) m->cheb_state = 2;
            if (s1 == 2 || s2 == 2) m->bcheb_state = 2;
        }
    } else {
        std::vector<double> v(static_cast<size_t>(nb));
        read_status(c, {{v.data(), rsums, sizeof(double) * size_t(nb)}});
        bn = sqrt_all(v);
        for (int32_t j = 0; j < nb; ++j) thr[j] = std::max(rtol * bn[j], atol);
        rn = bn;
    }
    for (int outer = rounds; outer < 12; ++outer) {
        bool done = true;
        // The inner solve reduces every column's residual by at least
        // inner_rtol: 1e-6 (about what FP32 iterations reach) when some
        // column is far from its threshold, else a quarter of the reduction
        // the strictest column still needs (the last round of a refinement
        // usually needs 1e-3 or 1e-4, not 1e-6).
        double need = 1.0;
        for (int32_t j = 0; j < nb; ++j) {
            done = done && rn[j] <= thr[j];
            if (rn[j] > thr[j]) need = std::min(need, thr[j] / rn[j]);
        }
        if (done) break;
        const double inner_rtol = std::min(0.1, std::max(std::max(rtol, 1e-6), 0.25 * need));
        const std::vector<double> before = rn;
        const int series = round(inner_rtol);
        out_ready = false;
        std::vector<double> v(static_cast<size_t>(nb));
        read_status(c, {{v.data(), rsums + nb, sizeof(double) * size_t(nb)}});
        rn = sqrt_all(v);
        // A series round whose interval misses the spectrum: some column that
        // still needed work reduced its residual by less than the square
        // root of the promised reduction (or not at all).  This metric's later
        // inner solves are PCG.
        if (series) {
            bool stalled = false;
            for (int32_t j = 0; j < nb; ++j)
                if (before[j] > thr[j] && !(rn[j] <= std::sqrt(inner_rtol) * before[j])) stalled = true;
            if (stalled) (series == 1 ? m->cheb_state : m->bcheb_state) = 2;
        }
        bool brk = false;
        for (int32_t j = 0; j < nb; ++j) brk = brk || agg[j].breakdown;
        if (brk) break;
    }
    ++m->calls;
    if (!rounds) launch_scale(c, KMSD_C128, bb, nb, yy, nb, n, nb, Cplx{0, 0});
    // finiteness of the stored solution per column
    bool fin_all;
    if (out_ready) {
        fin_all = out_nonfinite == 0;
    } else {
        launch_convert(c, KMSD_C128, yy, nb, dtype, Y, ldy, n, nb);
        fin_all = !launch_nonfinite(c, dtype, Y, ldy, n, nb);
    }
    for (auto &s : agg) s.fin = s.fin && fin_all;
    std::vector<double> qn = bn;
    store_report(m, agg, &rn, &qn, &thr);
    m->residual_checks += 0;
    bool ok = true, brk = false;
    for (size_t j = 0; j < m->rep_conv.size(); ++j) {
        ok = ok && m->rep_conv[j];
        brk = brk || m->rep_brk[j];
    }
    if (!ok)
        throw Error(KMSD_ENUMERIC, !fin_all ? "S solve solution is nonfinite or not representable after rescaling"
                                   : brk    ? "PCG non-positive/complex curvature breakdown"
                                            : "S solve failed true-residual acceptance");
}

void pcg_release(kmsd_metric *m) {
    delete m->work;
    m->work = nullptr;
}

}  // namespace kmsd
