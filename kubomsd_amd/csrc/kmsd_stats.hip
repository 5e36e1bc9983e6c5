// Device statistics of the stochastic estimator (reference
// src/kubo_msd/solvers/stochastic.py:52-119, observables/statistics.py:5-23,
// observables/diffusion.py:4-6; the GPU edition keeps the same block sums
// resident: kubo_msd_gpu/src/kubo_msd/solvers/stochastic_gpu.py:53-145).
//
//   kmsd_block_accumulate  per-sample sums into statistics blocks, in sample
//                          order (bitwise the reference's loop of += per sample)
//   kmsd_time_gradient     1/2 d/dt of MSD numerators with numpy.gradient's
//                          second-order formulas and operation order
//   kmsd_bootstrap_ratio   ratio of block sums and its paired-bootstrap
//                          standard error (resampling weights from the host)
//   kmsd_dos_validity      valid / reason_code masks of the DOS estimate
//
// Floating-point contraction is off in this file: every product and sum is
// rounded separately, as numpy evaluates the same expressions.
#pragma clang fp contract(off)

#include <cmath>
#include <vector>

#include "kmsd_internal.hpp"

using namespace kmsd;

namespace {

constexpr uint32_t R_LOW_DOS = 1, R_UNCERTAIN_TRACE = 2, R_INSUFFICIENT_BLOCKS = 32;

int64_t grid_for(int64_t total) { return std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 16384)); }

__global__ __launch_bounds__(256) void accumulate_kernel(double *dst, const double *src, int64_t nloc, int64_t nelem,
                                                         int64_t ld, int32_t *flags, int32_t bit) {
    bool bad = false;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < nelem; e += int64_t(gridDim.x) * 256) {
        double acc = dst[e];
        for (int64_t l = 0; l < nloc; ++l) {
            const double v = src[l * ld + e];
            bad |= !__builtin_isfinite(v);
            acc = acc + v;
        }
        dst[e] = acc;
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flags, bit);
}

// One output time index of numpy.gradient along a time axis.
struct GradCoef {
    int32_t kind;        // 0: ((w0 f[j0] + w1 f[j1]) + w2 f[j2]); 1: (f[j2] - f[j0]) / den
    int32_t j0, j1, j2;
    double w0, w1, w2, den;
};

__global__ __launch_bounds__(256) void gradient_kernel(const double *src, double *dst, int64_t outer, int64_t nt,
                                                       int64_t inner, const GradCoef *coef, double scale) {
    const int64_t total = outer * nt * inner;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += int64_t(gridDim.x) * 256) {
        const int64_t i = e % inner;
        const int64_t j = (e / inner) % nt;
        const int64_t o = e / (inner * nt);
        const double *f = src + o * nt * inner + i;
        const GradCoef c = coef[j];
        double g;
        if (c.kind == 0) {
            const double p0 = c.w0 * f[c.j0 * inner];
            const double p1 = c.w1 * f[c.j1 * inner];
            const double p2 = c.w2 * f[c.j2 * inner];
            g = (p0 + p1) + p2;
        } else {
            g = (f[c.j2 * inner] - f[c.j0 * inner]) / c.den;
        }
        dst[e] = scale * g;
    }
}

// Element e of A (nb blocks x nelem) pairs with column e % ne of D (nb x dcols,
// dcols = ne or 1).  value = sum_b A / sum_b D (NaN when sum_b D == 0);
// stderr = sample standard deviation (ddof 1) over the nres resampled ratios
// (sum_b w[r,b] A[b,e]) / (sum_b w[r,b] D[b,k]), NaN when any is undefined.
// valid (ne entries, optional) sets value and stderr to NaN where it is 0.
__global__ __launch_bounds__(256) void bootstrap_kernel(const double *A, const double *D, int64_t nb, int64_t nelem,
                                                        int64_t ne, int64_t dcols, const double *w, int64_t nres,
                                                        const uint8_t *valid, double *value, double *se,
                                                        double *total) {
    const double nan = __builtin_nan("");
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < nelem; e += int64_t(gridDim.x) * 256) {
        const int64_t k = dcols == 1 ? 0 : e % ne;
        double sa = A[e], sd = D[k];
        for (int64_t b = 1; b < nb; ++b) {
            sa = sa + A[b * nelem + e];
            sd = sd + D[b * dcols + k];
        }
        const bool keep = !valid || valid[e % ne];
        if (total) total[e] = sa;
        value[e] = keep && sd != 0.0 ? sa / sd : nan;
        if (!se) continue;
        auto ratio = [&](int64_t r) {
            const double *wr = w + r * nb;
            double ra = 0.0, rd = 0.0;
            for (int64_t b = 0; b < nb; ++b) {
                ra = __builtin_fma(wr[b], A[b * nelem + e], ra);
                rd = __builtin_fma(wr[b], D[b * dcols + k], rd);
            }
            return rd != 0.0 ? ra / rd : nan;
        };
        double sum = 0.0;
        for (int64_t r = 0; r < nres; ++r) sum = sum + ratio(r);
        const double mean = sum / double(nres);
        double ss = 0.0;
        for (int64_t r = 0; r < nres; ++r) {
            const double d = ratio(r) - mean;
            ss = ss + d * d;
        }
        se[e] = keep ? std::sqrt(ss / double(nres - 1)) : nan;
    }
}

__global__ __launch_bounds__(256) void dos_validity_kernel(const double *rho, const double *se, int64_t ne,
                                                           double threshold, int32_t insufficient, uint8_t *valid,
                                                           uint32_t *reason) {
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < ne; e += int64_t(gridDim.x) * 256) {
        const double r = rho[e];
        bool v = __builtin_isfinite(r) && r > threshold;
        uint32_t code = v ? 0u : R_LOW_DOS;
        if (se) {
            const double s = se[e];
            const double lower = r - 1.96 * s;
            const bool uncertain = !__builtin_isfinite(s) || lower <= 0.0;
            v = v && !uncertain;
            if (uncertain) code |= R_UNCERTAIN_TRACE;
        }
        if (insufficient) code |= R_INSUFFICIENT_BLOCKS;
        valid[e] = v ? 1 : 0;
        reason[e] = code;
    }
}

// numpy.gradient coefficients along a coordinate array t (numpy's own
// expressions and operation order; uniform spacing when every difference
// equals the first, as numpy reduces it to a scalar spacing).
std::vector<GradCoef> gradient_table(const double *t, int64_t nt, int edge_order) {
    std::vector<double> dx(size_t(nt - 1));
    for (int64_t j = 0; j + 1 < nt; ++j) dx[j] = t[j + 1] - t[j];
    bool uniform = true;
    for (double v : dx) uniform = uniform && v == dx[0];
    std::vector<GradCoef> c(static_cast<size_t>(nt));
    for (int64_t j = 1; j + 1 < nt; ++j) {
        GradCoef &g = c[j];
        g.j0 = int32_t(j - 1);
        g.j1 = int32_t(j);
        g.j2 = int32_t(j + 1);
        if (uniform) {
            g.kind = 1;
            g.den = 2. * dx[0];
        } else {
            const double dx1 = dx[j - 1], dx2 = dx[j];
            g.kind = 0;
            g.w0 = -(dx2) / (dx1 * (dx1 + dx2));
            g.w1 = (dx2 - dx1) / (dx1 * dx2);
            g.w2 = dx1 / (dx2 * (dx1 + dx2));
        }
    }
    GradCoef &a = c[0], &z = c[size_t(nt - 1)];
    if (edge_order == 1) {
        a.kind = z.kind = 1;
        a.j0 = 0;
        a.j2 = 1;
        a.den = dx[0];
        z.j0 = int32_t(nt - 2);
        z.j2 = int32_t(nt - 1);
        z.den = uniform ? dx[0] : dx[size_t(nt - 2)];
    } else {
        a.kind = z.kind = 0;
        a.j0 = 0;
        a.j1 = 1;
        a.j2 = 2;
        z.j0 = int32_t(nt - 3);
        z.j1 = int32_t(nt - 2);
        z.j2 = int32_t(nt - 1);
        if (uniform) {
            const double h = dx[0];
            a.w0 = -1.5 / h;
            a.w1 = 2. / h;
            a.w2 = -0.5 / h;
            z.w0 = 0.5 / h;
            z.w1 = -2. / h;
            z.w2 = 1.5 / h;
        } else {
            double dx1 = dx[0], dx2 = dx[1];
            a.w0 = -(2. * dx1 + dx2) / (dx1 * (dx1 + dx2));
            a.w1 = (dx1 + dx2) / (dx1 * dx2);
            a.w2 = -dx1 / (dx2 * (dx1 + dx2));
            dx1 = dx[size_t(nt - 3)];
            dx2 = dx[size_t(nt - 2)];
            z.w0 = (dx2) / (dx1 * (dx1 + dx2));
            z.w1 = -(dx2 + dx1) / (dx1 * dx2);
            z.w2 = (2. * dx2 + dx1) / (dx2 * (dx1 + dx2));
        }
    }
    return c;
}

}  // namespace

extern "C" int kmsd_block_accumulate(kmsd_ctx *c, double *dst, const double *src, int64_t nloc, int64_t nelem,
                                     int64_t ld_src, int32_t *flags, int32_t flag_bits) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] {
        if (nloc < 0 || nelem < 0 || ld_src < nelem) throw Error(KMSD_EINVAL, "invalid accumulation shape");
        if (nloc == 0 || nelem == 0) return;
        if (!dst || !src || !flags) throw Error(KMSD_EINVAL, "null device buffer");
        hipLaunchKernelGGL(accumulate_kernel, dim3(unsigned(grid_for(nelem))), dim3(256), 0, c->stream, dst, src,
                           nloc, nelem, ld_src, flags, flag_bits);
        HIPCHECK(hipGetLastError());
    });
}

extern "C" int kmsd_time_gradient(kmsd_ctx *c, const double *src, double *dst, int64_t outer, int64_t nt,
                                  int64_t inner, const double *time, double scale) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] {
        if (outer < 0 || inner < 0 || nt < 2 || !time) throw Error(KMSD_EINVAL, "time gradient needs at least 2 times");
        for (int64_t j = 0; j + 1 < nt; ++j)
            if (!(time[j + 1] > time[j])) throw Error(KMSD_EINVAL, "time grid must be strictly increasing");
        if (outer == 0 || inner == 0) return;
        if (!src || !dst || src == dst) throw Error(KMSD_EINVAL, "time gradient needs distinct device buffers");
        const std::vector<GradCoef> tab = gradient_table(time, nt, nt > 2 ? 2 : 1);
        const size_t bytes = sizeof(GradCoef) * tab.size();
        if (c->stats_tab_bytes < bytes) {  // grow-only table, reused by later calls
            HIPCHECK(hipStreamSynchronize(c->stream));
            dfree(c, c->stats_tab, c->stats_tab_bytes);
            c->stats_tab = nullptr;
            c->stats_tab_bytes = 0;
            c->stats_tab = dalloc(c, bytes);
            c->stats_tab_bytes = bytes;
        }
        // The table is copied from host memory that lives until the
        // synchronization below, which also orders it before a later call
        // rewrites it.
        copy_h2d(c, c->stats_tab, tab.data(), bytes);
        hipLaunchKernelGGL(gradient_kernel, dim3(unsigned(grid_for(outer * nt * inner))), dim3(256), 0, c->stream,
                           src, dst, outer, nt, inner, static_cast<const GradCoef *>(c->stats_tab), scale);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

extern "C" int kmsd_bootstrap_ratio(kmsd_ctx *c, const double *A, const double *D, int64_t nblocks, int64_t nelem,
                                    int64_t ne, int64_t dcols, const double *weights, int64_t nres,
                                    const uint8_t *valid, double *value, double *stderr_out, double *total) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] {
        if (nblocks < 1 || nelem < 0 || ne < 1 || nelem % ne || (dcols != 1 && dcols != ne))
            throw Error(KMSD_EINVAL, "statistical denominator must be (blocks,energy)");
        if (stderr_out && (nblocks < 2 || nres < 2 || !weights))
            throw Error(KMSD_EINVAL, "bootstrap needs at least two blocks, two resamples and weights");
        if (nelem == 0) return;
        if (!A || !D || !value) throw Error(KMSD_EINVAL, "null device buffer");
        hipLaunchKernelGGL(bootstrap_kernel, dim3(unsigned(grid_for(nelem))), dim3(256), 0, c->stream, A, D, nblocks,
                           nelem, ne, dcols, stderr_out ? weights : nullptr, stderr_out ? nres : 0, valid, value,
                           stderr_out, total);
        HIPCHECK(hipGetLastError());
    });
}

extern "C" int kmsd_dos_validity(kmsd_ctx *c, const double *rho, const double *stderr_in, int64_t ne,
                                 double threshold, int64_t nblocks, uint8_t *valid, uint32_t *reason) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] {
        if (ne < 0 || nblocks < 1) throw Error(KMSD_EINVAL, "invalid DOS validity shape");
        if (ne == 0) return;
        if (!rho || !valid || !reason) throw Error(KMSD_EINVAL, "null device buffer");
        hipLaunchKernelGGL(dos_validity_kernel, dim3(unsigned(grid_for(ne))), dim3(256), 0, c->stream, rho, stderr_in,
                           ne, threshold, int32_t(nblocks < 8), valid, reason);
        HIPCHECK(hipGetLastError());
    });
}
