// Host side of the C ABI declared in include/kmsd.h: device-resident
// matrices, the overlap metric (batched PCG), the engine's fused Chebyshev
// drivers and the launch-shape policy.  One HIP stream per context; every
// library allocation is charged against the context's HBM budget.
#include "kmsd.h"
#include "kmsd_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace kmsd;

namespace kmsd {
thread_local std::string g_create_error;
}

namespace {
// Live contexts of this process and the host thread they belong to: all of a
// process's contexts are created from one thread (see kmsd_ctx_create).
std::mutex g_owner_mu;
std::thread::id g_owner;
int64_t g_live = 0;
}  // namespace

// ---------------------------------------------------------------------------
// context

extern "C" int kmsd_ctx_create(int device, uint64_t budget, void *stream, kmsd_ctx **out) {
    if (!out) return KMSD_EINVAL;
    *out = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_owner_mu);
        if (g_live > 0 && g_owner != std::this_thread::get_id()) {
            g_create_error =
                "libkmsd contexts of one process belong to one host thread: another thread holds live "
                "contexts (run concurrent device work in separate processes)";
            return KMSD_EINVAL;
        }
        g_owner = std::this_thread::get_id();
        ++g_live;
    }
    kmsd_ctx *c = new kmsd_ctx();
    const int rc = guard(nullptr, [&] {
        int count = 0;
        HIPCHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw Error(KMSD_EINVAL, "invalid HIP device index");
        HIPCHECK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, device));
        c->device = device;
        c->cus = prop.multiProcessorCount;
        c->stream = static_cast<hipStream_t>(stream);
        size_t fr = 0, tot = 0;
        HIPCHECK(hipMemGetInfo(&fr, &tot));
        c->budget = budget ? budget : uint64_t(0.85 * double(fr));
        c->groups = int64_t(std::max(8, c->cus * 8));
        c->flag = static_cast<int32_t *>(dalloc(c, 256));
        HIPCHECK(hipMemset(c->flag, 0, 256));  // flag[1], flag[2] stay zero between uses (launch_convert_check)
        HIPCHECK(hipHostMalloc(&c->pin, kPinBytes, hipHostMallocDefault));
        noalloc_setup(device);
    });
    if (rc) {
        g_create_error = c->err;
        if (c->flag) (void)hipFree(c->flag);
        if (c->pin) (void)hipHostFree(c->pin);
        delete c;
        std::lock_guard<std::mutex> lock(g_owner_mu);
        --g_live;
        return rc;
    }
    *out = c;
    return KMSD_OK;
}

namespace kmsd {
void read_status(kmsd_ctx *c, std::initializer_list<StatusPart> parts) {
    size_t off = 0, staged[4] = {0, 0, 0, 0};
    bool pinned[4] = {false, false, false, false};
    int i = 0;
    for (const StatusPart &p : parts) {
        if (i < 4 && c->pin && off + p.bytes <= kPinBytes) {
            HIPCHECK(hipMemcpyAsync(static_cast<char *>(c->pin) + off, p.src, p.bytes, hipMemcpyDeviceToHost,
                                    c->stream));
            staged[i] = off;
            pinned[i] = true;
            off += (p.bytes + 63) & ~size_t(63);
        } else {
            HIPCHECK(hipMemcpyAsync(p.dst, p.src, p.bytes, hipMemcpyDeviceToHost, c->stream));
        }
        c->d2h_status_bytes += p.bytes;
        ++i;
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    i = 0;
    for (const StatusPart &p : parts) {
        if (i < 4 && pinned[i]) std::memcpy(p.dst, static_cast<const char *>(c->pin) + staged[i], p.bytes);
        ++i;
    }
}

void ctx_free(kmsd_ctx *c) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->flag) (void)hipFree(c->flag);
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->mu_part) (void)hipFree(c->mu_part);
    if (c->stats_tab) (void)hipFree(c->stats_tab);
    if (c->capture) (void)hipStreamDestroy(c->capture);
    delete c;
}
}  // namespace kmsd

extern "C" int kmsd_ctx_destroy(kmsd_ctx *c) {
    if (!c) return KMSD_OK;
    bool free_now = false;
    {
        std::lock_guard<std::recursive_mutex> lock(c->mu);
        if (c->closed) return KMSD_EINVAL;
        c->closed = true;
        free_now = c->children == 0;
    }
    {
        std::lock_guard<std::mutex> lock(g_owner_mu);
        --g_live;
    }
    if (free_now) ctx_free(c);
    return KMSD_OK;
}

extern "C" int kmsd_ctx_set_stream(kmsd_ctx *c, void *stream) {
    if (!c) return KMSD_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(c->mu);
    c->stream = static_cast<hipStream_t>(stream);
    return KMSD_OK;
}

extern "C" int kmsd_ctx_synchronize(kmsd_ctx *c) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] { HIPCHECK(hipStreamSynchronize(c->stream)); });
}

extern "C" int kmsd_ctx_transfers(const kmsd_ctx *c, uint64_t *out3) {
    if (!c || !out3) return KMSD_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(c->mu);
    out3[0] = c->h2d_bytes;
    out3[1] = c->d2h_data_bytes;
    out3[2] = c->d2h_status_bytes;
    return KMSD_OK;
}

extern "C" int kmsd_ctx_memory(const kmsd_ctx *c, uint64_t *used, uint64_t *budget) {
    if (!c) return KMSD_EINVAL;
    if (used) *used = c->used;
    if (budget) *budget = c->budget;
    return KMSD_OK;
}

extern "C" int kmsd_ctx_set_tuning(kmsd_ctx *c, const char *key, int64_t value) {
    if (!c || !key) return KMSD_EINVAL;
    return guard(c, [&] {
        const std::string k(key);
        if (k == "groups") {
            // 0: automatic (each fused step launches its kernel's resident capacity)
            if (value != 0 && (value < 8 || value % 8))
                throw Error(KMSD_EINVAL, "groups must be 0 (automatic) or a positive multiple of 8");
            c->auto_groups = value == 0;
            c->groups = value ? value : int64_t(std::max(8, c->cus * 8));
        } else if (k == "cpl" || k == "cpl_c64" || k == "cpl_c128") {
            if (value != 0 && value != 1 && value != 2 && value != 4)
                throw Error(KMSD_EINVAL, "cpl must be 0 (automatic), 1, 2 or 4");
            if (k != "cpl_c128") c->cpl[0] = int(value);
            if (k != "cpl_c64") c->cpl[1] = int(value);
        } else if (k == "moment_chunk") {
            if (value < 1 || value > 4096) throw Error(KMSD_EINVAL, "moment_chunk must be in [1, 4096]");
            c->moment_chunk = int(value);
        } else if (k == "fp32_math") {
            c->fp32_math = value != 0;
        } else if (k == "graphs") {
            c->graphs = value != 0;
        } else if (k == "ell_kernel") {
            c->ell_kernel = value != 0;
        } else if (k == "bsr_mfma") {
            c->bsr_mfma = value != 0;
        } else if (k == "bsr_lds") {
            c->bsr_lds = value != 0;
        } else if (k == "fuse_velocity") {
            c->fuse_velocity = value != 0;
        } else if (k == "tile_bytes") {
            if (value < 0 || value > (int64_t(1) << 30)) throw Error(KMSD_EINVAL, "tile_bytes must be in [0, 2^30]");
            c->tile_bytes = value;
        } else if (k == "pcg_poll") {
            if (value < 1) throw Error(KMSD_EINVAL, "pcg_poll must be positive");
            c->pcg_poll = int(value);
        } else if (k == "cheb_inner") {
            c->cheb_inner = value != 0;
        } else if (k == "bcheb_lo_pct") {
            if (value < 1 || value > 10000) throw Error(KMSD_EINVAL, "bcheb_lo_pct must be in [1, 10000]");
            c->bcheb_lo_pct = int(value);
        } else if (k == "min_blocks" || k == "pcg_min_blocks") {
            if (value < 1 || value > 64) throw Error(KMSD_EINVAL, "min_blocks must be in [1, 64]");
            (k == "min_blocks" ? c->min_blocks : c->pcg_min_blocks) = int(value);

        } else {
            throw Error(KMSD_EINVAL, "unknown tuning key");
        }
    });
}

extern "C" int kmsd_autotune(kmsd_ctx *c, const kmsd_matrix *A, int64_t nb, int dtype, int64_t *choice) {
    if (!c || !A || nb < 1 || nb > (int64_t(1) << 20) || !choice) return KMSD_EINVAL;
    const bool f32_values = (dtype & KMSD_TUNE_F32_VALUES) != 0;
    const bool tune_tile = (dtype & KMSD_TUNE_TILE) != 0;
    dtype &= ~int(KMSD_TUNE_F32_VALUES | KMSD_TUNE_TILE);
    return guard(c, [&] {
        if (dtype != KMSD_C64 && dtype != KMSD_C128) throw Error(KMSD_EINVAL, "dtype must be KMSD_C64 or KMSD_C128");
        if (f32_values && dtype != KMSD_C64) throw Error(KMSD_EINVAL, "FP32 matrix values need complex64 blocks");
        // Times the fused Chebyshev step of a moment chain (gathered X,
        // previous iterate P overwritten in place, projection W, moment
        // partials) for each candidate shape.
        const size_t bytes = size_t(A->n) * size_t(nb) * elem_bytes(dtype);
        const int64_t max_groups = std::max<int64_t>(4096, int64_t(c->cus) * 8);
        const size_t pbytes = size_t(max_groups) * size_t(nb) * 2 * sizeof(double);
        void *X = nullptr, *P = nullptr, *W = nullptr, *part = nullptr;
        const int slot = dtype == KMSD_C64 ? 0 : 1;  // the tuned precision's columns-per-lane setting
        const int saved_cpl = c->cpl[slot];
        const int64_t saved_groups = c->groups;
        const bool saved_auto = c->auto_groups;
        const int64_t saved_tile = c->tile_bytes;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            e0 = e1 = nullptr;
            (void)hipStreamSynchronize(c->stream);
            dfree(c, part, pbytes);
            dfree(c, W, bytes);
            dfree(c, P, bytes);
            dfree(c, X, bytes);
        };
        try {
            X = dalloc(c, bytes);
            P = dalloc(c, bytes);
            W = dalloc(c, bytes);
            part = dalloc(c, pbytes);
            HIPCHECK(hipMemsetAsync(X, 0, bytes, c->stream));
            HIPCHECK(hipMemsetAsync(P, 0, bytes, c->stream));
            HIPCHECK(hipMemsetAsync(W, 0, bytes, c->stream));
            HIPCHECK(hipEventCreate(&e0));
            HIPCHECK(hipEventCreate(&e1));
            const int def_cpl = default_cpl(dtype, A);
            const int cpls[3] = {1, 2, 4};
            const int64_t groups[4] = {0, 1024, 2048, 4096};  // 0 = resident capacity
            auto time_shape = [&]() -> double {
                Step s;
                s.A = A;
                s.X = X;
                s.ldx = nb;
                s.P = P;
                s.ldp = nb;
                s.OUT = P;
                s.ldo = nb;
                s.part = static_cast<double *>(part);
                s.alpha = Cplx{1.0, 0};
                if (f32_values) {
                    // the inner solves' product: S p stored, <p, S p> partials
                    s.W = X;
                    s.ldw = nb;
                    s.f32_matrix = true;
                } else {
                    // a moment-chain order: 2 H T_n - T_(n-1) in place, <W, T_(n+1)> partials
                    s.W = W;
                    s.ldw = nb;
                    s.beta = Cplx{-0.5, 0};
                    s.gamma = -1.0;
                    s.use_beta = s.use_p = true;
                }
                s.n = A->n;
                s.w = int32_t(nb);
                launch_step(c, dtype, s);  // warm-up
                float t[3];
                for (float &v : t) {
                    HIPCHECK(hipEventRecord(e0, c->stream));
                    launch_step(c, dtype, s);
                    HIPCHECK(hipEventRecord(e1, c->stream));
                    HIPCHECK(hipEventSynchronize(e1));
                    HIPCHECK(hipEventElapsedTime(&v, e0, e1));
                }
                std::sort(t, t + 3);
                return double(t[1]);
            };
            double best = 0, def_time = 0;
            int best_cpl = def_cpl;
            int64_t best_groups = 0;
            for (int cp : cpls) {
                if (dtype == KMSD_C128 && cp == 4) continue;
                if (nb % cp) continue;
                for (int64_t g : groups) {
                    c->cpl[slot] = cp;
                    c->auto_groups = g == 0;
                    c->groups = g ? g : int64_t(std::max(8, c->cus * 8));
                    const double med = time_shape();
                    if (cp == def_cpl && g == 0) def_time = med;
                    if (best == 0 || med < best) {
                        best = med;
                        best_cpl = cp;
                        best_groups = g;
                    }
                }
            }
            // Keep the default shape unless another is clearly (> 3 %) faster.
            if (def_time > 0 && best > 0.97 * def_time) {
                best_cpl = def_cpl;
                best_groups = 0;
            }
            c->cpl[slot] = best_cpl;
            c->auto_groups = best_groups == 0;
            c->groups = best_groups ? best_groups : int64_t(std::max(8, c->cus * 8));
            choice[0] = best_cpl;
            choice[1] = best_groups;
            if (tune_tile) {
                // tile size of the plane-tiled 4x4-block walk, with the chosen shape
                int64_t best_tile = c->tile_bytes;
                if (A->fmt == KMSD_BSR4 && A->plane) {
                    const int64_t tiles[4] = {int64_t(1) << 20, 0, int64_t(1) << 19, int64_t(1) << 21};
                    double tbest = 0, tdef = 0;
                    for (int64_t tb : tiles) {
                        c->tile_bytes = tb;
                        const double med = time_shape();
                        if (tb == (int64_t(1) << 20)) tdef = med;
                        if (tbest == 0 || med < tbest) {
                            tbest = med;
                            best_tile = tb;
                        }
                    }
                    if (tbest > 0.97 * tdef) best_tile = int64_t(1) << 20;
                }
                c->tile_bytes = best_tile;
                choice[2] = best_tile;
            }
        } catch (...) {
            c->cpl[slot] = saved_cpl;
            c->groups = saved_groups;
            c->auto_groups = saved_auto;
            c->tile_bytes = saved_tile;
            release();
            throw;
        }
        release();
    });
}

extern "C" int kmsd_time_step(kmsd_ctx *c, const kmsd_matrix *A, int64_t nb, int dtype, double *us) {
    if (!c || !A || nb < 1 || nb > (int64_t(1) << 20) || !us) return KMSD_EINVAL;
    return guard(c, [&] {
        if (dtype != KMSD_C64 && dtype != KMSD_C128) throw Error(KMSD_EINVAL, "dtype must be KMSD_C64 or KMSD_C128");
        // One fused Chebyshev step of a moment chain with the context's
        // launch shape: 2 A X - P into P, partials <W, P>; median of three
        // after a warm-up.
        const size_t bytes = size_t(A->n) * size_t(nb) * elem_bytes(dtype);
        const int64_t groups = std::max<int64_t>(step_groups(c, A->n, int32_t(nb), dtype, 4, A), 8);
        const size_t pbytes = size_t(groups) * size_t(nb) * 2 * sizeof(double);
        void *X = nullptr, *P = nullptr, *W = nullptr, *part = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            e0 = e1 = nullptr;
            (void)hipStreamSynchronize(c->stream);
            dfree(c, part, pbytes);
            dfree(c, W, bytes);
            dfree(c, P, bytes);
            dfree(c, X, bytes);
        };
        try {
            X = dalloc(c, bytes);
            P = dalloc(c, bytes);
            W = dalloc(c, bytes);
            part = dalloc(c, pbytes);
            HIPCHECK(hipMemsetAsync(X, 0, bytes, c->stream));
            HIPCHECK(hipMemsetAsync(P, 0, bytes, c->stream));
            HIPCHECK(hipMemsetAsync(W, 0, bytes, c->stream));
            HIPCHECK(hipEventCreate(&e0));
            HIPCHECK(hipEventCreate(&e1));
            Step s;
            s.A = A;
            s.X = X;
            s.ldx = nb;
            s.P = P;
            s.ldp = nb;
            s.OUT = P;
            s.ldo = nb;
            s.W = W;
            s.ldw = nb;
            s.part = static_cast<double *>(part);
            s.alpha = Cplx{1.0, 0};
            s.beta = Cplx{-0.5, 0};
            s.gamma = -1.0;
            s.use_beta = s.use_p = true;
            s.n = A->n;
            s.w = int32_t(nb);
            s.groups = groups;
            launch_step(c, dtype, s);  // warm-up
            float t[3];
            for (float &v : t) {
                HIPCHECK(hipEventRecord(e0, c->stream));
                launch_step(c, dtype, s);
                HIPCHECK(hipEventRecord(e1, c->stream));
                HIPCHECK(hipEventSynchronize(e1));
                HIPCHECK(hipEventElapsedTime(&v, e0, e1));
            }
            std::sort(t, t + 3);
            *us = 1000.0 * double(t[1]);
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

extern "C" const char *kmsd_last_error(const kmsd_ctx *c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

extern "C" int kmsd_build_info(char *buf, int64_t len) {
    if (!buf || len < 1) return KMSD_EINVAL;
    std::snprintf(buf, size_t(len),
                  "{\"language\":\"HIP C++17\",\"target\":\"gfx950\",\"precision\":[\"complex128\",\"complex64+fp64 "
                  "reductions\"],\"compiler\":\"%s\"}",
                  __VERSION__);
    return KMSD_OK;
}

// ---------------------------------------------------------------------------
// matrices

extern "C" int kmsd_matrix_create(kmsd_ctx *c, int fmt, int64_t n, int64_t nnz, const int64_t *rowptr,
                                  const int64_t *colidx, const double *values, kmsd_matrix **out) {
    if (!c || !out) return KMSD_EINVAL;
    *out = nullptr;
    kmsd_matrix *m = new kmsd_matrix();
    m->ctx = c;
    const int rc = guard(c, [&] {
        matrix_upload(c, m, fmt, n, nnz, rowptr, colidx, values);
        ctx_adopt(c);
    });
    if (rc) {
        matrix_release(m);
        delete m;
        return rc;
    }
    *out = m;
    return KMSD_OK;
}

extern "C" int kmsd_matrix_destroy(kmsd_matrix *m) {
    if (!m) return KMSD_OK;
    kmsd_ctx *c = m->ctx;
    bool free_ctx = false;
    {
        std::lock_guard<std::recursive_mutex> lock(c->mu);
        matrix_release(m);
        delete m;
        free_ctx = ctx_orphan(c);
    }
    if (free_ctx) ctx_free(c);
    return KMSD_OK;
}

extern "C" int kmsd_matrix_info(const kmsd_matrix *m, int64_t *out) {
    if (!m || !out) return KMSD_EINVAL;
    out[0] = m->fmt;
    out[1] = m->n;
    out[2] = m->nnz;
    out[3] = m->noff;
    out[4] = int64_t(m->bytes);
    out[5] = m->vkind;
    out[6] = m->uniform_slot >= 0 ? 1 : 0;
    out[7] = m->plane;
    return KMSD_OK;
}

extern "C" int kmsd_matrix_matmul(kmsd_matrix *m, const void *B, void *out, int64_t nb, int dtype) {
    if (!m) return KMSD_EINVAL;
    kmsd_ctx *c = m->ctx;
    return guard(c, [&] {
        check_block_args(B, out, nb, dtype);
        require_finite(c, dtype, B, nb, m->n, nb);
        spmm(c, dtype, m, B, nb, out, nb, nb, 1.0, 0.0, false);
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}
