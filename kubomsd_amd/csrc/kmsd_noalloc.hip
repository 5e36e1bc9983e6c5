// kmsd_moments_noalloc: the moment chain of kmsd_moments for callers that
// cannot allocate device memory (serving loops and harnesses that forbid
// allocation).  Its two chain blocks and the partial sums live in one
// fixed-size array in device memory, `scratch`, which every GPU the library
// runs on holds from the first context created on it until the process ends.
#include <mutex>

#include "kmsd.h"
#include "kmsd_internal.hpp"

using namespace kmsd;

// The scratch of kmsd_moments_noalloc (256 MiB per device).
static constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
// Device address of `scratch` per device (each GPU holds its own copy).  In
// host code the array's name is only a host-side stand-in: every use goes
// through these addresses.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

size_t round256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

namespace kmsd {
// Looks up the device address of `scratch` on `device` (with that device
// current), once per device; called by kmsd_ctx_create.
void noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw Error(KMSD_EINVAL, "invalid HIP device index");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    int current = 0;
    HIPCHECK(hipGetDevice(&current));
    if (current != device) HIPCHECK(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (rc != hipSuccess || !p)
        throw Error(KMSD_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
    g_scratch_ptr[device] = p;
}
}  // namespace kmsd

extern "C" int kmsd_moments_noalloc(kmsd_engine *e, const void *B, double *mu, int64_t L, int64_t nb, int dtype,
                                    double center, double half) {
    if (!e) return KMSD_EINVAL;
    kmsd_ctx *c = e->ctx;
    return guard(c, [&] {
        if (!e->metric->identity)
            throw Error(KMSD_EINVAL, "kmsd_moments_noalloc needs an identity metric (S = I, diagonal preconditioner)");
        check_block_args(B, mu, nb, dtype);
        if (L < 1) throw Error(KMSD_EINVAL, "moment count must be positive");
        void *base = nullptr;
        {
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            if (c->device >= 0 && c->device < MAX_DEVICES) base = g_scratch_ptr[c->device];
        }
        if (!base) throw Error(KMSD_EDEVICE, "kmsd_moments_noalloc: no scratch address for this device");
        MomentScratch sc = moment_scratch_need(e, B, L, nb, dtype, c->moment_chunk);
        const size_t blocks = 2 * round256(sc.block_bytes);
        const size_t per_order = sc.part_bytes / size_t(sc.chunk);
        if (blocks + per_order > NOALLOC_SCRATCH_BYTES)
            throw Error(KMSD_EINVAL, "kmsd_moments_noalloc: an n x nb block of " + std::to_string(sc.block_bytes) +
                                         " bytes needs more than its fixed scratch of " +
                                         std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes (two blocks and partials)");
        // as many orders per partial reduction as fit (the moments do not
        // depend on it: every order's partials reduce in the same fixed tree)
        sc.chunk = std::min<int64_t>(sc.chunk, int64_t((NOALLOC_SCRATCH_BYTES - blocks) / per_order));
        sc.part_bytes = per_order * size_t(sc.chunk);
        char *p = static_cast<char *>(base);
        sc.t1 = p;
        sc.t2 = p + round256(sc.block_bytes);
        sc.part = reinterpret_cast<double *>(p + blocks);
        moment_chain(e, B, mu, L, nb, dtype, center, half, &sc);
    });
}
