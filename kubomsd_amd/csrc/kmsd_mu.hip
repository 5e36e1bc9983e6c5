// Spectral double-moment tables mu_jk (tall-skinny complex GEMM over N).
#include "kmsd_internal.hpp"

using namespace kmsd;

extern "C" int kmsd_mu_table(kmsd_ctx *c, const void *left, const void *right, void *table, int64_t n, int64_t m,
                             int64_t k, int64_t nb, int dtype, int mode, int accumulate) {
    if (!c) return KMSD_EINVAL;
    return guard(c, [&] { throw Error(KMSD_EINVAL, "mu tables are not available in this build"); });
}
