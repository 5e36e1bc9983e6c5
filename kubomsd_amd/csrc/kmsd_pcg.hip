// Batched PCG on the device (reference Metric::solve, cpp/native.cpp:313-452).
#include "kmsd_pcg.hpp"

namespace kmsd {
void pcg_solve(kmsd_metric *, int, const void *, int64_t, void *, int64_t, int32_t, double, double, int) {
    throw Error(KMSD_EINVAL, "nonidentity metric solves are not available in this build");
}
void pcg_release(kmsd_metric *) {}
}  // namespace kmsd
