// Batched PCG on the device (reference Metric::solve, cpp/native.cpp:313-452).
#pragma once
#include "kmsd_internal.hpp"

namespace kmsd {
void pcg_solve(kmsd_metric *m, int dtype, const void *B, int64_t ldb, void *Y, int64_t ldy, int32_t nb, double rtol,
               double atol, int maxiter);
void pcg_release(kmsd_metric *m);
}  // namespace kmsd
