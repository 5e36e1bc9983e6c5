// Square double-moment tables on the matrix cores (gfx950 MFMA).
//
// C[j][q] = sum_i conj(L[j][i]) R[q][i] for complex64 L (m x N) and R (k x N)
// with one batch column, as a real part sum_i (Lr Rr + Li Ri) and an
// imaginary part sum_i (Lr Ri - Li Rr).  Workgroup tiles of 128 x 128 with the
// N reduction split across workgroups into FP64 partial tables (tiling,
// operand layout and the XCD-aware numbering are described in
// kmsd_mu_mfma.hip).
//   KMSD_MU_FP32  : v_mfma_f32_32x32x2_f32, FP32 products and accumulation
//   KMSD_MU_BF16X3: operands split x = hi + lo in bf16; three products
//                   hi*hi + hi*lo + lo*hi with v_mfma_f32_32x32x16_bf16
//                   (about 16 significant bits per product, FP32 accumulation)
#pragma once
#include "kmsd_internal.hpp"
#include "kmsd_kernels.hpp"

namespace kmsd {

void mu_mfma(kmsd_ctx *c, const void *left, const void *right, double *table, int64_t n, int64_t m, int64_t k,
             int mode, bool accumulate);

}  // namespace kmsd
