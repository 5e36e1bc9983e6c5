// Square double-moment tables on the matrix cores (gfx950 MFMA).
//
// C[j][q] = sum_i conj(L[j][i]) R[q][i] for complex64 L (m x N) and R (k x N)
// with one batch column.  Reading the interleaved complex rows as real rows
// of length 2N (L', R'), the real part is L' R'^T and the imaginary part is
// L' R''^T with R''[q][2i] = Im R, R''[q][2i+1] = -Re R, so one pass over the
// operands feeds two real GEMMs.  Workgroup tile 64 x 64 (four wavefronts of
// 32 x 32), K staged through LDS in slabs of KT complex entries; the N
// reduction is split across workgroups into partial tables.
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
