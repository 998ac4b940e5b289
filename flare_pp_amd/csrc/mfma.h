// mfma.h -- hand-written CDNA4 matrix-core contractions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace flare {

// K padding of the BF16 operands (multiple of the GEMM's K step).
int bf16_padded_k(int K);

constexpr int kGemmMaxBatch = 8;

// For q < nb independent products (one launch):
//   Z[a_row0[q] + i][:] = (Ah + Al)[a_row0[q] + i] . (Bh + Bl)[b_row0[q] + :]^T
// for i < M[q], N columns, FP32 result with leading dimension ldz, on BF16
// MFMA with FP32 accumulation, three products hi*hi + hi*lo + lo*hi.
// A and B rows have Kp BF16 values (a multiple of bf16_padded_k's step, the
// padding zero); row n of a symmetric right operand is its column n.
void gemm_bf16x3_batched(hipStream_t st, int nb, const int *M,
                         const int64_t *a_row0, const int64_t *b_row0, int N,
                         int Kp, const void *Ah, const void *Al, const void *Bh,
                         const void *Bl, float *Z, int ldz);

// hi/lo BF16 split of a rows x cols FP64 matrix into rows of kp (zero padded)
void split_bf16(hipStream_t st, int rows, int cols, int kp, const double *x,
                void *hi, void *lo);

}  // namespace flare
