// mfma.h -- hand-written CDNA4 matrix-core contractions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace flare {

// K padding of the BF16 operands (multiple of the GEMM's K step).
int bf16_padded_k(int K);

// Z (M x N, FP32, leading dimension ldz) = (Ah + Al) . (Bh + Bl)^T on BF16
// MFMA with FP32 accumulation, three products hi*hi + hi*lo + lo*hi.
// Ah/Al: M x Kp BF16 rows, Bh/Bl: N x Kp BF16 rows (row n of a symmetric
// right operand is its column n); Kp a multiple of bf16_padded_k's step and
// the padding zero.
void gemm_bf16x3(hipStream_t st, int M, int N, int Kp, const void *Ah,
                 const void *Al, const void *Bh, const void *Bl, float *Z,
                 int ldz);

// hi/lo BF16 split of a rows x cols FP64 matrix into rows of kp (zero padded)
void split_bf16(hipStream_t st, int rows, int cols, int kp, const double *x,
                void *hi, void *lo);

}  // namespace flare
