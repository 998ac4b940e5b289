// mfma.h -- hand-written CDNA4 matrix-core contractions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace flare {

// Z (M x N, row-major FP32) = X (M x K, row-major FP32) * Bsym, where Bsym
// (K x N, row-major FP32) is symmetric with K == N, on BF16 MFMA with FP32
// accumulation: each operand is split into hi + lo BF16 parts and the
// three products hi*hi + hi*lo + lo*hi are accumulated, which keeps ~16
// significant bits per operand.
void gemm_bf16x3(hipStream_t st, int M, int N, int K, const float *X,
                 const float *Bsym, float *Z, void *workspace);
size_t gemm_bf16x3_workspace(int M, int N, int K);

}  // namespace flare
