// mfma.h -- hand-written CDNA4 matrix-core contractions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace flare {

// K padding of the BF16 operands (multiple of the GEMM's K step).
int bf16_padded_k(int K);

constexpr int kGemmMaxBatch = 8;

// One product Z = (Ah + Al) . (Bh + Bl)^T: A has M rows and B has N rows of
// K BF16 values (K a multiple of bf16_padded_k's step, the padding zero),
// starting at element offsets a_off / b_off; Z (FP32) starts at z_off with
// leading dimension ldz.  Row n of a symmetric right operand is its
// column n.
struct GemmProduct {
  int M, N, K, ldz;
  int64_t a_off, b_off, z_off;
};

// Up to kGemmMaxBatch products in one launch on BF16 MFMA with FP32
// accumulation, three products hi*hi + hi*lo + lo*hi each.
void gemm_bf16x3_batched(hipStream_t st, int nb, const GemmProduct *prods,
                         const void *Ah, const void *Al, const void *Bh,
                         const void *Bl, float *Z);

// hi/lo BF16 split of a rows x cols FP64 matrix into rows of kp (zero padded)
void split_bf16(hipStream_t st, int rows, int cols, int kp, const double *x,
                void *hi, void *lo);

}  // namespace flare
