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
// rs / cs (optional, FP32, M and N values) multiply row i / column j of the
// result: the inverse scales of pre-scaled operand rows.
// ld_ab (wide-row products only; 0: K): row stride of A and B when the
// product covers only K of their columns (a K split).
struct GemmProduct {
  int M, N, K, ldz;
  int64_t a_off, b_off, z_off;
  const float *rs = nullptr, *cs = nullptr;
  int ld_ab = 0;
};

// Up to kGemmMaxBatch products in one launch on BF16 MFMA with FP32
// accumulation, three products hi*hi + hi*lo + lo*hi each.
void gemm_bf16x3_batched(hipStream_t st, int nb, const GemmProduct *prods,
                         const void *Ah, const void *Al, const void *Bh,
                         const void *Bl, float *Z);
// The same contraction with 256 x 352 workgroup tiles for wide rows (the
// mapped Z = D beta, N = F), row / column scales as gemm_f16x3_batched's.
void gemm_bf16x3_wide(hipStream_t st, int nb, const GemmProduct *prods,
                      const void *Ah, const void *Al, const void *Bh,
                      const void *Bl, float *Z);
void gemm_f16x3_wide(hipStream_t st, int nb, const GemmProduct *prods,
                     const void *Ah, const void *Al, const void *Bh,
                     const void *Bl, float *Z);
// Workgroup tiles of the wide-row products (one per CU at a time).
int gemm_x3_wide_tiles(int nb, const GemmProduct *prods);
// FP64 / FP32 products Z = A B^T on the matrix cores (row stride of A and
// B = K, rows a multiple of 16 bytes; no hi/lo parts).
void dgemm_batched(hipStream_t st, int nb, const GemmProduct *prods, const double *A,
                   const double *B, double *Z);
// FP32 on 256 x 352 tiles for wide rows (K a multiple of 4, 16-byte aligned
// rows of K values; no scales or strides)
void sgemm_wide(hipStream_t st, int nb, const GemmProduct *prods, const float *A,
                const float *B, float *Z);
void sgemm_batched(hipStream_t st, int nb, const GemmProduct *prods, const float *A,
                   const float *B, float *Z);
// Fused epilogues of the FP64 / FP32 matrix-core product Z = A B^T for the
// sparse GP (one per product of a batch).
//  kKernel (normalized_dot_product.cpp:27-87, 115-210): the product is
//   B_i . P_u (or d_i . p_hat_u when normalized); t = product / (|B_i||P_u|)
//   (the product itself when normalized), k = sig2 t^xi into kval (FP64;
//   skipped when null),
//   c = alpha_u t^(xi-1) into cmat (the GEMM's type), both with leading
//   dimension ldk and zero in columns N .. ldk-1, and per row the partial
//   sums of alpha_u t^xi over each 64-column group into
//   part[row * nparts + group]; k = c = 0 where |B_i| or |P_u| < 1e-8.
//  kSquares (sparse_gp.cpp:93-133): part[row * nparts + group] = the sum of
//   the squared products over the group's columns; with tri set, row n of
//   B is zero past column n (a lower-triangular factor), so a tile ends its
//   K loop at its last column.
// nparts >= ceil(N / 64) is the row pitch of part; groups past ceil(N / 64)
// are written as zero.
struct DEpilogue {
  enum Mode { kKernel = 1, kSquares = 2 };
  int mode = kKernel;
  const double *norm2_rows = nullptr, *pnorm = nullptr, *alpha = nullptr;
  double sig2 = 1, power = 2;
  int normalized = 0;
  double *kval = nullptr;
  void *cmat = nullptr;
  int ldk = 0;
  double *part = nullptr;
  int nparts = 0;
  int tri = 0;
};
void dgemm_epi_batched(hipStream_t st, int nb, const GemmProduct *prods,
                       const DEpilogue *ep, const double *A, const double *B);
void sgemm_epi_batched(hipStream_t st, int nb, const GemmProduct *prods,
                       const DEpilogue *ep, const float *A, const float *B);
// The same on FP16 MFMA (operands from split_f16_*).
void gemm_f16x3_batched(hipStream_t st, int nb, const GemmProduct *prods,
                        const void *Ah, const void *Al, const void *Bh,
                        const void *Bl, float *Z);
// FP16 hi/lo split with one power-of-two scale per row (largest element in
// [0.5, 1)); inv_scale[r] = 1 / scale.  x rows of stride ldx, out rows of kp.
void split_f16_rows(hipStream_t st, int rows, int cols, int ldx, int kp,
                    const double *x, void *hi, void *lo, float *inv_scale);
// the same for the transpose of a rows x cols FP64 block (cols output rows
// of ldo values, one scale per output row)
void split_f16_cols(hipStream_t st, int rows, int cols, int ldo, const double *x,
                    void *hi, void *lo, float *inv_scale);

// hi/lo BF16 split of a rows x cols FP64 matrix into rows of kp (zero padded)
void split_bf16(hipStream_t st, int rows, int cols, int kp, const double *x,
                void *hi, void *lo);

}  // namespace flare
