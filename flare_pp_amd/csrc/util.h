// util.h -- launchers of the bookkeeping kernels in util.hip.
#pragma once

#include "common.h"

namespace flare {

void launch_sqrt(hipStream_t st, int n, const double *x, double *y);
// dst[r * ldd + c] = src[r * srs + c * scs] for c < cols, zero up to ldd
void launch_pad_block(hipStream_t st, int rows, int cols, const double *src, int64_t srs,
                      int64_t scs, double *dst, int ldd);
void launch_pad_block_f32(hipStream_t st, int rows, int cols, const double *src,
                          int64_t srs, int64_t scs, float *dst, int ldd);
// the (c1, c2, l) table of the B2 features, as flare_ctx::b2_index uploads it
void launch_b2_index(hipStream_t st, int nr, int L, int *code);
void launch_rows_to_atoms(hipStream_t st, int n, const int *atom_of_row,
                          const double *x, double *y);
void launch_b2_filter_count(hipStream_t st, const B2Dev &p, const NListDev &nl,
                            const int *species, int *count);
void launch_b2_fill_fpair(hipStream_t st, const B2Dev &p, const NListDev &nl,
                          const int *species, const int *row_of,
                          const int64_t *q0, int64_t *fpair);
void launch_normalize_rows_f32(hipStream_t st, int rows, int F, const float *X,
                               const double *norm2, float *out);
void launch_normalize_rows_f64(hipStream_t st, int rows, int F, const double *X,
                               const double *norm2, double *out);
void launch_to_f64(hipStream_t st, int64_t n, const float *x, double *y);
// y (rows x F) = hi + lo of BF16 rows with stride kp
void launch_bf16_pair_to_f64(hipStream_t st, int rows, int F, int kp,
                             const void *hi, const void *lo, double *y);
// zero rows whose |B|^2 is below thresh
void launch_mask_rows(hipStream_t st, int rows, int F, const double *norm2,
                      double thresh, double *X);
// in-place kernel epilogue of one K_uu block (leading dimension ld)
void launch_kuu_block(hipStream_t st, int Ms, int ld, double *K,
                      const double *pnorm, double sig2, double power);
void launch_add_diag(hipStream_t st, int M, double *A, int ld, double v);
void launch_square(hipStream_t st, int n, const double *x, double *y);
void launch_copy_block(hipStream_t st, int r, int m, const double *src,
                       int lds, double *dst, int ldd);
void launch_scale(hipStream_t st, int n, double a, double *x);
// b[j][i] = a[i][j], M x M
void launch_transpose(hipStream_t st, int M, const double *a, double *b);

}  // namespace flare
