// util.hip -- small elementwise and bookkeeping kernels of the host layer.
#include "common.h"
#include "kernels.h"
#include "util.h"

namespace flare {

static inline unsigned nblk(int64_t n, int b) {
  return (unsigned)((n + b - 1) / b);
}

__global__ void sqrt_kernel(int n, const double *__restrict__ x,
                            double *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = sqrt(x[i]);
}

__global__ void rows_to_atoms_kernel(int n, const int *__restrict__ atom_of_row,
                                     const double *__restrict__ x,
                                     double *__restrict__ y) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) y[atom_of_row[r]] = x[r];
}

// pairs of each atom inside the per-species B2 cutoff (b2.cpp:264-281)
__global__ void b2_filter_count_kernel(B2Dev p, NListDev nl,
                                       const int *__restrict__ species,
                                       int *__restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl.n) return;
  int si = species[i], c = 0;
  for (int k = nl.cum[i]; k < nl.cum[i + 1]; k++)
    c += pair_inside(p, si, nl.nspec[k], nl.rel + 4 * (int64_t)k) ? 1 : 0;
  count[i] = c;
}

// structure-list index of every B2 pair, species-grouped row order;
// q0[row] = first slot of that row
__global__ void b2_fill_fpair_kernel(B2Dev p, NListDev nl,
                                     const int *__restrict__ species,
                                     const int *__restrict__ row_of,
                                     const int64_t *__restrict__ q0,
                                     int64_t *__restrict__ fpair) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl.n) return;
  int si = species[i];
  int64_t q = q0[row_of[i]];
  for (int k = nl.cum[i]; k < nl.cum[i + 1]; k++)
    if (pair_inside(p, si, nl.nspec[k], nl.rel + 4 * (int64_t)k)) fpair[q++] = k;
}

__global__ void to_f64_kernel(int64_t n, const float *__restrict__ x,
                              double *__restrict__ y) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) y[q] = (double)x[q];
}

__global__ void bf16_pair_to_f64_kernel(int rows, int F, int kp,
                                        const __bf16 *__restrict__ hi,
                                        const __bf16 *__restrict__ lo,
                                        double *__restrict__ y) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)rows * F) return;
  int64_t r = q / F, c = q - r * F;
  y[q] = (double)((float)hi[r * kp + c] + (float)lo[r * kp + c]);
}

__global__ void mask_rows_kernel(int rows, int F, const double *__restrict__ norm2,
                                 double thresh, double *__restrict__ X) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)rows * F) return;
  if (norm2[q / F] < thresh) X[q] = 0.0;
}

// K_uu block epilogue: t = dot / (|p_u||p_v|), sig2 t^xi, zero for empty
// environments (normalized_dot_product.cpp:27-87)
__global__ void kuu_block_kernel(int Ms, int ld, double *__restrict__ K,
                                 const double *__restrict__ pnorm, double sig2,
                                 double power) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)Ms * Ms) return;
  int u = (int)(q / Ms), v = (int)(q - (int64_t)u * Ms);
  double nu = pnorm[u], nv = pnorm[v];
  double *e = K + (int64_t)u * ld + v;
  double k = 0;
  if (!(nu < 1e-8) && !(nv < 1e-8)) {
    double t = *e / (nu * nv);
    k = sig2 * (power == 2.0 ? t * t : (power == 1.0 ? t : pow(t, power)));
  }
  *e = k;
}

__global__ void add_diag_kernel(int M, double *__restrict__ A, int ld, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) A[(int64_t)i * ld + i] += v;
}

__global__ void square_kernel(int n, const double *__restrict__ x,
                              double *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] * x[i];
}

__global__ void copy_block_kernel(int r, int m, const double *__restrict__ src,
                                  int lds, double *__restrict__ dst, int ldd) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)r * m) return;
  int i = (int)(q / m), j = (int)(q - (int64_t)i * m);
  dst[(int64_t)i * ldd + j] = src[(int64_t)i * lds + j];
}

__global__ void scale_kernel(int n, double a, double *__restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

template <typename TI, typename TO>
__global__ void normalize_rows_kernel(int rows, int F, const TI *__restrict__ X,
                                      const double *__restrict__ norm2,
                                      TO *__restrict__ out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)rows * F) return;
  int r = (int)(q / F);
  double n2 = norm2[r];
  double s = n2 > 0 ? 1.0 / sqrt(n2) : 0.0;
  out[q] = (TO)((double)X[q] * s);
}

__global__ void transpose_kernel(int M, const double *__restrict__ a,
                                 double *__restrict__ b) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)M * M) return;
  int i = (int)(q / M), j = (int)(q - (int64_t)i * M);
  b[(int64_t)j * M + i] = a[q];
}

// B2 feature table: entry k = (c1 << 20) | (c2 << 8) | l for the packed
// pair p(c1 <= c2) and l of feature k = p * (L + 1) + l (b2.cpp:209-233)
__global__ void b2_index_kernel(int nr, int L, int *__restrict__ code) {
  const int F = nr * (nr + 1) / 2 * (L + 1);
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= F) return;
  int l = k % (L + 1), p = k / (L + 1), c1 = 0;
  while (p >= nr - c1) {  // row c1 of the packed upper triangle holds nr - c1 pairs
    p -= nr - c1;
    c1++;
  }
  code[k] = (c1 << 20) | ((c1 + p) << 8) | l;
}

// dst[r * ldd + c] = src[r * srs + c * scs] for c < cols, 0 for cols <= c <
// ldd: strided block copies with zero-padded rows (transposes included)
template <typename D>
__global__ void pad_block_kernel(int rows, int cols, const double *__restrict__ src,
                                 int64_t srs, int64_t scs, D *__restrict__ dst, int ldd) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)rows * ldd) return;
  const int r = (int)(q / ldd), c = (int)(q - (int64_t)r * ldd);
  dst[q] = c < cols ? (D)src[r * srs + c * scs] : (D)0;
}

void launch_sqrt(hipStream_t st, int n, const double *x, double *y) {
  if (n == 0) return;
  hipLaunchKernelGGL(sqrt_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, x, y);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_b2_index(hipStream_t st, int nr, int L, int *code) {
  const int F = nr * (nr + 1) / 2 * (L + 1);
  hipLaunchKernelGGL(b2_index_kernel, dim3(nblk(F, 256)), dim3(256), 0, st, nr, L, code);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_pad_block(hipStream_t st, int rows, int cols, const double *src, int64_t srs,
                      int64_t scs, double *dst, int ldd) {
  if (rows == 0 || ldd == 0) return;
  hipLaunchKernelGGL(pad_block_kernel<double>, dim3(nblk((int64_t)rows * ldd, 256)), dim3(256),
                     0, st, rows, cols, src, srs, scs, dst, ldd);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_pad_block_f32(hipStream_t st, int rows, int cols, const double *src,
                          int64_t srs, int64_t scs, float *dst, int ldd) {
  if (rows == 0 || ldd == 0) return;
  hipLaunchKernelGGL(pad_block_kernel<float>, dim3(nblk((int64_t)rows * ldd, 256)), dim3(256),
                     0, st, rows, cols, src, srs, scs, dst, ldd);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_rows_to_atoms(hipStream_t st, int n, const int *atom_of_row,
                          const double *x, double *y) {
  if (n == 0) return;
  hipLaunchKernelGGL(rows_to_atoms_kernel, dim3(nblk(n, 256)), dim3(256), 0, st,
                     n, atom_of_row, x, y);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_b2_filter_count(hipStream_t st, const B2Dev &p, const NListDev &nl,
                            const int *species, int *count) {
  if (nl.n == 0) return;
  hipLaunchKernelGGL(b2_filter_count_kernel, dim3(nblk(nl.n, 256)), dim3(256),
                     0, st, p, nl, species, count);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_b2_fill_fpair(hipStream_t st, const B2Dev &p, const NListDev &nl,
                          const int *species, const int *row_of,
                          const int64_t *q0, int64_t *fpair) {
  if (nl.n == 0) return;
  hipLaunchKernelGGL(b2_fill_fpair_kernel, dim3(nblk(nl.n, 256)), dim3(256), 0,
                     st, p, nl, species, row_of, q0, fpair);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_normalize_rows_f32(hipStream_t st, int rows, int F, const float *X,
                               const double *norm2, float *out) {
  if (rows == 0) return;
  hipLaunchKernelGGL((normalize_rows_kernel<float, float>),
                     dim3(nblk((int64_t)rows * F, 256)), dim3(256), 0, st, rows,
                     F, X, norm2, out);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_normalize_rows_f64(hipStream_t st, int rows, int F, const double *X,
                               const double *norm2, double *out) {
  if (rows == 0) return;
  hipLaunchKernelGGL((normalize_rows_kernel<double, double>),
                     dim3(nblk((int64_t)rows * F, 256)), dim3(256), 0, st, rows,
                     F, X, norm2, out);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_to_f64(hipStream_t st, int64_t n, const float *x, double *y) {
  if (n == 0) return;
  hipLaunchKernelGGL(to_f64_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, x, y);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_bf16_pair_to_f64(hipStream_t st, int rows, int F, int kp,
                             const void *hi, const void *lo, double *y) {
  if (rows == 0) return;
  hipLaunchKernelGGL(bf16_pair_to_f64_kernel, dim3(nblk((int64_t)rows * F, 256)),
                     dim3(256), 0, st, rows, F, kp, (const __bf16 *)hi,
                     (const __bf16 *)lo, y);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_mask_rows(hipStream_t st, int rows, int F, const double *norm2,
                      double thresh, double *X) {
  if (rows == 0) return;
  hipLaunchKernelGGL(mask_rows_kernel, dim3(nblk((int64_t)rows * F, 256)),
                     dim3(256), 0, st, rows, F, norm2, thresh, X);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_kuu_block(hipStream_t st, int Ms, int ld, double *K,
                      const double *pnorm, double sig2, double power) {
  if (Ms == 0) return;
  hipLaunchKernelGGL(kuu_block_kernel, dim3(nblk((int64_t)Ms * Ms, 256)),
                     dim3(256), 0, st, Ms, ld, K, pnorm, sig2, power);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_add_diag(hipStream_t st, int M, double *A, int ld, double v) {
  if (M == 0) return;
  hipLaunchKernelGGL(add_diag_kernel, dim3(nblk(M, 256)), dim3(256), 0, st, M,
                     A, ld, v);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_square(hipStream_t st, int n, const double *x, double *y) {
  if (n == 0) return;
  hipLaunchKernelGGL(square_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, x, y);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_copy_block(hipStream_t st, int r, int m, const double *src,
                       int lds, double *dst, int ldd) {
  if (r == 0 || m == 0) return;
  hipLaunchKernelGGL(copy_block_kernel, dim3(nblk((int64_t)r * m, 256)),
                     dim3(256), 0, st, r, m, src, lds, dst, ldd);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_scale(hipStream_t st, int n, double a, double *x) {
  if (n == 0) return;
  hipLaunchKernelGGL(scale_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, a, x);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_transpose(hipStream_t st, int M, const double *a, double *b) {
  if (M == 0) return;
  hipLaunchKernelGGL(transpose_kernel, dim3(nblk((int64_t)M * M, 256)),
                     dim3(256), 0, st, M, a, b);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
