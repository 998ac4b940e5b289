// gp.hip -- NormalizedDotProduct sparse-GP pieces.
//
//  gp_epilogue   k(i,u) = sig2 (B_i.P_u / (|B_i||P_u|))^xi on the raw
//                GEMM output, zero when either norm < 1e-8
//                (normalized_dot_product.cpp:27-87, 115-210); also the
//                adjoint weights c = alpha t^(xi-1) and e = sum alpha t^xi.
//  gp_w          dE_i/dB_i = sig2 xi / |B| (sum_u c_u P_hat_u - e B_hat),
//                the chain rule of envs_struc's force kernel
//                (normalized_dot_product.cpp:176-190) applied to alpha.
//  kuu / chol    K_uu + jitter I (sparse_gp.cpp:581-582), blocked Cholesky
//                panel; the triangular solves/updates go to rocBLAS.
//  pack/unpack   mapping coefficient file convention: packed upper
//                triangle, off-diagonals doubled (normalized_dot_product.cpp
//                :736-753, pair_flare.cpp:378-395).
#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace flare {

__device__ inline double powi(double t, double power) {
  if (power == 2.0) return t * t;
  if (power == 1.0) return t;
  return pow(t, power);
}

__global__ void __launch_bounds__(64)
gp_epilogue_kernel(int rows, int M, int ld, int normalized, const double *__restrict__ K,
                   const double *__restrict__ norm2_rows,
                   const double *__restrict__ pnorm,
                   const double *__restrict__ alpha, double sig2, double power,
                   double *__restrict__ kval, double *__restrict__ cmat,
                   double *__restrict__ e_row) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const double nb = sqrt(norm2_rows[r]);
  const bool empty = nb < 1e-8;
  double e = 0;
  // rows of ld >= M (K read, kval / cmat written); zero past M
  for (int u = lane; u < ld; u += kWave) {
    double kv = 0, cv = 0;
    if (u < M && !empty && !(pnorm[u] < 1e-8)) {
      const double dot = K[(int64_t)r * ld + u];
      const double t = normalized ? dot : dot / (nb * pnorm[u]);
      const double tp = powi(t, power - 1.0);
      kv = sig2 * powi(t, power);
      cv = alpha ? alpha[u] * tp : 0.0;
      if (alpha) e += alpha[u] * powi(t, power);
    }
    if (kval) kval[(int64_t)r * ld + u] = kv;
    if (cmat) cmat[(int64_t)r * ld + u] = cv;
  }
  e = wave_sum(e);
  if (lane == 0 && e_row) e_row[r] = e;
}

// gp_w_kernel for rows of an even F on 16-byte boundaries: one wavefront per
// row, four rows per workgroup, G, B and w moved as double2; the same
// arithmetic per element.
__global__ void __launch_bounds__(256)
gp_w2_kernel(int rows, int F, const double *__restrict__ G,
             const double *__restrict__ e_part, int nparts, const double *__restrict__ B,
             const double *__restrict__ norm2, double sig2, double power,
             double *__restrict__ w, double *__restrict__ E_row) {
  const int lane = threadIdx.x & (kWave - 1);
  const int r = blockIdx.x * 4 + (int)(threadIdx.x / kWave);
  if (r >= rows) return;
  const double nb = sqrt(norm2[r]);
  const bool empty = nb < 1e-8;
  double e = 0;
  for (int q = 0; q < nparts; q++) e += e_part[(int64_t)r * nparts + q];
  const double s = empty ? 0.0 : sig2 * power / nb;
  const int F2 = F / 2;
  const double2 *g2 = reinterpret_cast<const double2 *>(G + (int64_t)r * F);
  const double2 *b2 = reinterpret_cast<const double2 *>(B + (int64_t)r * F);
  double2 *w2 = reinterpret_cast<double2 *>(w + (int64_t)r * F);
  for (int f = lane; f < F2; f += kWave) {
    double2 o = {0.0, 0.0};
    if (!empty) {
      const double2 g = g2[f], x = b2[f];
      o.x = s * (g.x - e * x.x / nb);
      o.y = s * (g.y - e * x.y / nb);
    }
    w2[f] = o;
  }
  if (lane == 0) E_row[r] = empty ? 0.0 : sig2 * e;
}

__global__ void __launch_bounds__(64)
gp_w_kernel(int rows, int F, const double *__restrict__ G,
            const double *__restrict__ e_part, int nparts, const double *__restrict__ B,
            const double *__restrict__ norm2, double sig2, double power,
            double *__restrict__ w, double *__restrict__ E_row) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const double nb = sqrt(norm2[r]);
  const bool empty = nb < 1e-8;
  double e = 0;
  for (int q = 0; q < nparts; q++) e += e_part[(int64_t)r * nparts + q];
  const double s = empty ? 0.0 : sig2 * power / nb;
  for (int f = lane; f < F; f += kWave) {
    double v = 0;
    if (!empty)
      v = s * (G[(int64_t)r * F + f] - e * B[(int64_t)r * F + f] / nb);
    w[(int64_t)r * F + f] = v;
  }
  if (lane == 0) E_row[r] = empty ? 0.0 : sig2 * e;
}

__global__ void __launch_bounds__(64)
row_sqsum_sub_kernel(int rows, int M, const double *__restrict__ Q,
                     const double *__restrict__ kself,
                     double *__restrict__ var) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  double s = 0;
  for (int u = lane; u < M; u += kWave) {
    double q = Q[(int64_t)r * M + u];
    s += q * q;
  }
  s = wave_sum(s);
  if (lane == 0) var[r] = kself[r] - s;
}

// var = k(i, i) - |L^-1 k_i|^2 from per-row partial sums of squares
__global__ void var_from_parts_kernel(int rows, int nparts, const double *__restrict__ part,
                                      const double *__restrict__ kself,
                                      double *__restrict__ var) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  double s = 0;
  for (int q = 0; q < nparts; q++) s += part[(int64_t)r * nparts + q];
  var[r] = kself[r] - s;
}

__global__ void __launch_bounds__(64)
kself_kernel(int rows, int F, const double *__restrict__ B,
             const double *__restrict__ norm2, double sig2, double power,
             double *__restrict__ kself) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const double nb = sqrt(norm2[r]);
  double d = 0;
  for (int f = lane; f < F; f += kWave) {
    double b = B[(int64_t)r * F + f];
    d += b * b;
  }
  d = wave_sum(d);
  if (lane == 0) kself[r] = (nb < 1e-8) ? 0.0 : sig2 * powi(d / (nb * nb), power);
}

__global__ void __launch_bounds__(64)
row_norms_kernel(int rows, int F, const double *__restrict__ X,
                 double *__restrict__ norm, double *__restrict__ xhat) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  double s = 0;
  for (int f = lane; f < F; f += kWave) {
    double x = X[(int64_t)r * F + f];
    s += x * x;
  }
  s = wave_sum(s);
  const double nrm = sqrt(s);
  if (lane == 0 && norm) norm[r] = nrm;
  if (xhat)
    for (int f = lane; f < F; f += kWave)
      xhat[(int64_t)r * F + f] = nrm < 1e-8 ? 0.0 : X[(int64_t)r * F + f] / nrm;
}

// Cholesky of one nb x nb diagonal block (column-major, lower), nb <= 64.
__global__ void __launch_bounds__(256)
chol_panel_kernel(int nb, double *__restrict__ A, int ld, int *info) {
  __shared__ double a[64][65];
  const int t = threadIdx.x;
  for (int q = t; q < nb * nb; q += blockDim.x) {
    int i = q % nb, j = q / nb;
    a[i][j] = A[(int64_t)j * ld + i];
  }
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    if (t == 0) {
      double d = a[j][j];
      if (!(d > 0)) {
        atomicExch(info, 1);
        d = 1.0;
      }
      a[j][j] = sqrt(d);
    }
    __syncthreads();
    for (int i = j + 1 + t; i < nb; i += blockDim.x) a[i][j] /= a[j][j];
    __syncthreads();
    const int m = nb - j - 1;
    for (int q = t; q < m * m; q += blockDim.x) {
      int i = j + 1 + q / m, k = j + 1 + q % m;
      if (k <= i) a[i][k] -= a[i][j] * a[k][j];
    }
    __syncthreads();
  }
  for (int q = t; q < nb * nb; q += blockDim.x) {
    int i = q % nb, j = q / nb;
    A[(int64_t)j * ld + i] = (i >= j) ? a[i][j] : 0.0;
  }
}

__global__ void zero_upper_kernel(int M, double *__restrict__ A, int ld) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)M * M) return;
  int j = (int)(q / M), i = (int)(q - (int64_t)j * M);  // column-major (i, j)
  if (i < j) A[(int64_t)j * ld + i] = 0.0;
}

__global__ void scale_rows_kernel(int rows, int F, const double *__restrict__ X,
                                  const double *__restrict__ s,
                                  double *__restrict__ out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)rows * F) return;
  out[q] = X[q] * s[q / F];
}

// full (F x F symmetric) -> packed file order (one thread per row i)
__global__ void pack_beta_kernel(int F, const double *__restrict__ full,
                                 double *__restrict__ packed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F) return;
  int64_t base = (int64_t)i * F - (int64_t)i * (i - 1) / 2;
  for (int j = i; j < F; j++) {
    double v = full[(int64_t)i * F + j];
    packed[base + (j - i)] = (i == j) ? v : 2 * v;
  }
}

__global__ void unpack_beta_kernel(int F, const double *__restrict__ packed,
                                   double *__restrict__ full) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F) return;
  int64_t base = (int64_t)i * F - (int64_t)i * (i - 1) / 2;
  for (int j = i; j < F; j++) {
    double v = packed[base + (j - i)];
    if (i == j) {
      full[(int64_t)i * F + j] = v;
    } else {
      full[(int64_t)i * F + j] = v / 2;
      full[(int64_t)j * F + i] = v / 2;
    }
  }
}

// Low-precision forms.  K holds t = d_i . p_hat_u directly (FP32, from the
// FP32 or split FP16 product of normalised rows); everything derived from it
// is formed in FP64: k = sig2 t^xi (kept in FP64 for the variance), the
// adjoint weights c = alpha t^(xi-1) (written FP32, or as FP16 hi + lo rows
// of Mp scaled by a power of two that brings the row's largest |c| into
// [0.5, 1), with the inverse scale in c_is) and e = sum alpha t^xi.
__global__ void __launch_bounds__(64)
gp_epilogue_lp_kernel(int rows, int M, int Mp, const float *__restrict__ K,
                      const double *__restrict__ norm2_rows,
                      const double *__restrict__ pnorm,
                      const double *__restrict__ alpha, double sig2,
                      double power, double *__restrict__ kval, int ldk,
                      float *__restrict__ cf, _Float16 *__restrict__ ch,
                      _Float16 *__restrict__ cl, float *__restrict__ c_is,
                      double *__restrict__ e_row) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const bool empty = sqrt(norm2_rows[r]) < 1e-8;
  auto cval = [&](int u, double &kv, double &cv, double &ev) {
    kv = cv = ev = 0;
    if (u < M && !empty && !(pnorm[u] < 1e-8)) {
      const double t = (double)K[(int64_t)r * M + u];
      const double tx = powi(t, power);
      kv = sig2 * tx;
      cv = alpha[u] * powi(t, power - 1.0);
      ev = alpha[u] * tx;
    }
  };
  double e = 0;
  float cmax = 0.0f;
  // kval and cf rows of ldk (>= M), zero past M
  for (int u = lane; u < max(M, ldk); u += kWave) {
    double kv, cv, ev;
    cval(u, kv, cv, ev);
    e += ev;
    if (kval) kval[(int64_t)r * ldk + u] = kv;
    if (cf) cf[(int64_t)r * ldk + u] = (float)cv;
    cmax = fmaxf(cmax, fabsf((float)cv));
  }
  e = wave_sum(e);
  if (lane == 0) e_row[r] = e;
  if (!ch) return;
  for (int off = 32; off > 0; off >>= 1) cmax = fmaxf(cmax, __shfl_xor(cmax, off));
  int ex;
  const bool nonzero = f16_row_scale(cmax, ex);
  const float sc = ldexpf(1.0f, ex);
  for (int u = lane; u < Mp; u += kWave) {
    double kv, cv, ev;
    cval(u, kv, cv, ev);
    const float x = nonzero ? (float)cv * sc : 0.0f;
    const _Float16 h = (_Float16)x;
    ch[(int64_t)r * Mp + u] = h;
    cl[(int64_t)r * Mp + u] = (_Float16)(x - (float)h);
  }
  if (lane == 0) c_is[r] = ldexpf(1.0f, -ex);
}

// gp_epilogue_lp_kernel for every species in one launch, one wavefront per
// row and KR values of the row per lane held in registers (one pass over
// K): the same values, sums in the same order, the same bits.  With K2, t is
// the FP32 sum of the two halves of a K-split product.
struct GpEpiLpBatch {
  int nsp, rows;
  GpEpiLp s[kMaxSpecies];
};

template <int KR>
__global__ void __launch_bounds__(256)
gp_epilogue_lp_rows_kernel(GpEpiLpBatch b, double sig2, double power) {
  const int lane = threadIdx.x & (kWave - 1);
  const int gr = blockIdx.x * 4 + (int)(threadIdx.x / kWave);
  if (gr >= b.rows) return;
  int sp = 0;
  while (sp + 1 < b.nsp && gr >= b.s[sp + 1].row0) sp++;
  const GpEpiLp &q = b.s[sp];
  const int r = gr - q.row0;
  const int M = q.M, U = max(M, q.ldk);
  const bool empty = sqrt(q.norm2_rows[r]) < 1e-8;
  float t[KR], t2[KR];
  double pn[KR], al[KR];
  // every load of the row issued before any is used (the branch on K2 is
  // outside the batch)
  auto load = [&](auto two_c) {
    constexpr bool two = decltype(two_c)::value;
#pragma unroll
    for (int j = 0; j < KR; j++) {
      const int u = lane + kWave * j;
      t[j] = t2[j] = 0.0f;
      pn[j] = al[j] = 0.0;
      if (u < M) {
        t[j] = q.K[(int64_t)r * M + u];
        if (two) t2[j] = q.K2[(int64_t)r * M + u];
        pn[j] = q.pnorm[u];
        al[j] = q.alpha[u];
      }
    }
    if (two)
#pragma unroll
      for (int j = 0; j < KR; j++) t[j] += t2[j];
  };
  if (q.K2)
    load(std::true_type{});
  else
    load(std::false_type{});
  double e = 0;
  float cmax = 0.0f, cvs[KR];
#pragma unroll
  for (int j = 0; j < KR; j++) {
    const int u = lane + kWave * j;
    double kv = 0, cv = 0, ev = 0;
    if (u < M && !empty && !(pn[j] < 1e-8)) {
      const double tt = (double)t[j];
      const double tx = powi(tt, power);
      kv = sig2 * tx;
      cv = al[j] * powi(tt, power - 1.0);
      ev = al[j] * tx;
    }
    cvs[j] = (float)cv;
    if (u < U) {
      e += ev;
      if (q.kval) q.kval[(int64_t)r * q.ldk + u] = kv;
      if (q.cf) q.cf[(int64_t)r * q.ldk + u] = (float)cv;
      cmax = fmaxf(cmax, fabsf((float)cv));
    }
  }
  e = wave_sum(e);
  if (lane == 0) q.e_row[r] = e;
  if (!q.ch) return;
  for (int off = 32; off > 0; off >>= 1) cmax = fmaxf(cmax, __shfl_xor(cmax, off));
  int ex;
  const bool nonzero = f16_row_scale(cmax, ex);
  const float sc = ldexpf(1.0f, ex);
  _Float16 *ch = reinterpret_cast<_Float16 *>(q.ch), *cl = reinterpret_cast<_Float16 *>(q.cl);
#pragma unroll
  for (int j = 0; j < KR; j++) {
    const int u = lane + kWave * j;
    if (u < q.Mp) {
      const float x = nonzero ? cvs[j] * sc : 0.0f;
      const _Float16 h = (_Float16)x;
      ch[(int64_t)r * q.Mp + u] = h;
      cl[(int64_t)r * q.Mp + u] = (_Float16)(x - (float)h);
    }
  }
  if (lane == 0) q.c_is[r] = ldexpf(1.0f, -ex);
}

// w = sig2 xi / |B| (G - e d) with d = B/|B| the FP32 row and G = c P_hat.
__global__ void __launch_bounds__(64)
gp_w_lp_kernel(int rows, int F, const float *__restrict__ G,
               const double *__restrict__ e_part, int nparts, const float *__restrict__ d,
               const double *__restrict__ norm2, double sig2, double power,
               float *__restrict__ w, double *__restrict__ E_row) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const double nb = sqrt(norm2[r]);
  const bool empty = nb < 1e-8;
  double e = 0;
  for (int q = 0; q < nparts; q++) e += e_part[(int64_t)r * nparts + q];
  const double s = empty ? 0.0 : sig2 * power / nb;
  for (int f = lane; f < F; f += kWave) {
    double v = 0;
    if (!empty)
      v = s * ((double)G[(int64_t)r * F + f] - e * (double)d[(int64_t)r * F + f]);
    w[(int64_t)r * F + f] = (float)v;
  }
  if (lane == 0) E_row[r] = empty ? 0.0 : sig2 * e;
}

// k(i, i) = sig2 (d.d)^xi from the FP32 normalised row
__global__ void __launch_bounds__(64)
kself_lp_kernel(int rows, int F, const float *__restrict__ d,
                const double *__restrict__ norm2, double sig2, double power,
                double *__restrict__ kself) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  double s = 0;
  for (int f = lane; f < F; f += kWave) {
    const double x = (double)d[(int64_t)r * F + f];
    s += x * x;
  }
  s = wave_sum(s);
  if (lane == 0) kself[r] = sqrt(norm2[r]) < 1e-8 ? 0.0 : sig2 * powi(s, power);
}

__global__ void to_f32_kernel(int64_t n, const double *__restrict__ x,
                              float *__restrict__ y) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) y[q] = (float)x[q];
}

// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_gp_epilogue(hipStream_t st, int rows, int M, int normalized,
                        const double *K, const double *norm2_rows,
                        const double *pnorm, const double *alpha, double sig2,
                        double power, double *kval, double *cmat,
                        double *e_row, int ld) {
  if (rows == 0) return;
  hipLaunchKernelGGL(gp_epilogue_kernel, dim3(rows), dim3(kWave), 0, st, rows,
                     M, ld > 0 ? ld : M, normalized, K, norm2_rows, pnorm, alpha, sig2,
                     power, kval, cmat, e_row);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_gp_w(hipStream_t st, int rows, int F, const double *G,
                 const double *e_part, int nparts, const double *B, const double *norm2,
                 double sig2, double power, double *w, double *E_row) {
  if (rows == 0) return;
  const bool v2 = F % 2 == 0 && ((uintptr_t)G | (uintptr_t)B | (uintptr_t)w) % 16 == 0;
  if (v2)
    hipLaunchKernelGGL(gp_w2_kernel, dim3((rows + 3) / 4), dim3(4 * kWave), 0, st, rows, F,
                       G, e_part, nparts, B, norm2, sig2, power, w, E_row);
  else
    hipLaunchKernelGGL(gp_w_kernel, dim3(rows), dim3(kWave), 0, st, rows, F, G,
                       e_part, nparts, B, norm2, sig2, power, w, E_row);
  FLARE_HIP_CHECK(hipGetLastError());
}

// ---- compute flare/std/atom (compute_flare_std_atom.cpp:172-212) ---------

__device__ inline double signed_sqrt(double v) {
  return v >= 0.0 ? sqrt(v) : -sqrt(fabs(v));
}

__global__ void __launch_bounds__(64)
std_map_kernel(int rows, int F, const double *__restrict__ B,
               const double *__restrict__ Z, const double *__restrict__ norm2,
               double sig2, const int *__restrict__ atom_of_row,
               double *__restrict__ stds) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  double d = 0;
  for (int f = lane; f < F; f += kWave)
    d += B[(int64_t)r * F + f] * Z[(int64_t)r * F + f];
  d = wave_sum(d);
  if (lane == 0) {
    const double n2 = norm2[r];
    double v = 0;
    if (!(n2 < 1e-8)) {
      v = d / n2;  // compute_energy_and_u, power 2 (lammps_descriptor.cpp:267)
      v /= sig2;
      v = signed_sqrt(v);
    }
    stds[atom_of_row[r]] = v;
  }
}

__global__ void __launch_bounds__(64)
std_kvec_kernel(int rows, int M, double *__restrict__ T,
                const double *__restrict__ norm2, double power) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const double n2 = norm2[r];
  const bool empty = n2 < 1e-8;
  const double nb = sqrt(n2);
  for (int u = lane; u < M; u += kWave) {
    const double t = T[(int64_t)r * M + u];
    T[(int64_t)r * M + u] = empty ? 0.0 : powi(t / nb, power);
  }
}

__global__ void __launch_bounds__(64)
std_linv_kernel(int rows, int M, const double *__restrict__ Q,
                const double *__restrict__ norm2, double sig2,
                const int *__restrict__ atom_of_row, double *__restrict__ stds) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  double s = 0;
  for (int u = lane; u < M; u += kWave) {
    const double q = Q[(int64_t)r * M + u];
    s += q * q;
  }
  s = wave_sum(s);
  if (lane == 0) {
    double v = 0;
    if (!(norm2[r] < 1e-8)) v = signed_sqrt(1.0 - sig2 * s);
    stds[atom_of_row[r]] = v;
  }
}

__global__ void add_diag_rm_kernel(int F, double *__restrict__ out, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F) out[(int64_t)i * F + i] += v;
}

void launch_std_map(hipStream_t st, int rows, int F, const double *B,
                    const double *Z, const double *norm2, double sig2,
                    const int *atom_of_row, double *stds) {
  if (rows == 0) return;
  hipLaunchKernelGGL(std_map_kernel, dim3(rows), dim3(kWave), 0, st, rows, F, B,
                     Z, norm2, sig2, atom_of_row, stds);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_std_kvec(hipStream_t st, int rows, int M, double *T,
                     const double *norm2, double power) {
  if (rows == 0 || M == 0) return;
  hipLaunchKernelGGL(std_kvec_kernel, dim3(rows), dim3(kWave), 0, st, rows, M, T,
                     norm2, power);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_std_linv(hipStream_t st, int rows, int M, const double *Q,
                     const double *norm2, double sig2, const int *atom_of_row,
                     double *stds) {
  if (rows == 0) return;
  hipLaunchKernelGGL(std_linv_kernel, dim3(rows), dim3(kWave), 0, st, rows, M, Q,
                     norm2, sig2, atom_of_row, stds);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_add_diag_rm(hipStream_t st, int F, double *out, double v) {
  if (F == 0) return;
  hipLaunchKernelGGL(add_diag_rm_kernel, dim3((F + 255) / 256), dim3(256), 0, st,
                     F, out, v);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_var_from_parts(hipStream_t st, int rows, int nparts, const double *part,
                           const double *kself, double *var) {
  if (rows == 0) return;
  hipLaunchKernelGGL(var_from_parts_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, rows,
                     nparts, part, kself, var);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_row_sqsum_sub(hipStream_t st, int rows, int M, const double *Q,
                          const double *kself, double *var) {
  if (rows == 0) return;
  hipLaunchKernelGGL(row_sqsum_sub_kernel, dim3(rows), dim3(kWave), 0, st,
                     rows, M, Q, kself, var);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_kself(hipStream_t st, int rows, int F, const double *B,
                  const double *norm2, double sig2, double power,
                  double *kself) {
  if (rows == 0) return;
  hipLaunchKernelGGL(kself_kernel, dim3(rows), dim3(kWave), 0, st, rows, F, B,
                     norm2, sig2, power, kself);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_row_norms(hipStream_t st, int rows, int F, const double *X,
                      double *norm, double *xhat) {
  if (rows == 0) return;
  hipLaunchKernelGGL(row_norms_kernel, dim3(rows), dim3(kWave), 0, st, rows, F,
                     X, norm, xhat);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_chol_panel(hipStream_t st, int nb, double *A, int ld, int *info) {
  FLARE_REQUIRE(nb > 0 && nb <= 64, "panel width must be in [1, 64]");
  hipLaunchKernelGGL(chol_panel_kernel, dim3(1), dim3(256), 0, st, nb, A, ld,
                     info);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_zero_upper(hipStream_t st, int M, double *A, int ld) {
  if (M == 0) return;
  hipLaunchKernelGGL(zero_upper_kernel, dim3(nblk((int64_t)M * M, 256)),
                     dim3(256), 0, st, M, A, ld);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_scale_rows(hipStream_t st, int rows, int F, const double *X,
                       const double *s, double *out) {
  if (rows == 0) return;
  hipLaunchKernelGGL(scale_rows_kernel, dim3(nblk((int64_t)rows * F, 256)),
                     dim3(256), 0, st, rows, F, X, s, out);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_pack_beta(hipStream_t st, int F, int power, const double *full,
                      double *packed) {
  if (power == 1) {
    FLARE_HIP_CHECK(hipMemcpyAsync(packed, full, sizeof(double) * F,
                                   hipMemcpyDeviceToDevice, st));
    return;
  }
  hipLaunchKernelGGL(pack_beta_kernel, dim3(nblk(F, 64)), dim3(64), 0, st, F,
                     full, packed);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_unpack_beta(hipStream_t st, int F, int power, const double *packed,
                        double *full) {
  if (power == 1) {
    FLARE_HIP_CHECK(hipMemcpyAsync(full, packed, sizeof(double) * F,
                                   hipMemcpyDeviceToDevice, st));
    return;
  }
  hipLaunchKernelGGL(unpack_beta_kernel, dim3(nblk(F, 64)), dim3(64), 0, st, F,
                     packed, full);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_gp_epilogue_lp(hipStream_t st, int rows, int M, int Mp,
                           const float *K, const double *norm2_rows,
                           const double *pnorm, const double *alpha, double sig2,
                           double power, double *kval, int ldk, float *cf, void *ch,
                           void *cl, float *c_is, double *e_row) {
  if (rows == 0) return;
  hipLaunchKernelGGL(gp_epilogue_lp_kernel, dim3(rows), dim3(kWave), 0, st, rows,
                     M, Mp, K, norm2_rows, pnorm, alpha, sig2, power, kval, ldk, cf,
                     (_Float16 *)ch, (_Float16 *)cl, c_is, e_row);
  FLARE_HIP_CHECK(hipGetLastError());
}

bool launch_gp_epilogue_lp_rows(hipStream_t st, int nsp, const GpEpiLp *sp, double sig2,
                                double power) {
  FLARE_REQUIRE(nsp >= 0 && nsp <= kMaxSpecies, "gp epilogue: too many species");
  GpEpiLpBatch b;
  b.nsp = nsp;
  b.rows = 0;
  int width = 0;
  for (int q = 0; q < nsp; q++) {
    b.s[q] = sp[q];
    FLARE_REQUIRE(sp[q].row0 == b.rows, "gp epilogue: species rows must be packed");
    b.rows += sp[q].rows;
    width = std::max(width, std::max(sp[q].M, std::max(sp[q].ldk, sp[q].Mp)));
  }
  if (width > 16 * kWave) return false;
  if (b.rows == 0) return true;
  const dim3 grid((b.rows + 3) / 4), block(4 * kWave);
  if (width <= 4 * kWave)
    hipLaunchKernelGGL(gp_epilogue_lp_rows_kernel<4>, grid, block, 0, st, b, sig2, power);
  else if (width <= 8 * kWave)
    hipLaunchKernelGGL(gp_epilogue_lp_rows_kernel<8>, grid, block, 0, st, b, sig2, power);
  else
    hipLaunchKernelGGL(gp_epilogue_lp_rows_kernel<16>, grid, block, 0, st, b, sig2, power);
  FLARE_HIP_CHECK(hipGetLastError());
  return true;
}

// gp_w_lp_kernel for rows of F % 4 == 0 floats on 16-byte boundaries: one
// wavefront per row, four rows per workgroup, G, d and w moved as float4;
// the same arithmetic per element, so the same bits.
__global__ void __launch_bounds__(256)
gp_w_lp4_kernel(int rows, int F, const float *__restrict__ G,
                const double *__restrict__ e_part, int nparts, const float *__restrict__ d,
                const double *__restrict__ norm2, double sig2, double power,
                float *__restrict__ w, double *__restrict__ E_row) {
  const int lane = threadIdx.x & (kWave - 1);
  const int r = blockIdx.x * 4 + (int)(threadIdx.x / kWave);
  if (r >= rows) return;
  const double nb = sqrt(norm2[r]);
  const bool empty = nb < 1e-8;
  double e = 0;
  for (int q = 0; q < nparts; q++) e += e_part[(int64_t)r * nparts + q];
  const double s = empty ? 0.0 : sig2 * power / nb;
  const int F4 = F / 4;
  const float4 *g4 = reinterpret_cast<const float4 *>(G + (int64_t)r * F);
  const float4 *d4 = reinterpret_cast<const float4 *>(d + (int64_t)r * F);
  float4 *w4 = reinterpret_cast<float4 *>(w + (int64_t)r * F);
  for (int f = lane; f < F4; f += kWave) {
    float4 o = {0.0f, 0.0f, 0.0f, 0.0f};
    if (!empty) {
      const float4 g = g4[f], x = d4[f];
      o.x = (float)(s * ((double)g.x - e * (double)x.x));
      o.y = (float)(s * ((double)g.y - e * (double)x.y));
      o.z = (float)(s * ((double)g.z - e * (double)x.z));
      o.w = (float)(s * ((double)g.w - e * (double)x.w));
    }
    w4[f] = o;
  }
  if (lane == 0) E_row[r] = empty ? 0.0 : sig2 * e;
}

void launch_gp_w_lp(hipStream_t st, int rows, int F, const float *G,
                    const double *e_part, int nparts, const float *d, const double *norm2,
                    double sig2, double power, float *w, double *E_row) {
  if (rows == 0) return;
  const bool v4 = F % 4 == 0 && ((uintptr_t)G | (uintptr_t)d | (uintptr_t)w) % 16 == 0;
  if (v4)
    hipLaunchKernelGGL(gp_w_lp4_kernel, dim3((rows + 3) / 4), dim3(4 * kWave), 0, st, rows,
                       F, G, e_part, nparts, d, norm2, sig2, power, w, E_row);
  else
    hipLaunchKernelGGL(gp_w_lp_kernel, dim3(rows), dim3(kWave), 0, st, rows, F, G,
                       e_part, nparts, d, norm2, sig2, power, w, E_row);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_kself_lp(hipStream_t st, int rows, int F, const float *d,
                     const double *norm2, double sig2, double power,
                     double *kself) {
  if (rows == 0) return;
  hipLaunchKernelGGL(kself_lp_kernel, dim3(rows), dim3(kWave), 0, st, rows, F, d,
                     norm2, sig2, power, kself);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_to_f32(hipStream_t st, int64_t n, const double *x, float *y) {
  if (n == 0) return;
  hipLaunchKernelGGL(to_f32_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, x,
                     y);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
