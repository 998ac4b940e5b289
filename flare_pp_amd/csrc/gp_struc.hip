// gp_struc.hip -- kernel matrices of a structure's outputs and the SOR / DTC
// variances of SparseGP (sparse_gp.cpp:630-680).
//
//  envs_struc     k(sparse u, output o) for o = [E, F (3n), S (6)]
//                 (normalized_dot_product.cpp:115-210): after the GEMMs
//                 DOT = P D^T and FD = P dB^T, one workgroup per sparse row
//                 forms the per-pair values, the energy and stress entries;
//                 the force entries are gathered per atom (own pairs minus
//                 the pairs that have the atom as neighbor), so no float
//                 atomics are used.
//  self kernel    k(o, o) (normalized_dot_product.cpp:446-578).  Energy and
//                 stress sum over ordered atom pairs of one species; the
//                 stress sums over neighbor pairs factor through
//                 G_i^{mn} = sum_k dB_{k,m} r_{k,n} and g_i^{mn} =
//                 sum_k (B_i . dB_{k,m}) r_{k,n}, leaving n_s x n_s GEMMs.
//                 Force entries: one workgroup per atom a over the pairs
//                 that touch a (as centre or neighbor), sign + when a has
//                 the same role in both pairs.
#include "common.h"
#include "kernels.h"

namespace flare {

namespace {

__device__ inline double powr(double t, double p) {
  if (p == 2.0) return t * t;
  if (p == 1.0) return t;
  if (p == 0.0) return 1.0;
  return pow(t, p);
}

// deterministic block sum over 256 threads (fixed tree)
template <int NV>
__device__ inline void block_sum256(double (&v)[NV], double *sh) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double x = wave_sum(v[q]);
    if (lane == 0) sh[w * NV + q] = x;
  }
  __syncthreads();
  if (t < NV) {
    double s = 0;
    for (int k = 0; k < 4; k++) s += sh[k * NV + t];
    sh[4 * NV + t] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; q++) v[q] = sh[4 * NV + q];
  __syncthreads();
}

}  // namespace

constexpr double kEmpty = 1e-8;

// One workgroup per sparse row u of one species.  DOT: Ms x ns, FD (in,
// overwritten with the per-pair values): Ms x 3Q.  pair_row: Q rows (within
// the species).  K: output row stride n_out.
__global__ void __launch_bounds__(256)
envs_struc_epilogue_kernel(int Ms, int ns, int64_t Q, const double *__restrict__ DOT,
                           double *__restrict__ FD, const double *__restrict__ pnorm,
                           const double *__restrict__ rnorm2,
                           const double *__restrict__ fdots,
                           const double *__restrict__ coords,
                           const int *__restrict__ pair_row, double sig2,
                           double power, double vol_inv, double *__restrict__ K,
                           int64_t n_out, int n_atoms) {
  __shared__ double sh[5 * 7];
  const int u = blockIdx.x;
  const double norm_i = pnorm[u];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const bool skip = norm_i < kEmpty;
  if (!skip) {
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
      const double norm_j = sqrt(rnorm2[j]);
      if (norm_j < kEmpty) continue;
      acc[0] += sig2 * powr(DOT[(int64_t)u * ns + j] / (norm_i * norm_j), power);
    }
  }
  for (int64_t e = threadIdx.x; e < 3 * Q; e += blockDim.x) {
    const int64_t q = e / 3;
    const int c = (int)(e - 3 * q);
    const int j = pair_row[q];
    const double norm_j = sqrt(rnorm2[j]);
    double val = 0;
    if (!skip && !(norm_j < kEmpty)) {
      const double norm_ij = norm_i * norm_j;
      const double norm_ij3 = norm_ij * norm_j * norm_j;
      const double dot_val = DOT[(int64_t)u * ns + j];
      const double dval = power * powr(dot_val / norm_ij, power - 1);
      const double f1 = FD[(int64_t)u * 3 * Q + e] / norm_ij;
      const double f2 = dot_val * fdots[e] / norm_ij3;
      val = sig2 * dval * (f1 - f2);
      int sc = c == 0 ? 1 : (c == 1 ? 4 : 6);
      for (int c2 = c; c2 < 3; c2++, sc++)
        acc[sc] -= val * coords[3 * q + c2] * vol_inv;
    }
    FD[(int64_t)u * 3 * Q + e] = val;
  }
  block_sum256<7>(acc, sh);
  if (threadIdx.x == 0 && !skip) {
    K[(int64_t)u * n_out] = acc[0];
    for (int q = 0; q < 6; q++) K[(int64_t)u * n_out + 1 + 3 * n_atoms + q] = acc[1 + q];
  }
}

// K[u][1 + 3a + c] = sum over own pairs of a - sum over pairs whose neighbor
// is a, of the per-pair values of one species.  own: atom -> [own0, own0 +
// ownc) (pairs within the species, -1 count for atoms of other species).
__global__ void __launch_bounds__(256)
envs_struc_gather_kernel(int Ms, int n_atoms, int64_t Q, const double *__restrict__ VAL,
                         const int *__restrict__ own0, const int *__restrict__ ownc,
                         const int *__restrict__ in_cum,
                         const int *__restrict__ in_pairs, double *__restrict__ K,
                         int64_t n_out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)Ms * n_atoms) return;
  const int u = (int)(t / n_atoms), a = (int)(t - (int64_t)u * n_atoms);
  const double *v = VAL + (int64_t)u * 3 * Q;
  double f[3] = {0, 0, 0};
  for (int q = own0[a], e = own0[a] + ownc[a]; q < e; q++)
    for (int c = 0; c < 3; c++) f[c] += v[3 * (int64_t)q + c];
  for (int r = in_cum[a]; r < in_cum[a + 1]; r++) {
    const int64_t q = in_pairs[r];
    for (int c = 0; c < 3; c++) f[c] -= v[3 * q + c];
  }
  for (int c = 0; c < 3; c++) K[(int64_t)u * n_out + 1 + 3 * a + c] += f[c];
}

// G_i^{mn} (ns x F, combo q = (m, n >= m)) and g_i^{mn} over the pairs of
// row i: one workgroup per row.
__global__ void __launch_bounds__(256)
self_g_kernel(int ns, int F, int m, int nn, const int64_t *__restrict__ q0,
              const int *__restrict__ cnt, const double *__restrict__ dB,
              const double *__restrict__ fdots, const double *__restrict__ coords,
              double *__restrict__ G, double *__restrict__ g) {
  const int i = blockIdx.x;
  const int64_t b = q0[i], e = b + cnt[i];
  for (int f = threadIdx.x; f < F; f += blockDim.x) {
    double s = 0;
    for (int64_t q = b; q < e; q++) s += dB[(3 * q + m) * F + f] * coords[3 * q + nn];
    G[(int64_t)i * F + f] = s;
  }
  if (threadIdx.x == 0) {
    double s = 0;
    for (int64_t q = b; q < e; q++) s += fdots[3 * q + m] * coords[3 * q + nn];
    g[i] = s;
  }
}

// Per row i of one species: sum over j of the energy term (combo < 0) or of
// the stress term of one (m, n) combo from T = B B^T, U = G B^T, V = G G^T.
__global__ void __launch_bounds__(256)
self_rows_kernel(int ns, const double *__restrict__ T, const double *__restrict__ U,
                 const double *__restrict__ V, const double *__restrict__ g,
                 const double *__restrict__ rnorm2, double sig2, double power,
                 int energy, double *__restrict__ row_out) {
  __shared__ double sh[5];
  const int i = blockIdx.x;
  const double ni = sqrt(rnorm2[i]);
  double acc[1] = {0};
  if (!(ni < kEmpty)) {
    const double ni2 = ni * ni, ni3 = ni2 * ni;
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
      const double nj = sqrt(rnorm2[j]);
      if (nj < kEmpty) continue;
      const double nij = ni * nj;
      const double t = T[(int64_t)i * ns + j] / nij;
      if (energy) {
        acc[0] += sig2 * powr(t, power);
        continue;
      }
      const double nj2 = nj * nj, nj3 = nj2 * nj;
      const double c1 = (power - 1) * power * powr(t, power - 2);
      const double c2 = power * powr(t, power - 1);
      const double Uij = U[(int64_t)i * ns + j], Uji = U[(int64_t)j * ns + i];
      const double gi = g[i], gj = g[j];
      const double v1 = Uij / nij - t * gi / ni2;
      const double v2 = Uji / nij - t * gj / nj2;
      const double v3 = V[(int64_t)i * ns + j] / nij;
      const double v4 = gi * Uji / (ni3 * nj);
      const double v5 = gj * Uij / (ni * nj3);
      const double v6 = gi * gj * t / (ni2 * nj2);
      acc[0] += sig2 * (c1 * v1 * v2 + c2 * (v3 - v4 - v5 + v6));
    }
  }
  block_sum256<1>(acc, sh);
  if (threadIdx.x == 0) row_out[i] = acc[0];
}

// Force entries of the self kernel: one workgroup per atom a.  Entries are
// the pairs with centre a (role 0) and with neighbor a (role 1); pairs of
// centres of the same species interact.  D: all rows (n x F, grouped), dB:
// 3Q x F, pair_rowg: global row of each pair's centre, row_species.
constexpr int kSelfMaxEntries = kSelfMaxPairs;

__global__ void __launch_bounds__(256)
self_force_kernel(int n_atoms, int F, const int *__restrict__ row_of,
                  const int64_t *__restrict__ q0, const int *__restrict__ cnt,
                  const int *__restrict__ in_cum, const int *__restrict__ in_pairs,
                  const int *__restrict__ pair_rowg, const int *__restrict__ row_species,
                  const double *__restrict__ D, const double *__restrict__ rnorm2,
                  const double *__restrict__ dB, const double *__restrict__ fdots,
                  double sig2, double power, double *__restrict__ out,
                  int *__restrict__ overflow) {
  __shared__ int eq[kSelfMaxEntries];
  __shared__ signed char erole[kSelfMaxEntries];
  __shared__ double sh[5 * 3];
  const int a = blockIdx.x;
  const int ra = row_of[a];
  const int64_t ob = q0[ra];
  const int no = cnt[ra];
  const int ni_ = in_cum[a + 1] - in_cum[a];
  int E = no + ni_;
  if (E > kSelfMaxEntries) {
    if (threadIdx.x == 0) atomicMax(overflow, E);
    E = 0;
  }
  for (int k = threadIdx.x; k < E; k += blockDim.x) {
    if (k < no) {
      eq[k] = (int)(ob + k);
      erole[k] = 0;
    } else {
      eq[k] = in_pairs[in_cum[a] + (k - no)];
      erole[k] = 1;
    }
  }
  __syncthreads();
  double acc[3] = {0, 0, 0};
  const int64_t total = (int64_t)E * E * 3;
  for (int64_t w = threadIdx.x; w < total; w += blockDim.x) {
    const int m = (int)(w % 3);
    const int64_t pr = w / 3;
    const int e1 = (int)(pr / E), e2 = (int)(pr - (int64_t)e1 * E);
    const int64_t k = eq[e1], l = eq[e2];
    const int i = pair_rowg[k], j = pair_rowg[l];
    if (row_species[i] != row_species[j]) continue;
    const double ni = sqrt(rnorm2[i]), nj = sqrt(rnorm2[j]);
    if (ni < kEmpty || nj < kEmpty) continue;
    const double *Bi = D + (int64_t)i * F, *Bj = D + (int64_t)j * F;
    const double *dk = dB + (3 * k + m) * F, *dl = dB + (3 * l + m) * F;
    double bij = 0, fdj = 0, fdi = 0, ff = 0;
    for (int f = 0; f < F; f++) {
      bij += Bi[f] * Bj[f];
      fdj += dk[f] * Bj[f];
      fdi += dl[f] * Bi[f];
      ff += dk[f] * dl[f];
    }
    const double nij = ni * nj, ni2 = ni * ni, nj2 = nj * nj;
    const double ni3 = ni2 * ni, nj3 = nj2 * nj;
    const double t = bij / nij;
    const double c1 = (power - 1) * power * powr(t, power - 2);
    const double c2 = power * powr(t, power - 1);
    const double sfd1 = fdots[3 * k + m], sfd2 = fdots[3 * l + m];
    const double v1 = fdj / nij - t * sfd1 / ni2;
    const double v2 = fdi / nij - t * sfd2 / nj2;
    const double v3 = ff / nij;
    const double v4 = sfd1 * fdi / (ni3 * nj);
    const double v5 = sfd2 * fdj / (ni * nj3);
    const double v6 = sfd1 * sfd2 * t / (ni2 * nj2);
    const double kv = sig2 * (c1 * v1 * v2 + c2 * (v3 - v4 - v5 + v6));
    acc[m] += erole[e1] == erole[e2] ? kv : -kv;
  }
  block_sum256<3>(acc, sh);
  if (threadIdx.x == 0)
    for (int m = 0; m < 3; m++) out[1 + 3 * a + m] = acc[m];
}

// out[c] (+)= sum_r X[r][c]^2 (sign -1 subtracts); one thread per column,
// rows in order
__global__ void col_sqsum_kernel(int rows, int64_t cols, const double *__restrict__ X,
                                 double sign, int accumulate, double *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double s = 0;
  for (int r = 0; r < rows; r++) {
    const double x = X[(int64_t)r * cols + c];
    s += x * x;
  }
  out[c] = accumulate ? out[c] + sign * s : sign * s;
}

// ---- launchers ----------------------------------------------------------------

void launch_envs_struc_epilogue(hipStream_t st, int Ms, int ns, int64_t Q,
                                const double *DOT, double *FD, const double *pnorm,
                                const double *rnorm2, const double *fdots,
                                const double *coords, const int *pair_row,
                                double sig2, double power, double vol_inv, double *K,
                                int64_t n_out, int n_atoms) {
  if (Ms == 0) return;
  hipLaunchKernelGGL(envs_struc_epilogue_kernel, dim3(Ms), dim3(256), 0, st, Ms, ns,
                     Q, DOT, FD, pnorm, rnorm2, fdots, coords, pair_row, sig2, power,
                     vol_inv, K, n_out, n_atoms);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_envs_struc_gather(hipStream_t st, int Ms, int n_atoms, int64_t Q,
                              const double *VAL, const int *own0, const int *ownc,
                              const int *in_cum, const int *in_pairs, double *K,
                              int64_t n_out) {
  const int64_t tot = (int64_t)Ms * n_atoms;
  if (tot == 0) return;
  hipLaunchKernelGGL(envs_struc_gather_kernel, dim3((unsigned)((tot + 255) / 256)),
                     dim3(256), 0, st, Ms, n_atoms, Q, VAL, own0, ownc, in_cum,
                     in_pairs, K, n_out);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_self_g(hipStream_t st, int ns, int F, int m, int nn, const int64_t *q0,
                   const int *cnt, const double *dB, const double *fdots,
                   const double *coords, double *G, double *g) {
  if (ns == 0) return;
  hipLaunchKernelGGL(self_g_kernel, dim3(ns), dim3(256), 0, st, ns, F, m, nn, q0, cnt,
                     dB, fdots, coords, G, g);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_self_rows(hipStream_t st, int ns, const double *T, const double *U,
                      const double *V, const double *g, const double *rnorm2,
                      double sig2, double power, int energy, double *row_out) {
  if (ns == 0) return;
  hipLaunchKernelGGL(self_rows_kernel, dim3(ns), dim3(256), 0, st, ns, T, U, V, g,
                     rnorm2, sig2, power, energy, row_out);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_self_force(hipStream_t st, int n_atoms, int F, const int *row_of,
                       const int64_t *q0, const int *cnt, const int *in_cum,
                       const int *in_pairs, const int *pair_rowg,
                       const int *row_species, const double *D, const double *rnorm2,
                       const double *dB, const double *fdots, double sig2,
                       double power, double *out, int *overflow) {
  if (n_atoms == 0) return;
  hipLaunchKernelGGL(self_force_kernel, dim3(n_atoms), dim3(256), 0, st, n_atoms, F,
                     row_of, q0, cnt, in_cum, in_pairs, pair_rowg, row_species, D,
                     rnorm2, dB, fdots, sig2, power, out, overflow);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_col_sqsum(hipStream_t st, int rows, int64_t cols, const double *X,
                      double sign, int accumulate, double *out) {
  if (cols == 0) return;
  hipLaunchKernelGGL(col_sqsum_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256),
                     0, st, rows, cols, X, sign, accumulate, out);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
