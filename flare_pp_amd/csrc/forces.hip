// forces.hip -- energy, u = dE/dA and pair forces.
//
//  force_kernel      per atom: E_i and w = dE_i/dB_i of the mapped
//                    potential (lammps_descriptor.cpp:250-304; power 2
//                    E = B.beta B/|B|^2, w = 2(beta B - E B)/|B|^2, power 1
//                    E = B.beta/|B|, w = beta/|B| - E B/|B|^2) unless w is
//                    given, then u[c,lm] = sum_c' w[p(c,c'),l] A[c',lm] (1+d_cc')
//                    (lammps_descriptor.cpp:272-303), then for every pair
//                    f_ik = sum_{n,lm} u[s_k N + n, lm] dA[n,lm]/dr_ik.
//                    plain:   dA/dr read from the materialised tensor
//                             (the pair_flare / Kokkos formulation);
//                    adjoint: the pair basis is recomputed and the
//                             contraction factorised as
//                             f = r_hat sum_n g_n' v_n + sum_lm q_lm grad Y_lm,
//                             v_n = sum_lm u Y, q_lm = sum_n u g_n.
//  gather_kernel     F_a = sum_k (f_ak - f_rev(ak)): the scatter
//                    F_i += f, F_j -= f of pair_flare.cpp:151-156 as a
//                    deterministic gather over the full list.
#include "common.h"
#include "kernels.h"

namespace flare {

template <typename T>
struct WSource {
  int normalized;
  const T *X, *Z, *beta1, *w;
  const double *norm2;
  double *E_row;
};

namespace {

// Adjoint contraction for one pair, everything recomputed in registers.
template <typename T, int NMAX>
__device__ inline void adjoint_pair(const B2Dev &p, T r, T x, T y, T z,
                                    T rcut, const T *__restrict__ ub,
                                    const T *__restrict__ ynorm, T &f0, T &f1,
                                    T &f2) {
  const int N = p.N, H = p.H;
  const T rinv = T(1) / r;
  const T ux = x * rinv, uy = y * rinv, uz = z * rinv;
  T g[NMAX], dg[NMAX], v[NMAX];
  radial_basis<T, NMAX>(r, rcut, (T)p.r1, N, p.cutoff_kind, g, dg);
#pragma unroll
  for (int n = 0; n < NMAX; n++) v[n] = 0;
  T a0 = 0, a1 = 0, a2 = 0;
  ylm_visit<T, true>(ux, uy, uz, rinv, p.L, ynorm,
                     [&](int lm, T Y, T Yx, T Yy, T Yz) {
                       T q = 0;
#pragma unroll
                       for (int n = 0; n < NMAX; n++) {
                         if (n < N) {
                           T un = ub[n * H + lm];
                           q += un * g[n];
                           v[n] += un * Y;
                         }
                       }
                       a0 += q * Yx;
                       a1 += q * Yy;
                       a2 += q * Yz;
                     });
  T fr = 0;
#pragma unroll
  for (int n = 0; n < NMAX; n++)
    if (n < N) fr += v[n] * dg[n];
  f0 = a0 + fr * ux;
  f1 = a1 + fr * uy;
  f2 = a2 + fr * uz;
}

}  // namespace

// One wavefront per atom.  Dynamic LDS (T): A[W], w[F], u[W], x[F].
// WM selects where w = dE_i/dB_i comes from:
//   0  precomputed (sparse-GP path);
//   1  mapped power 2 from the GEMM rows X and Z = beta X: E = X.Z/|X|^2,
//      w = 2 (Z - E X) / (|B| |X|)   (X = B, or B/|B| when normalized);
//   2  mapped power 1 from X and beta_s: E = X.beta/|X|,
//      w = (beta - E B/|B|) / |B|    (lammps_descriptor.cpp:262-270).
// Environments with |B|^2 < 1e-8 carry no energy (pair_flare.cpp:135).
template <typename T, int MODE, int NMAX, int WM>
__global__ void __launch_bounds__(64)
force_kernel(B2Dev p, NListDev nl, const int *__restrict__ species,
             const int *__restrict__ row_of, const T *__restrict__ ynorm,
             const T *__restrict__ A_in, WSource<T> src,
             const double *__restrict__ dA, double *__restrict__ pair_f,
             double *__restrict__ vir_atom) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int H = p.H, NH = p.NH, W = p.W, F = p.F, nr = p.nr, L1 = p.L + 1;
  T *A = reinterpret_cast<T *>(smem);
  T *w = A + W;
  T *u = w + F;
  const T *Ai = A_in + (int64_t)i * W;
  const int row = row_of[i];
  for (int o = lane; o < W; o += kWave) A[o] = Ai[o];
  if (WM == 0) {
    const T *wi = src.w + (int64_t)row * F;
    for (int f = lane; f < F; f += kWave) w[f] = wi[f];
  } else {
    const double n2 = src.norm2[row];
    const T *x = src.X + (int64_t)row * F;
    if (n2 < 1e-8) {
      for (int f = lane; f < F; f += kWave) w[f] = 0;
      if (lane == 0) src.E_row[row] = 0;
    } else {
      const double bn = sqrt(n2);
      const T *z = WM == 1 ? src.Z + (int64_t)row * F
                           : src.beta1 + (int64_t)species[i] * F;
      double d = 0;
      for (int f = lane; f < F; f += kWave) d += (double)x[f] * (double)z[f];
      d = wave_sum(d);
      double E, sx, sz;
      if (WM == 1) {
        E = src.normalized ? d : d / n2;
        sz = src.normalized ? 2.0 / bn : 2.0 / n2;
        sx = -E * sz;
      } else {
        E = src.normalized ? d : d / bn;
        sz = 1.0 / bn;
        sx = -E / bn * (src.normalized ? 1.0 : 1.0 / bn);
      }
      for (int f = lane; f < F; f += kWave)
        w[f] = (T)(sz * (double)z[f] + sx * (double)x[f]);
      if (lane == 0) src.E_row[row] = E;
    }
  }
  __syncthreads();
  for (int o = lane; o < W; o += kWave) {
    const int c = o / H, lm = o - (o / H) * H;
    const int l = lm_to_l(lm);
    T acc = 0;
    for (int c2 = 0; c2 < nr; c2++) {
      T t = w[b2_pair_index(c, c2, nr) * L1 + l] * A[c2 * H + lm];
      acc += (c2 == c) ? 2 * t : t;
    }
    u[o] = acc;
  }
  __syncthreads();
  const int si = species[i];
  const int b = nl.cum[i], e = nl.cum[i + 1];
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (MODE == 0) {
    for (int k = b; k < e; k++) {
      const double *rv = nl.rel + 4 * (int64_t)k;
      const int sj = nl.nspec[k];
      double f0 = 0, f1 = 0, f2 = 0;
      if (pair_inside(p, si, sj, rv)) {
        const double *d = dA + (int64_t)k * 3 * NH;
        const T *ub = u + sj * NH;
        for (int o = lane; o < NH; o += kWave) {
          double uo = (double)ub[o];
          f0 += d[o] * uo;
          f1 += d[NH + o] * uo;
          f2 += d[2 * NH + o] * uo;
        }
        f0 = wave_sum(f0);
        f1 = wave_sum(f1);
        f2 = wave_sum(f2);
      }
      if (lane == 0) {
        pair_f[3 * (int64_t)k + 0] = f0;
        pair_f[3 * (int64_t)k + 1] = f1;
        pair_f[3 * (int64_t)k + 2] = f2;
        const double r[3] = {rv[1], rv[2], rv[3]};
        const double f[3] = {f0, f1, f2};
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) v[3 * a + c] -= r[a] * f[c];
      }
    }
  } else {
    for (int k = b + lane; k < e; k += kWave) {
      const double *rv = nl.rel + 4 * (int64_t)k;
      const int sj = nl.nspec[k];
      double f[3] = {0, 0, 0};
      if (pair_inside(p, si, sj, rv)) {
        T f0, f1, f2;
        adjoint_pair<T, NMAX>(p, (T)rv[0], (T)rv[1], (T)rv[2], (T)rv[3],
                              (T)p.cut[si * p.S + sj], u + sj * NH, ynorm, f0,
                              f1, f2);
        f[0] = f0;
        f[1] = f1;
        f[2] = f2;
      }
      pair_f[3 * (int64_t)k + 0] = f[0];
      pair_f[3 * (int64_t)k + 1] = f[1];
      pair_f[3 * (int64_t)k + 2] = f[2];
      const double r[3] = {rv[1], rv[2], rv[3]};
      for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) v[3 * a + c] -= r[a] * f[c];
    }
#pragma unroll
    for (int q = 0; q < 9; q++) v[q] = wave_sum(v[q]);
  }
  if (lane == 0)
    for (int q = 0; q < 9; q++) vir_atom[9 * (int64_t)i + q] = v[q];
}

// 16 lanes per atom, four atoms per wavefront
__global__ void __launch_bounds__(256)
gather_kernel(int n, const int *__restrict__ cum, const int *__restrict__ rev,
              const double *__restrict__ pf, double *__restrict__ forces) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = t >> 4, sub = t & 15;
  double f0 = 0, f1 = 0, f2 = 0;
  if (a < n) {
    for (int k = cum[a] + sub; k < cum[a + 1]; k += 16) {
      const int r = rev[k];
      f0 += pf[3 * (int64_t)k];
      f1 += pf[3 * (int64_t)k + 1];
      f2 += pf[3 * (int64_t)k + 2];
      if (r >= 0) {
        f0 -= pf[3 * (int64_t)r];
        f1 -= pf[3 * (int64_t)r + 1];
        f2 -= pf[3 * (int64_t)r + 2];
      }
    }
  }
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) {
    f0 += __shfl_xor(f0, off, 16);
    f1 += __shfl_xor(f1, off, 16);
    f2 += __shfl_xor(f2, off, 16);
  }
  if (a < n && sub == 0) {
    forces[3 * (int64_t)a] = f0;
    forces[3 * (int64_t)a + 1] = f1;
    forces[3 * (int64_t)a + 2] = f2;
  }
}

// Deterministic two-stage sum of the per-row energies and per-atom virial
// rows: stage 1 gives kPartials partial sums over fixed contiguous ranges.
constexpr int kPartials = 128;

__global__ void __launch_bounds__(256)
partial_sums_kernel(int n, const double *__restrict__ e_row,
                    const double *__restrict__ vir, double *__restrict__ part) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(n, lo + per);
  for (int q = 0; q < 10; q++) {
    double s = 0;
    for (int r = lo + t; r < hi; r += blockDim.x)
      s += q == 0 ? e_row[r] : vir[(int64_t)r * 9 + (q - 1)];
    sh[t] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (t < off) sh[t] += sh[t + off];
      __syncthreads();
    }
    if (t == 0) part[blockIdx.x * 10 + q] = sh[0];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(64)
partial_final_kernel(int nparts, const double *__restrict__ part,
                     double *__restrict__ sums) {
  const int q = threadIdx.x;
  if (q >= 10) return;
  double s = 0;
  for (int b = 0; b < nparts; b++) s += part[b * 10 + q];
  sums[q] = s;
}

// one workgroup per column; fixed partition => deterministic
__global__ void __launch_bounds__(1024)
colsum_kernel(int n, int k, const double *__restrict__ x,
              double *__restrict__ out) {
  __shared__ double part[1024];
  const int col = blockIdx.x;
  const int t = threadIdx.x;
  double s = 0;
  for (int r = t; r < n; r += blockDim.x) s += x[(int64_t)r * k + col];
  part[t] = s;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) part[t] += part[t + off];
    __syncthreads();
  }
  if (t == 0) out[col] = part[0];
}

__global__ void finish_kernel(const double *__restrict__ sums,
                              double inv_volume, double *energy,
                              double *virial6, double *stress6) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double *V = sums + 1;  // V[a*3+b] = -sum r_a f_b
  if (energy) energy[0] = sums[0];
  if (virial6) {
    virial6[0] = V[0];
    virial6[1] = V[4];
    virial6[2] = V[8];
    virial6[3] = V[1];
    virial6[4] = V[2];
    virial6[5] = V[5];
  }
  if (stress6) {
    // (xx, xy, xz, yy, yz, zz) of -(1/V) sum f_c r_c2, c2 >= c
    // (normalized_dot_product.cpp:193-202)
    stress6[0] = V[0] * inv_volume;
    stress6[1] = V[3] * inv_volume;
    stress6[2] = V[6] * inv_volume;
    stress6[3] = V[4] * inv_volume;
    stress6[4] = V[7] * inv_volume;
    stress6[5] = V[8] * inv_volume;
  }
}

// ---------------------------------------------------------------------------
template <typename T, int MODE, int WM>
static void launch_forces_t(hipStream_t st, const B2Dev &p, const NListDev &nl,
                            const int *species, const int *row_of,
                            const T *ynorm, const void *A, const WSource<T> &src,
                            const double *dA, double *pair_f, double *vir_atom) {
  size_t lds = sizeof(T) * (2 * (size_t)p.W + (size_t)p.F);
  FLARE_REQUIRE(lds <= 64 * 1024, "descriptor settings too large for the force kernel");
  if (p.N <= 8)
    hipLaunchKernelGGL((force_kernel<T, MODE, 8, WM>), dim3(nl.n), dim3(kWave),
                       lds, st, p, nl, species, row_of, ynorm, (const T *)A, src,
                       dA, pair_f, vir_atom);
  else
    hipLaunchKernelGGL((force_kernel<T, MODE, kMaxN, WM>), dim3(nl.n),
                       dim3(kWave), lds, st, p, nl, species, row_of, ynorm,
                       (const T *)A, src, dA, pair_f, vir_atom);
}

template <typename T, int MODE>
static void launch_forces_w(hipStream_t st, const B2Dev &p, const NListDev &nl,
                            const int *species, const int *row_of,
                            const T *ynorm, const void *A, const WInput &wi,
                            const double *dA, double *pair_f, double *vir_atom) {
  WSource<T> src{wi.normalized, (const T *)wi.X, (const T *)wi.Z,
                 (const T *)wi.beta1, (const T *)wi.w, wi.norm2, wi.E_row};
  if (wi.mode == 0)
    launch_forces_t<T, MODE, 0>(st, p, nl, species, row_of, ynorm, A, src, dA,
                                pair_f, vir_atom);
  else if (wi.mode == 1)
    launch_forces_t<T, MODE, 1>(st, p, nl, species, row_of, ynorm, A, src, dA,
                                pair_f, vir_atom);
  else
    launch_forces_t<T, MODE, 2>(st, p, nl, species, row_of, ynorm, A, src, dA,
                                pair_f, vir_atom);
}

void launch_forces(hipStream_t st, int prec, int mode, const B2Dev &p,
                   const NListDev &nl, const int *species, const int *row_of,
                   const double *ynorm, const float *ynorm_f, const void *A,
                   const WInput &wi, const double *dA, double *pair_f,
                   double *vir_atom) {
  if (nl.n == 0) return;
  if (prec == kPrecF64) {
    if (mode == 0)
      launch_forces_w<double, 0>(st, p, nl, species, row_of, ynorm, A, wi, dA,
                                 pair_f, vir_atom);
    else
      launch_forces_w<double, 1>(st, p, nl, species, row_of, ynorm, A, wi, dA,
                                 pair_f, vir_atom);
  } else {
    FLARE_REQUIRE(mode == 1, "the plain (materialised) form runs in FP64 only");
    launch_forces_w<float, 1>(st, p, nl, species, row_of, ynorm_f, A, wi, dA,
                              pair_f, vir_atom);
  }
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_gather(hipStream_t st, const NListDev &nl, const double *pair_f,
                   double *forces) {
  if (nl.n == 0) return;
  hipLaunchKernelGGL(gather_kernel, dim3((nl.n * 16 + 255) / 256), dim3(256), 0,
                     st, nl.n, nl.cum, nl.rev, pair_f, forces);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_energy_virial_sums(hipStream_t st, int n, const double *e_row,
                               const double *vir, double *part, double *sums) {
  hipLaunchKernelGGL(partial_sums_kernel, dim3(kPartials), dim3(256), 0, st, n,
                     e_row, vir, part);
  hipLaunchKernelGGL(partial_final_kernel, dim3(1), dim3(64), 0, st, kPartials,
                     part, sums);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_colsum(hipStream_t st, int n, int k, const double *x, double *out) {
  hipLaunchKernelGGL(colsum_kernel, dim3(k), dim3(1024), 0, st, n, k, x, out);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_finish(hipStream_t st, const double *sums, double inv_volume,
                   double *energy, double *virial6, double *stress6) {
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(64), 0, st, sums, inv_volume,
                     energy, virial6, stress6);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
