// descriptors.hip -- single bond and B2 power spectrum.
//
//  b2_forward_kernel   per atom: Chebyshev x cutoff and Y_lm per pair,
//                      A[c, lm] = sum_pairs g_n Y_lm (b2.cpp:315-398,
//                      lammps_descriptor.cpp:7-107), then
//                      B[p(c1,c2) (L+1) + l] = sum_m A[c1,lm] A[c2,lm]
//                      (b2.cpp:203-241) and |B|^2.  One wavefront per atom;
//                      per-pair values are staged in LDS, each lane owns a
//                      set of (n, lm) outputs.  FP32 instantiation uses
//                      compensated (Kahan) sums.
//  pair_dA_kernel      plain form: materialises dA/dr per pair in the
//                      compact layout [pair][xyz][n][lm] (the neighbor's
//                      species block only; b2.cpp:381-391).
//  b2_dervs_kernel     DescriptorValues force derivatives dB/dr and
//                      B . dB/dr per pair (b2.cpp:219-240).
#include "common.h"
#include "kernels.h"
#include "small_mma.h"

namespace flare {

// Output stage shared by the forward kernels: B2 contraction from the
// single-bond values in LDS (b2.cpp:203-241), |B|^2 in FP64, and the rows
// handed to the contraction with beta / the sparse set:
//   out.mode 0: B itself;   1: d = B/|B| (zero when |B|^2 < 1e-8, the
//   empty-environment threshold of pair_flare.cpp:135), plus optionally its
//   BF16 hi + lo split into rows of out.kp (zero padded);  2: only that
//   split (the mixed-precision form reads d nowhere else).
// Rows the forward kernel writes beside A and norm2: mode 0 B, 1 d = B/|B|,
// 2 none; xh / xl (when set) d as hi + lo rows of kp, BF16 (f16 == 0) or FP16
// with one power-of-two scale per row (f16 == 1, inv_scale[row] = 2^-e).
struct FwdOut {
  int mode;
  int kp;
  __bf16 *xh, *xl;
  int f16;
  float *inv_scale;
};

// FP16 hi + lo of the d row x_f = Bs[f] sc, its largest |x| brought into
// [0.5, 1) by a power of two (split_f16_rows on the FP32 d rows, bit for bit)
template <typename T>
__device__ inline void store_f16_split(const T *Bs, T sc, int F, int row, const FwdOut &o) {
  const int lane = threadIdx.x;
  float m = 0.0f;
  for (int f = lane; f < F; f += kWave) m = fmaxf(m, fabsf((float)(Bs[f] * sc)));
  for (int off = kWave / 2; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  int e;
  const bool nonzero = f16_row_scale(m, e);
  const float s2 = ldexpf(1.0f, e);
  uint2 *xh = reinterpret_cast<uint2 *>(o.xh + (int64_t)row * o.kp);
  uint2 *xl = reinterpret_cast<uint2 *>(o.xl + (int64_t)row * o.kp);
  for (int f4 = lane; 4 * f4 < o.kp; f4 += kWave) {
    union {
      _Float16 h[4];
      uint2 u;
    } H, Lo;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int f = 4 * f4 + j;
      const float v = nonzero && f < F ? (float)(Bs[f] * sc) * s2 : 0.0f;
      const _Float16 h = (_Float16)v;
      H.h[j] = h;
      Lo.h[j] = (_Float16)(v - (float)h);
    }
    xh[f4] = H.u;
    xl[f4] = Lo.u;
  }
  if (lane == 0) o.inv_scale[row] = ldexpf(1.0f, -e);
}

template <typename T>
__device__ inline void b2_contract_store(const B2Dev &p, const T *Al, T *Bs,
                                         const int *__restrict__ b2_idx,
                                         int row, const FwdOut &o,
                                         T *__restrict__ B_out,
                                         double *__restrict__ norm2_out) {
  const int lane = threadIdx.x;
  const int H = p.H, F = p.F;
  double nrm = 0;
  for (int f = lane; f < F; f += kWave) {
    const int code = b2_idx[f];
    const int c1 = code >> 20, c2 = (code >> 8) & 0xfff, l = code & 0xff;
    const T *a1 = Al + c1 * H + l * l;
    const T *a2 = Al + c2 * H + l * l;
    T sacc = 0;
    for (int m = 0; m < 2 * l + 1; m++) sacc += a1[m] * a2[m];
    Bs[f] = sacc;
    nrm += (double)sacc * (double)sacc;
  }
  nrm = wave_sum(nrm);
  if (lane == 0) norm2_out[row] = nrm;
  __syncthreads();
  if (o.mode == 0) {
    T *Bi = B_out + (int64_t)row * F;
    for (int f = lane; f < F; f += kWave) Bi[f] = Bs[f];
    return;
  }
  const T sc = nrm < 1e-8 ? T(0) : (T)(1.0 / sqrt(nrm));
  if (o.mode == 1) {
    T *Bi = B_out + (int64_t)row * F;
    for (int f = lane; f < F; f += kWave) Bi[f] = Bs[f] * sc;
  }
  if (o.xh && o.f16) {
    store_f16_split<T>(Bs, sc, F, row, o);
  } else if (o.xh) {
    __bf16 *xh = o.xh + (int64_t)row * o.kp, *xl = o.xl + (int64_t)row * o.kp;
    for (int f = lane; f < o.kp; f += kWave) {
      float x = f < F ? (float)(Bs[f] * sc) : 0.0f;
      __bf16 h = (__bf16)x;
      xh[f] = h;
      xl[f] = (__bf16)(x - (float)h);
    }
  }
}

template <typename T>
struct KahanAcc {
  T s = 0, c = 0;
  __device__ inline void add(T v) {
    T y = v - c;
    T t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// One wavefront (64 threads) per atom.  Dynamic LDS layout (T):
//   A[W], Acomp[W], g[64*N], Y[64*H]   and int spec[64]
template <typename T, bool KAHAN>
__global__ void __launch_bounds__(64)
b2_forward_kernel(B2Dev p, NListDev nl, const int *__restrict__ species,
                  const int *__restrict__ row_of, const T *__restrict__ ynorm,
                  const int *__restrict__ b2_idx, FwdOut out,
                  T *__restrict__ A_out, T *__restrict__ B_out,
                  double *__restrict__ norm2_out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = p.N, H = p.H, NH = p.NH, W = p.W, F = p.F, S = p.S;
  T *A = reinterpret_cast<T *>(smem);
  T *Ac = A + W;
  T *gb = Ac + W;
  T *yb = gb + kWave * N;
  int *spec = reinterpret_cast<int *>(yb + kWave * H);
  T *Bs = reinterpret_cast<T *>(spec + kWave);  // F

  for (int o = lane; o < W; o += kWave) {
    A[o] = 0;
    Ac[o] = 0;
  }
  const int si = species[i];
  const int b = nl.cum[i], e = nl.cum[i + 1];
  for (int c0 = b; c0 < e; c0 += kWave) {
    const int cnt = min(kWave, e - c0);
    __syncthreads();
    if (lane < cnt) {
      const int q = c0 + lane;
      const double *rv = nl.rel + 4 * (int64_t)q;
      const int sj = nl.nspec[q];
      int sflag = -1;
      if (pair_inside(p, si, sj, rv)) {
        sflag = sj;
        T r = (T)rv[0];
        T rinv = T(1) / r;
        T ux = (T)rv[1] * rinv, uy = (T)rv[2] * rinv, uz = (T)rv[3] * rinv;
        T g[kMaxN];
        radial_basis<T>(r, (T)p.cut[si * S + sj], (T)p.r1, N, p.cutoff_kind, g,
                        (T *)nullptr);
        for (int n = 0; n < N; n++) gb[lane * N + n] = g[n];
        T *yrow = yb + lane * H;
        ylm_visit<T, false>(ux, uy, uz, rinv, p.L, ynorm,
                            [&](int idx, T y, T, T, T) { yrow[idx] = y; });
      }
      spec[lane] = sflag;
    }
    __syncthreads();
    for (int o = lane; o < NH; o += kWave) {
      const int n = o / H, lm = o - (o / H) * H;
      for (int s = 0; s < S; s++) {
        if (KAHAN) {
          KahanAcc<T> acc;
          for (int k = 0; k < cnt; k++)
            if (spec[k] == s) acc.add(gb[k * N + n] * yb[k * H + lm]);
          // merge into the running (sum, compensation) pair
          KahanAcc<T> run;
          run.s = A[s * NH + o];
          run.c = Ac[s * NH + o];
          run.add(acc.s);
          run.add(-acc.c);
          A[s * NH + o] = run.s;
          Ac[s * NH + o] = run.c;
        } else {
          T acc = 0;
          for (int k = 0; k < cnt; k++)
            if (spec[k] == s) acc += gb[k * N + n] * yb[k * H + lm];
          A[s * NH + o] += acc;
        }
      }
    }
  }
  __syncthreads();
  T *Ai = A_out + (int64_t)i * W;
  for (int o = lane; o < W; o += kWave) Ai[o] = A[o];
  // b2_idx[f] = (c1 << 20) | (c2 << 8) | l
  b2_contract_store<T>(p, A, Bs, b2_idx, row_of[i], out, B_out, norm2_out);
}

// ---------------------------------------------------------------------------
// Matrix-core single bond.  Per atom the accumulation is a small product
//   A (nr x H) = G^T (nr x K) . Y (K x H),  K = pairs,
// G[k][s*N + n] = g_n(r_k) when the neighbor's species is s, else 0.
// One wavefront per atom; 64 pairs are staged per chunk (lane = pair) and
// consumed by v_mfma_*_16x16x4 in K steps of 4 pairs.  Operand layouts
// (64 lanes): A[i][k] lane = i + 16 k; B[k][j] lane = j + 16 k; the
// result D[i][j] sits in lane j + 16 (i / 4), element i % 4 for the FP32
// instruction and in lane j + 16 (i % 4), element i / 4 for the FP64 one
// (both measured on gfx950).
// FP32: every two K steps (8 pairs) are accumulated from zero and
// Kahan-added into the running (sum, compensation) pair; FP64 accumulates
// in the matrix core.

// Output rows from Bs (LDS) and |B|^2 (see b2_contract_store); the BF16
// split is written four values (8 bytes) per lane store.
template <typename T>
__device__ inline void b2_store_rows(const B2Dev &p, const T *Bs, double nrm,
                                     int row, const FwdOut &o,
                                     T *__restrict__ B_out,
                                     double *__restrict__ norm2_out) {
  const int lane = threadIdx.x;
  const int F = p.F;
  if (lane == 0) norm2_out[row] = nrm;
  if (o.mode == 0) {
    T *Bi = B_out + (int64_t)row * F;
    for (int f = lane; f < F; f += kWave) Bi[f] = Bs[f];
    return;
  }
  const T sc = nrm < 1e-8 ? T(0) : (T)(1.0 / sqrt(nrm));
  if (o.mode == 1) {
    T *Bi = B_out + (int64_t)row * F;
    for (int f = lane; f < F; f += kWave) Bi[f] = Bs[f] * sc;
  }
  if (o.xh && o.f16) {
    store_f16_split<T>(Bs, sc, F, row, o);
  } else if (o.xh) {
    uint2 *xh = reinterpret_cast<uint2 *>(o.xh + (int64_t)row * o.kp);
    uint2 *xl = reinterpret_cast<uint2 *>(o.xl + (int64_t)row * o.kp);
    for (int f4 = lane; 4 * f4 < o.kp; f4 += kWave) {
      union {
        __bf16 h[4];
        uint2 u;
      } H, Lo;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int f = 4 * f4 + j;
        float x = 0.0f;
        if (f < F) x = (float)(Bs[f] * sc);
        const __bf16 h = (__bf16)x;
        H.h[j] = h;
        Lo.h[j] = (__bf16)(x - (float)h);
      }
      xh[f4] = H.u;
      xl[f4] = Lo.u;
    }
  }
}

// Pair staging of b2_forward_mfma_kernel, in T values.  General layout: a
// G row of MT*16 channels and a Y row of NT*16 harmonics per pair (each
// padded by one).  Compact layout (n_max == NMAX and l_max fixed): the
// NMAX radial values and the neighbor's species per pair (the channel mask
// s = c / NMAX is applied when the operand is read) and an unpadded row of
// the (L+1)^2 harmonics (bounds applied when read), for chunks of 48 pairs:
// most atoms of dense solids and liquids at rc 5-5.5 A have 30-45 pairs, and
// the smaller LDS footprint admits a third more wavefronts per CU.
template <int MT, int NT, int NMAX, int LFIX, bool EXACT>
struct FwdStage {
  static constexpr bool kCompact = EXACT && LFIX >= 0;
  static constexpr int kRows = kCompact ? 48 : kWave;  // pairs per chunk
  static constexpr int kGP = kCompact ? NMAX + 1 : MT * 16 + 1;
  static constexpr int kYP = kCompact ? (LFIX + 1) * (LFIX + 1) : NT * 16 + 1;
  static constexpr int kStage = kRows * (kGP + kYP);
  static_assert(kRows % 4 == 0 && kRows <= kWave, "whole K steps, one pair per lane");
};

// LDS values of b2_forward_mfma_kernel: the staging rows, and B (F) then A
// (nr x H) from the start once the staging rows are dead (a barrier ends
// the single-bond loop before A is written).  At l_max 6 with 24 channels
// this is what bounds the wavefronts per CU.
__host__ __device__ inline size_t fwd_lds_values(int stage, int F, int A) {
  return F + A <= stage ? (size_t)stage : (size_t)(F + A);
}

// LFIX >= 0: l_max fixed at compile time; EXACT: n_max == NMAX.
template <typename T, int MT, int NT, bool KAHAN, int NMAX, int LFIX, bool EXACT>
__global__ void __launch_bounds__(64)
b2_forward_mfma_kernel(B2Dev p, NListDev nl, const int *__restrict__ species,
                       const int *__restrict__ row_of,
                       const T *__restrict__ ynorm, FwdOut out,
                       T *__restrict__ A_out, T *__restrict__ B_out,
                       double *__restrict__ norm2_out) {
  typedef typename Mma<T>::V V;
  typedef FwdStage<MT, NT, NMAX, LFIX, EXACT> St;
  constexpr bool kCompact = St::kCompact;
  constexpr int KP = St::kGP;  // staged row of G
  constexpr int YP = St::kYP;  // staged row of Y
  constexpr int CH = St::kRows;  // pairs per chunk
  extern __shared__ __align__(16) unsigned char smem[];
  // B (F values) and A (nr x H) reuse the pair staging rows, dead once A is
  // formed (fwd_lds_values)
  T *Gs = reinterpret_cast<T *>(smem);
  T *Ys = Gs + CH * KP;
  T *Bs = Gs;
  T *Al = Gs + p.F;
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = EXACT ? NMAX : p.N;
  const int H = LFIX >= 0 ? (LFIX + 1) * (LFIX + 1) : p.H;
  const int W = p.W, S = p.S, nr = p.nr;
  V acc[MT][NT], cmp[MT][NT];
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < NT; b++) {
      acc[a][b] = V{0, 0, 0, 0};
      cmp[a][b] = V{0, 0, 0, 0};
    }
  const int si = species[i];
  const int beg = nl.cum[i], end = nl.cum[i + 1];
  const int col = lane & 15, krow = lane >> 4;
  for (int c0 = beg; c0 < end; c0 += CH) {
    const int cnt = min(CH, end - c0);
    __syncthreads();
    if (lane < CH) {
      T *gr = Gs + lane * KP;
      T *yr = Ys + lane * YP;
      bool ok = false;
      int sj = 0;
      const double *rv = nullptr;
      if (lane < cnt) {
        const int q = c0 + lane;
        rv = nl.rel + 4 * (int64_t)q;
        sj = nl.nspec[q];
        ok = pair_inside(p, si, sj, rv);
      }
      if (ok) {
        const T r = (T)rv[0];
        const T rinv = T(1) / r;
        T g[NMAX];
        radial_basis<T, NMAX, EXACT>(r, (T)p.cut[si * S + sj], (T)p.r1, N,
                                     p.cutoff_kind, g, (T *)nullptr);
        if (kCompact) {
#pragma unroll
          for (int n = 0; n < NMAX; n++) gr[n] = g[n];
          gr[NMAX] = (T)sj;
        } else if (EXACT) {
          // channel c = s NMAX + n holds g_n for the neighbor's species only
#pragma unroll
          for (int c = 0; c < MT * 16; c++)
            gr[c] = (c / NMAX == sj) ? g[c % NMAX] : T(0);
        } else {
          for (int c = 0; c < MT * 16; c++) gr[c] = 0;
#pragma unroll
          for (int n = 0; n < NMAX; n++)
            if (n < N) gr[sj * N + n] = g[n];
        }
        if (LFIX < 0)
          for (int c = 0; c < NT * 16; c++) yr[c] = 0;
        ylm_visit<T, false, LFIX>((T)rv[1] * rinv, (T)rv[2] * rinv,
                                  (T)rv[3] * rinv, rinv, p.L, ynorm,
                                  [&](int idx, T y, T, T, T) { yr[idx] = y; });
        if (LFIX >= 0 && !kCompact) {
#pragma unroll
          for (int c = (LFIX + 1) * (LFIX + 1); c < NT * 16; c++) yr[c] = 0;
        }
      } else if (kCompact) {
        // no species matches -1: the G operand is zero; Y zero keeps the
        // products finite
        gr[NMAX] = T(-1);
#pragma unroll
        for (int c = 0; c < YP; c++) yr[c] = 0;
      } else {
#pragma unroll
        for (int c = 0; c < MT * 16; c++) gr[c] = 0;
#pragma unroll
        for (int c = 0; c < NT * 16; c++) yr[c] = 0;
      }
    }
    __syncthreads();
    // K steps of 4 pairs.  FP32 sums KG steps (4 KG pairs) in the matrix
    // core and Kahan-adds them into the running (sum, compensation) pair.
    constexpr int KG = KAHAN ? 4 : 2;
    const int steps = (cnt + 3) >> 2;
    for (int st = 0; st < steps; st += KG) {
      T a[KG][MT], b[KG][NT];
#pragma unroll
      for (int u = 0; u < KG; u++) {
        const int kk = 4 * (st + u) + krow;
        const bool on = st + u < steps;  // uniform
        if (kCompact) {
          const T sk = on ? Gs[kk * KP + NMAX] : T(-1);
#pragma unroll
          for (int m = 0; m < MT; m++) {
            const int c = m * 16 + col;
            a[u][m] = sk == (T)(c / NMAX) ? Gs[kk * KP + (c % NMAX)] : T(0);
          }
#pragma unroll
          for (int n = 0; n < NT; n++) {
            const int lm = n * 16 + col;
            b[u][n] = (on && lm < YP) ? Ys[kk * YP + lm] : T(0);
          }
        } else {
#pragma unroll
          for (int m = 0; m < MT; m++) a[u][m] = on ? Gs[kk * KP + m * 16 + col] : T(0);
#pragma unroll
          for (int n = 0; n < NT; n++) b[u][n] = on ? Ys[kk * YP + n * 16 + col] : T(0);
        }
      }
#pragma unroll
      for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++) {
          if (KAHAN) {
            V part = Mma<T>::run(a[0][m], b[0][n], V{0, 0, 0, 0});
#pragma unroll
            for (int u = 1; u < KG; u++) part = Mma<T>::run(a[u][m], b[u][n], part);
            V y = part - cmp[m][n];
            V t = acc[m][n] + y;
            cmp[m][n] = (t - acc[m][n]) - y;
            acc[m][n] = t;
          } else {
#pragma unroll
            for (int u = 0; u < KG; u++) acc[m][n] = Mma<T>::run(a[u][m], b[u][n], acc[m][n]);
          }
        }
    }
  }
  __syncthreads();
  T *Ai = A_out + (int64_t)i * W;
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int n = 0; n < NT; n++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int c = m * 16 + Mma<T>::row(krow, j);
        const int lm = n * 16 + col;
        if (c < nr && lm < H) {
          T v = KAHAN ? acc[m][n][j] - cmp[m][n][j] : acc[m][n][j];
          Al[c * H + lm] = v;
          Ai[c * H + lm] = v;
        }
      }
  __syncthreads();
  double nrm = b2_gram_mfma<T, MT, LFIX>(p, Al, Bs);
  nrm = wave_sum(nrm);
  __syncthreads();
  b2_store_rows<T>(p, Bs, nrm, row_of[i], out, B_out, norm2_out);
}

template <typename T, bool KAHAN, int MT, int NT, int NMAX, int LFIX, bool EXACT>
static void launch_mfma_k(hipStream_t st, const B2Dev &p, const NListDev &nl,
                          const int *species, const int *row_of, const T *ynorm,
                          const FwdOut &out, void *A_out, void *B_out,
                          double *norm2_out) {
  typedef FwdStage<MT, NT, NMAX, LFIX, EXACT> St;
  const size_t lds = sizeof(T) * fwd_lds_values(St::kStage, p.F, p.nr * p.H);
  FLARE_REQUIRE(lds <= 64 * 1024, "descriptor settings too large");
  hipLaunchKernelGGL((b2_forward_mfma_kernel<T, MT, NT, KAHAN, NMAX, LFIX, EXACT>),
                     dim3(nl.n), dim3(kWave), lds, st, p, nl, species, row_of,
                     ynorm, out, (T *)A_out, (T *)B_out, norm2_out);
}

template <typename T, bool KAHAN, int MT, int NT>
static void launch_mfma_t(hipStream_t st, const B2Dev &p, const NListDev &nl,
                          const int *species, const int *row_of, const T *ynorm,
                          const FwdOut &out, void *A_out, void *B_out,
                          double *norm2_out) {
  // compile-time (n_max, l_max) variants of the common settings
  if (p.N == 8 && p.L == 3 && NT == 1) {
    launch_mfma_k<T, KAHAN, MT, NT, 8, 3, true>(st, p, nl, species, row_of, ynorm,
                                                out, A_out, B_out, norm2_out);
  } else if (p.N == 8 && p.L == 4 && NT == 2) {
    launch_mfma_k<T, KAHAN, MT, NT, 8, 4, true>(st, p, nl, species, row_of, ynorm,
                                                out, A_out, B_out, norm2_out);
  } else if (p.N == 8 && p.L == 6 && NT == 4) {
    launch_mfma_k<T, KAHAN, MT, NT, 8, 6, true>(st, p, nl, species, row_of, ynorm,
                                                out, A_out, B_out, norm2_out);
  } else if (p.N <= 8) {
    launch_mfma_k<T, KAHAN, MT, NT, 8, -1, false>(st, p, nl, species, row_of, ynorm,
                                                  out, A_out, B_out, norm2_out);
  } else {
    launch_mfma_k<T, KAHAN, MT, NT, kMaxN, -1, false>(st, p, nl, species, row_of,
                                                      ynorm, out, A_out, B_out,
                                                      norm2_out);
  }
}

template <typename T, bool KAHAN>
static bool launch_mfma(hipStream_t st, const B2Dev &p, const NListDev &nl,
                        const int *species, const int *row_of, const T *ynorm,
                        const FwdOut &out, void *A_out, void *B_out,
                        double *norm2_out) {
  const int MT = (p.nr + 15) / 16, NT = (p.H + 15) / 16;
#define FLARE_MFMA_CASE(m, n)                                                 \
  if (MT == m && NT == n) {                                                   \
    launch_mfma_t<T, KAHAN, m, n>(st, p, nl, species, row_of, ynorm, out,     \
                                  A_out, B_out, norm2_out);                   \
    return true;                                                              \
  }
  FLARE_MFMA_CASE(1, 1)
  FLARE_MFMA_CASE(1, 2)
  FLARE_MFMA_CASE(1, 3)
  FLARE_MFMA_CASE(1, 4)
  FLARE_MFMA_CASE(2, 1)
  FLARE_MFMA_CASE(2, 2)
  FLARE_MFMA_CASE(2, 3)
  FLARE_MFMA_CASE(2, 4)
#undef FLARE_MFMA_CASE
  return false;
}

// Plain form: dA/dr for `pb` (<= 64) pairs per workgroup of 256 threads,
// pairs p_begin <= p < p_end: out[((p - p_begin)*3 + c)*NH + n*H + lm];
// zero for pairs outside the species cutoff.
__global__ void __launch_bounds__(256)
pair_dA_kernel(B2Dev p, NListDev nl, const int *__restrict__ species,
               const double *__restrict__ ynorm, int pb, int64_t p_begin,
               int64_t p_end, double *__restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = p.N, H = p.H, NH = p.NH;
  double *g = reinterpret_cast<double *>(smem);  // pb*N
  double *dg = g + pb * N;                       // pb*N
  double *Y = dg + pb * N;                       // pb*H
  double *Yg = Y + pb * H;                       // pb*3*H
  double *uhat = Yg + pb * 3 * H;                // pb*3
  int *ins = reinterpret_cast<int *>(uhat + pb * 3);
  const int64_t p0 = p_begin + (int64_t)blockIdx.x * pb;
  const int t = threadIdx.x;
  if (t < pb) {
    const int64_t q = p0 + t;
    int inside = 0;
    if (q < p_end) {
      const double *rv = nl.rel + 4 * q;
      const int si = species[nl.center[q]], sj = nl.nspec[q];
      if (pair_inside(p, si, sj, rv)) {
        inside = 1;
        double r = rv[0], rinv = 1.0 / r;
        double ux = rv[1] * rinv, uy = rv[2] * rinv, uz = rv[3] * rinv;
        double gl[kMaxN], dgl[kMaxN];
        radial_basis<double>(r, p.cut[si * p.S + sj], p.r1, N, p.cutoff_kind,
                             gl, dgl);
        for (int n = 0; n < N; n++) {
          g[t * N + n] = gl[n];
          dg[t * N + n] = dgl[n];
        }
        double *yr = Y + t * H;
        double *ygr = Yg + t * 3 * H;
        ylm_visit<double, true>(ux, uy, uz, rinv, p.L, ynorm,
                                [&](int idx, double y, double yx, double yy,
                                    double yz) {
                                  yr[idx] = y;
                                  ygr[idx] = yx;
                                  ygr[H + idx] = yy;
                                  ygr[2 * H + idx] = yz;
                                });
        uhat[3 * t] = ux;
        uhat[3 * t + 1] = uy;
        uhat[3 * t + 2] = uz;
      }
    }
    ins[t] = inside;
  }
  __syncthreads();
  // each thread owns fixed (n, lm) columns and walks the block's pairs, so
  // consecutive threads write consecutive addresses of every (pair, xyz) row
  const int per = 3 * NH;
  const int npair = (int)min((int64_t)pb, p_end - p0);
  double *o = out + (p0 - p_begin) * per;
  for (int col = t; col < NH; col += blockDim.x) {
    const int n = col / H, lm = col - n * H;
    for (int k = 0; k < npair; k++) {
      double *ok = o + (int64_t)k * per + col;
      if (ins[k]) {
        const double gn = g[k * N + n], dgn = dg[k * N + n], y = Y[k * H + lm];
        const double *yg = Yg + k * 3 * H + lm;
        ok[0] = dgn * uhat[3 * k] * y + gn * yg[0];
        ok[NH] = dgn * uhat[3 * k + 1] * y + gn * yg[H];
        ok[2 * NH] = dgn * uhat[3 * k + 2] * y + gn * yg[2 * H];
      } else {
        ok[0] = 0;
        ok[NH] = 0;
        ok[2 * NH] = 0;
      }
    }
  }
}

// DescriptorValues force derivatives.  One wavefront per filtered pair q
// (pairs of the B2 list in species-grouped order, fpair[q] = index in the
// Structure list).  Writes dB[(3q + c)*F + f], fdots[3q + c] = B . dB,
// coords[3q + c] and nidx[q].
__global__ void __launch_bounds__(64)
b2_dervs_kernel(B2Dev p, NListDev nl, const int *__restrict__ species,
                const int *__restrict__ row_of, const double *__restrict__ ynorm,
                const int *__restrict__ b2_idx, int64_t nq,
                const int64_t *__restrict__ fpair,
                const double *__restrict__ A_in, const double *__restrict__ B_in,
                double *__restrict__ dB, double *__restrict__ fdots,
                double *__restrict__ coords, int *__restrict__ nidx) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = p.N, H = p.H, NH = p.NH, W = p.W, F = p.F;
  double *A = reinterpret_cast<double *>(smem);  // W
  double *dA = A + W;                            // 3*NH
  double *g = dA + 3 * NH;                       // N
  double *dg = g + N;                            // N
  double *Y = dg + N;                            // H
  double *Yg = Y + H;                            // 3H
  double *uh = Yg + 3 * H;                       // 3
  const int64_t q = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t ps = fpair[q];
  const int i = nl.center[ps];
  const int si = species[i], sj = nl.nspec[ps];
  const double *rv = nl.rel + 4 * ps;
  for (int o = lane; o < W; o += kWave) A[o] = A_in[(int64_t)i * W + o];
  if (lane == 0) {
    double r = rv[0], rinv = 1.0 / r;
    double ux = rv[1] * rinv, uy = rv[2] * rinv, uz = rv[3] * rinv;
    double gl[kMaxN], dgl[kMaxN];
    radial_basis<double>(r, p.cut[si * p.S + sj], p.r1, N, p.cutoff_kind, gl,
                         dgl);
    for (int n = 0; n < N; n++) {
      g[n] = gl[n];
      dg[n] = dgl[n];
    }
    ylm_visit<double, true>(ux, uy, uz, rinv, p.L, ynorm,
                            [&](int idx, double y, double yx, double yy,
                                double yz) {
                              Y[idx] = y;
                              Yg[idx] = yx;
                              Yg[H + idx] = yy;
                              Yg[2 * H + idx] = yz;
                            });
    uh[0] = ux;
    uh[1] = uy;
    uh[2] = uz;
    coords[3 * q + 0] = rv[1];
    coords[3 * q + 1] = rv[2];
    coords[3 * q + 2] = rv[3];
    nidx[q] = nl.idx[ps];
  }
  __syncthreads();
  for (int v = lane; v < 3 * NH; v += kWave) {
    int c = v / NH, rem = v - c * NH;
    int n = rem / H, lm = rem - (rem / H) * H;
    dA[v] = dg[n] * uh[c] * Y[lm] + g[n] * Yg[c * H + lm];
  }
  __syncthreads();
  const double *Bi = B_in + (int64_t)row_of[i] * F;
  double fd[3] = {0, 0, 0};
  for (int f = lane; f < F; f += kWave) {
    const int code = b2_idx[f];
    const int c1 = code >> 20, c2 = (code >> 8) & 0xfff, l = code & 0xff;
    const bool in1 = (c1 / N) == sj, in2 = (c2 / N) == sj;
    const int n1 = c1 - (c1 / N) * N, n2 = c2 - (c2 / N) * N;
    for (int c = 0; c < 3; c++) {
      double s = 0;
      for (int m = 0; m < 2 * l + 1; m++) {
        int lm = l * l + m;
        double d2 = in2 ? dA[c * NH + n2 * H + lm] : 0.0;
        double d1 = in1 ? dA[c * NH + n1 * H + lm] : 0.0;
        s += A[c1 * H + lm] * d2 + d1 * A[c2 * H + lm];
      }
      dB[(3 * q + c) * F + f] = s;
      fd[c] += s * Bi[f];
    }
  }
  for (int c = 0; c < 3; c++) {
    double v = wave_sum(fd[c]);
    if (lane == 0) fdots[3 * q + c] = v;
  }
}

// ---------------------------------------------------------------------------
// Host launchers

void launch_b2_forward(hipStream_t st, const B2Dev &p, const NListDev &nl,
                       const int *species, const int *row_of,
                       const double *ynorm, const float *ynorm_f,
                       const int *b2_idx, int prec, int out_mode, int kp,
                       void *xh, void *xl, void *A_out, void *B_out,
                       double *norm2_out, float *f16_inv_scale) {
  if (nl.n == 0) return;
  FLARE_REQUIRE(!f16_inv_scale || (prec != kPrecF64 && xh && xl),
                "FP16 row split: FP32 rows with hi / lo outputs");
  FwdOut out{out_mode, kp, (__bf16 *)xh, (__bf16 *)xl, f16_inv_scale ? 1 : 0, f16_inv_scale};
  // matrix-core path when the channel and harmonic counts fit its tiles.
  // FP32 single-bond sums are Kahan-compensated in the FP32 form; the mixed
  // form (out_mode 2) sums plainly: its error is set by the BF16 hi + lo
  // product (force RMS relative error 4.68e-6 with compensation, 4.69e-6
  // without, at 32,000 atoms) and the forward kernel runs 65 -> 61 us.
  bool done;
  if (prec == kPrecF64)
    done = launch_mfma<double, false>(st, p, nl, species, row_of, ynorm, out, A_out, B_out,
                                      norm2_out);
  else if (out_mode == 2)
    done = launch_mfma<float, false>(st, p, nl, species, row_of, ynorm_f, out, A_out, B_out,
                                     norm2_out);
  else
    done = launch_mfma<float, true>(st, p, nl, species, row_of, ynorm_f, out, A_out, B_out,
                                    norm2_out);
  if (!done) {
    size_t tsz = prec == kPrecF64 ? sizeof(double) : sizeof(float);
    size_t lds = tsz * (2 * (size_t)p.W + kWave * (size_t)(p.N + p.H) + p.F) +
                 kWave * sizeof(int);
    FLARE_REQUIRE(lds <= 64 * 1024,
                  "descriptor settings too large for the single-bond kernel");
    if (prec == kPrecF64)
      hipLaunchKernelGGL((b2_forward_kernel<double, false>), dim3(nl.n),
                         dim3(kWave), lds, st, p, nl, species, row_of, ynorm,
                         b2_idx, out, (double *)A_out, (double *)B_out, norm2_out);
    else
      hipLaunchKernelGGL((b2_forward_kernel<float, true>), dim3(nl.n),
                         dim3(kWave), lds, st, p, nl, species, row_of, ynorm_f,
                         b2_idx, out, (float *)A_out, (float *)B_out, norm2_out);
  }
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_pair_dA(hipStream_t st, const B2Dev &p, const NListDev &nl,
                    const int *species, const double *ynorm, int64_t p_begin,
                    int64_t p_end, double *out) {
  if (p_end <= p_begin) return;
  // pairs per workgroup: as many as fit 64 KB of LDS, at most 64
  const size_t per_pair =
      sizeof(double) * (2 * (size_t)p.N + 4 * (size_t)p.H + 3) + sizeof(int);
  int pb = (int)std::min<size_t>(kWave, (64 * 1024) / per_pair);
  FLARE_REQUIRE(pb >= 1, "descriptor settings too large for the plain form");
  size_t lds = per_pair * pb;
  int64_t blocks = (p_end - p_begin + pb - 1) / pb;
  hipLaunchKernelGGL(pair_dA_kernel, dim3((unsigned)blocks), dim3(256), lds,
                     st, p, nl, species, ynorm, pb, p_begin, p_end, out);
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_b2_dervs(hipStream_t st, const B2Dev &p, const NListDev &nl,
                     const int *species, const int *row_of,
                     const double *ynorm, const int *b2_idx, int64_t nq,
                     const int64_t *fpair, const double *A, const double *B,
                     double *dB, double *fdots, double *coords, int *nidx) {
  if (nq == 0) return;
  size_t lds = sizeof(double) *
               ((size_t)p.W + 3 * (size_t)p.NH + 2 * (size_t)p.N + 4 * (size_t)p.H + 3);
  hipLaunchKernelGGL(b2_dervs_kernel, dim3((unsigned)nq), dim3(kWave), lds, st,
                     p, nl, species, row_of, ynorm, b2_idx, nq, fpair, A, B, dB,
                     fdots, coords, nidx);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
