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
#include "small_mma.h"

namespace flare {

template <typename T>
struct WSource {
  int normalized;
  const T *X, *Z, *beta1, *w;
  const __bf16 *xh, *xl;  // X as BF16 hi + lo rows of kp (mixed form)
  int kp;
  const double *norm2;
  double *E_row;
};

// dst[o] = load(o) for o < n (dst in LDS); each lane issues KB loads
// before storing any of them.
template <int KB, typename T, typename Fn>
__device__ inline void stage_lds(int n, Fn &&load, T *dst) {
  const int lane = threadIdx.x;
  for (int base = 0; base < n; base += KB * kWave) {
    T t[KB];
#pragma unroll
    for (int j = 0; j < KB; j++) {
      const int o = base + lane + j * kWave;
      t[j] = 0;
      if (o < n) t[j] = load(o);
    }
#pragma unroll
    for (int j = 0; j < KB; j++) {
      const int o = base + lane + j * kWave;
      if (o < n) dst[o] = t[j];
    }
  }
}

// dst[0..n) = src[0..n) (dst in LDS) in 16-byte vectors, KB per lane per
// batch, all loads of a batch issued before its stores; scalar when src or
// dst is not 16-byte aligned or n * sizeof(T) is not a multiple of 16.
template <int KB, typename T>
__device__ inline void stage_lds_vec(int n, const T *__restrict__ src, T *dst) {
  const bool vec = ((n * (int)sizeof(T)) & 15) == 0 &&
                   ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
  if (!vec) {
    stage_lds<KB>(n, [&](int o) { return src[o]; }, dst);
    return;
  }
  const int lane = threadIdx.x;
  const int nv = n * (int)sizeof(T) / 16;
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
  uint4 *d4 = reinterpret_cast<uint4 *>(dst);
  for (int base = 0; base < nv; base += KB * kWave) {
    uint4 t[KB];
#pragma unroll
    for (int j = 0; j < KB; j++) {
      const int o = base + lane + j * kWave;
      t[j] = uint4{0, 0, 0, 0};
      if (o < nv) t[j] = s4[o];
    }
#pragma unroll
    for (int j = 0; j < KB; j++) {
      const int o = base + lane + j * kWave;
      if (o < nv) d4[o] = t[j];
    }
  }
}

// u[c, lm] = sum_c2 W_l[c][c2] A[c2, lm] with W_l[c][c2] =
// w[p(c,c2) (L+1) + l] (1 + d_{c c2}) (lammps_descriptor.cpp:272-303), one
// matrix-core product per l: M = channels (c), K = channels (c2), N = the
// 2l+1 harmonics of that l.  Result stored as u2[(s H + lm) N + n] for
// channel c = s N + n, so the per-pair loop reads the n of one lm together.
// SCALE: each u[., lm] is multiplied by the harmonic's normalisation
// nrm_lm[lm] = N_l|m|, the form adjoint_pair_fixed reads.
// ptab[c nr + c2] = p(c, c2) (L+1): the feature offset of a channel pair.
// CHANNEL_MAJOR: store u[c H + lm] (the dA/dr layout of the plain form)
// instead of u2[(s H + lm) N + n].
template <typename T, int LFIX, bool SCALE, int NFIX = -1, bool CHANNEL_MAJOR = false>
__device__ inline void u_mfma(const B2Dev &p, const T *A, const T *w,
                              const T *nrm_lm, const int *ptab, T *u2) {
  typedef typename Mma<T>::V V;
  const int lane = threadIdx.x;
  const int col = lane & 15, g = lane >> 4;
  const int H = LFIX >= 0 ? (LFIX + 1) * (LFIX + 1) : p.H;
  const int L1 = LFIX >= 0 ? LFIX + 1 : p.L + 1;
  const int N = NFIX > 0 ? NFIX : p.N, nr = p.nr;
  const int MT = (nr + 15) >> 4, KS = (nr + 3) >> 2;
  for (int mt = 0; mt < MT; mt++) {
    const int c = mt * 16 + col;
    // destination (s H) N + n of this lane's four result rows
    int dst[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int cr = mt * 16 + Mma<T>::row(g, q);
      const int s = cr / N, n = cr - s * N;
      dst[q] = cr < nr ? (CHANNEL_MAJOR ? cr * H : s * H * N + n) : -1;
    }
#pragma unroll 1
    for (int l = 0; l < L1; l++) {
      const int nl = 2 * l + 1, l2 = l * l;
      for (int nt = 0; nt * 16 < nl; nt++) {
        const int j = nt * 16 + col;
        V acc = V{0, 0, 0, 0};
        for (int ks = 0; ks < KS; ks++) {
          const int c2 = 4 * ks + g;
          T a = 0, b = 0;
          if (c < nr && c2 < nr) {
            const T wv = w[ptab[c * nr + c2] + l];
            a = c == c2 ? 2 * wv : wv;
          }
          if (c2 < nr && j < nl) b = A[c2 * H + l2 + j];
          acc = Mma<T>::run(a, b, acc);
        }
        if (j < nl) {
          const T sc = SCALE ? nrm_lm[l2 + j] : T(1);
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (dst[q] >= 0) u2[dst[q] + (l2 + j) * (CHANNEL_MAJOR ? 1 : N)] = acc[q] * sc;
        }
      }
    }
  }
}

// u_mfma for l_max = L, n_max = 8 and NR (8, 16 or 24) radial channels:
// the adjoint forms' fixed settings with one to three species.
// Harmonic normalisation applied, u2 layout as u_mfma's.  Each lane's matrix
// operand W_l[c][c2] for its (c, c2 = 4 ks + g) and every l is read once into
// registers (the L + 1 values of a channel pair are adjacent in w), so the
// MFMA loop only loads the A operand, and all bounds are compile-time.
template <typename T, int L, int NR>
__device__ inline void u_mfma_fixed(const T *A, const T *w, const T *nrm_lm, T *u2) {
  static_assert(NR % 8 == 0 && NR <= 24, "one or two M tiles of channels");
  typedef typename Mma<T>::V V;
  constexpr int L1 = L + 1, H = L1 * L1, N = 8, KS = NR / 4, MT = (NR + 15) / 16;
  const int lane = threadIdx.x;
  const int col = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
    const int c = mt * 16 + col;
    const bool crow = c < NR;
    T wv[KS][L1];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int c2 = 4 * ks + g;
      const int off = crow ? b2_pair_index(c, c2, NR) * L1 : 0;
      const T fac = c == c2 ? T(2) : T(1);
#pragma unroll
      for (int l = 0; l < L1; l++) wv[ks][l] = crow ? fac * w[off + l] : T(0);
    }
    // this lane's four result rows cr = mt 16 + row(g, q) < NR: channel (s, n)
    int dst[4];
    bool drow[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int cr = mt * 16 + Mma<T>::row(g, q);
      drow[q] = cr < NR;
      dst[q] = (cr >> 3) * H * N + (cr & 7);
    }
#pragma unroll
    for (int l = 0; l < L1; l++) {
      constexpr int kMaxJ = 16;
      const int nl = 2 * l + 1, l2 = l * l;
      static_assert(2 * L + 1 <= kMaxJ, "one N tile of harmonics");
      const int j = col;
      const bool jok = j < nl;
      V acc = V{0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int c2 = 4 * ks + g;
        const T b = jok ? A[c2 * H + l2 + j] : T(0);
        acc = Mma<T>::run(wv[ks][l], b, acc);
      }
      if (jok) {
        const T sc = nrm_lm[l2 + j];
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (drow[q]) u2[dst[q] + (l2 + j) * N] = acc[q] * sc;
      }
    }
  }
}

// u_mfma_fixed with A given in u2's layout, Au[(s H + lm) N + n] for channel
// c = s N + n (N = 8), and u written over it: for each l every M tile's
// product is formed (all reads of A's columns l^2 .. l^2 + 2l) before the
// tile results are stored into those same columns, and a wavefront's LDS
// accesses complete in issue order.  The W_l operand is read from w per
// step instead of being held in registers.  Same products, same bits as
// u_mfma_fixed.
template <typename T, int L, int NR>
__device__ inline void u_mfma_fixed_inplace(T *Au, const T *w, const T *nrm_lm) {
  static_assert(NR % 8 == 0 && NR <= 24, "one or two M tiles of channels");
  typedef typename Mma<T>::V V;
  constexpr int L1 = L + 1, H = L1 * L1, N = 8, KS = NR / 4, MT = (NR + 15) / 16;
  const int lane = threadIdx.x;
  const int col = lane & 15, g = lane >> 4;
#pragma unroll 1
  for (int l = 0; l < L1; l++) {
    const int nl = 2 * l + 1, l2 = l * l;
    const int j = col;
    const bool jok = j < nl;
    V acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      const int c = mt * 16 + col;
      const bool crow = c < NR;
      acc[mt] = V{0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int c2 = 4 * ks + g;
        const T fac = c == c2 ? T(2) : T(1);
        const T a = crow ? fac * w[b2_pair_index(c, c2, NR) * L1 + l] : T(0);
        const T b = jok ? Au[((c2 >> 3) * H + l2 + j) * N + (c2 & 7)] : T(0);
        acc[mt] = Mma<T>::run(a, b, acc[mt]);
      }
    }
    if (jok) {
      const T sc = nrm_lm[l2 + j];
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int cr = mt * 16 + Mma<T>::row(g, q);
          if (cr < NR) Au[((cr >> 3) * H + l2 + j) * N + (cr & 7)] = acc[mt][q] * sc;
        }
    }
  }
}

namespace {

// Adjoint contraction for one pair, everything recomputed in registers.
// LFIX >= 0: l_max fixed at compile time; EXACT: n_max == NMAX.
template <typename T, int NMAX, int LFIX, bool EXACT>
__device__ inline void adjoint_pair(const B2Dev &p, T r, T x, T y, T z,
                                    T rcut, const T *__restrict__ ub,
                                    const T *__restrict__ ynorm, T &f0, T &f1,
                                    T &f2) {
  const int N = EXACT ? NMAX : p.N;
  const T rinv = T(1) / r;
  const T ux = x * rinv, uy = y * rinv, uz = z * rinv;
  T g[NMAX], dg[NMAX], v[NMAX];
  radial_basis<T, NMAX, EXACT>(r, rcut, (T)p.r1, N, p.cutoff_kind, g, dg);
#pragma unroll
  for (int n = 0; n < NMAX; n++) v[n] = 0;
  T a0 = 0, a1 = 0, a2 = 0;
  ylm_visit<T, true, LFIX>(ux, uy, uz, rinv, p.L, ynorm,
                     [&](int lm, T Y, T Yx, T Yy, T Yz) {
                       T q = 0;
                       const T *ul = ub + lm * N;
#pragma unroll
                       for (int n = 0; n < NMAX; n++) {
                         if (EXACT || n < N) {
                           T un = ul[n];
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
    if (EXACT || n < N) fr += v[n] * dg[n];
  f0 = a0 + fr * ux;
  f1 = a1 + fr * uy;
  f2 = a2 + fr * uz;
}

// The same contraction with l_max = L and n_max = NMAX fixed at compile time
// and the harmonics left unnormalised: with u' = u N_l|m| (see u_mfma) and
// Ytilde = Q_l^|m|(u_z) {Re,Im}(u_x + i u_y)^|m|,
//   v_n = sum_lm u'_n,lm Ytilde_lm,   q'_lm = sum_n u'_n,lm g_n,
// and the harmonic gradient sum_lm q_lm grad Y_lm = (G - u (u.G)) / r with
// G = sum_lm q'_lm dYtilde/du: for m = 0 dYtilde/du = (0, 0, Q'),
// cosine type (m Q C_{m-1}, -m Q S_{m-1}, Q' C_m), sine type
// (m Q S_{m-1}, m Q C_{m-1}, Q' S_m), Q' = dQ/du_z.  The tangential
// projection is applied once per pair instead of once per harmonic.
template <typename T>
__device__ inline void load_u_row8(const T *p, T *out);
template <>
__device__ inline void load_u_row8<float>(const float *p, float *out) {
  const float4 a = reinterpret_cast<const float4 *>(p)[0];
  const float4 b = reinterpret_cast<const float4 *>(p)[1];
  out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
  out[4] = b.x; out[5] = b.y; out[6] = b.z; out[7] = b.w;
}
template <>
__device__ inline void load_u_row8<double>(const double *p, double *out) {
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const double2 a = reinterpret_cast<const double2 *>(p)[h];
    out[2 * h] = a.x;
    out[2 * h + 1] = a.y;
  }
}

template <typename T, int L>
__device__ inline void adjoint_pair_fixed(const B2Dev &p, T r, T x, T y, T z,
                                          T rcut, const T *__restrict__ ub,
                                          T &f0, T &f1, T &f2) {
  constexpr int NM = 8;
  const T rinv = T(1) / r;
  const T ux = x * rinv, uy = y * rinv, uz = z * rinv;
  T g[NM], dg[NM], v[NM];
  radial_basis<T, NM, true>(r, rcut, (T)p.r1, NM, p.cutoff_kind, g, dg);
#pragma unroll
  for (int n = 0; n < NM; n++) v[n] = 0;
  T G0 = 0, G1 = 0, G2 = 0;
  T Cm = 1, Sm = 0, Cm1 = 0, Sm1 = 0, dfact = 1;
  auto term = [&](int lm, T Yt, T &qo) {
    T u[NM];
    load_u_row8<T>(ub + lm * NM, u);
    T q = u[0] * g[0];
#pragma unroll
    for (int n = 1; n < NM; n++) q += u[n] * g[n];
#pragma unroll
    for (int n = 0; n < NM; n++) v[n] += u[n] * Yt;
    qo = q;
  };
#pragma unroll
  for (int m = 0; m <= L; m++) {
    T q1 = 0, q2 = 0, p1 = 0, p2 = 0;
#pragma unroll
    for (int l = m; l <= L; l++) {
      T Q, P;
      if (l == m) {
        Q = dfact;
        P = 0;
      } else if (l == m + 1) {
        Q = T(2 * m + 1) * uz * dfact;
        P = dfact * T(2 * m + 1);
      } else {
        Q = (T(2 * l - 1) * uz * q1 - T(l + m - 1) * q2) * (T(1) / T(l - m));
        if (l == m + 2)
          P = T(2 * m + 3) * uz * p1;
        else
          P = (T(2 * l - 1) * uz * p1 - T(l + m) * p2) * (T(1) / T(l - m - 1));
      }
      q2 = q1;
      q1 = Q;
      p2 = p1;
      p1 = P;
      const int base = l * l + l;
      T qc;
      if (m == 0) {
        term(base, Q, qc);
        G2 += qc * P;
      } else {
        const T Qf = Q * T(m);
        term(base + m, Q * Cm, qc);  // cosine type
        T t = qc * Qf;
        G0 += t * Cm1;
        G1 -= t * Sm1;
        G2 += qc * P * Cm;
        term(base - m, Q * Sm, qc);  // sine type
        t = qc * Qf;
        G0 += t * Sm1;
        G1 += t * Cm1;
        G2 += qc * P * Sm;
      }
    }
    const T cn = Cm * ux - Sm * uy;
    const T sn = Sm * ux + Cm * uy;
    Cm1 = Cm;
    Sm1 = Sm;
    Cm = cn;
    Sm = sn;
    dfact *= T(2 * m + 1);
  }
  T fr = 0;
#pragma unroll
  for (int n = 0; n < NM; n++) fr += v[n] * dg[n];
  const T uG = ux * G0 + uy * G1 + uz * G2;
  f0 = fr * ux + rinv * (G0 - ux * uG);
  f1 = fr * uy + rinv * (G1 - uy * uG);
  f2 = fr * uz + rinv * (G2 - uz * uG);
}

}  // namespace

// One wavefront per atom.  Dynamic LDS (T): A[W], w[F], then u[W] sharing
// max(W, F) values with the X row x[F] (x is read only before w is formed,
// u written only after), n[H], then int ptab[nr * nr] unless the u
// contraction's channel-pair offsets come from registers.  The l_max 6
// instances (8, 16 or 24 channels) instead hold A in u's layout and
// overwrite it with u, and keep the X row in registers (BF16 rows packed, 34
// registers per lane; F <= 2,100 there): A/u[W], w[F], n[H], 13.3 KB in FP32
// at F = 2,100, so that a
// CU holds twelve wavefronts at 168 registers, FP64 six at 248 (eight and
// four with a separate u and 18 / 36 KB of LDS).  The kernel's time follows
// the resident wavefronts: 3.62 / 4.05 / 5.39 ms at eight / seven / six per
// CU, 2.47 ms at twelve, for 262,144 atoms in the mixed form.
// WM 3 is WM 1 with X given as BF16 hi + lo rows (the GEMM's operand).
// WM selects where w = dE_i/dB_i comes from:
//   0  precomputed (sparse-GP path);
//   1  mapped power 2 from the GEMM rows X and Z = beta X: E = X.Z/|X|^2,
//      w = 2 (Z - E X) / (|B| |X|)   (X = B, or B/|B| when normalized);
//   2  mapped power 1 from X and beta_s: E = X.beta/|X|,
//      w = (beta - E B/|B|) / |B|    (lammps_descriptor.cpp:262-270).
// Environments with |B|^2 < 1e-8 carry no energy (pair_flare.cpp:135).
// FP32 instances up to l_max 4 are held to 128 registers (four wavefronts
// per SIMD, 124 used at l_max 4 without spills): the latency of the
// prologue's loads is then hidden by more resident atoms (79 vs 94 us per
// launch at 32,000 atoms, B2 {2,8,4}).  l_max 6 would spill ~50.
constexpr int force_waves_per_eu(bool f32, int lfix) {
  return !f32 ? (lfix == 6 ? 2 : 1) : lfix <= 4 ? 4 : lfix == 6 ? 3 : 1;
}

template <typename T, int MODE, int NMAX, int WM, int LFIX, bool EXACT>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(force_waves_per_eu(sizeof(T) == 4, LFIX))))
force_kernel(B2Dev p, NListDev nl, const int *__restrict__ species,
             const int *__restrict__ row_of, const T *__restrict__ ynorm,
             const T *__restrict__ A_in, WSource<T> src,
             const double *__restrict__ dA, int64_t dA_p0, int a0,
             T *__restrict__ pair_f) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int i = a0 + blockIdx.x;
  const int lane = threadIdx.x;
  const int H = p.H, NH = p.NH, W = p.W, F = p.F;
  // the fixed-size l_max 6 instances (8, 16 or 24 channels, checked at
  // launch): A staged in u's layout and overwritten by u
  constexpr bool kOverlay = MODE == 1 && LFIX == 6 && EXACT && NMAX == 8;
  T *A = reinterpret_cast<T *>(smem);
  T *w = A + W;
  T *u = kOverlay ? A : w + F;
  T *xs = u;  // X row (rebuilt from hi + lo when BF16), other instances
  T *nrm_lm = kOverlay ? w + F : u + (W > F ? W : F);  // harmonic normalisation per lm
  int *ptab = reinterpret_cast<int *>(nrm_lm + H);  // nr x nr feature offsets
  const T *Ai = A_in + (int64_t)i * W;
  const int row = row_of[i];
  constexpr bool kFixed = MODE == 1 && LFIX >= 0 && EXACT && NMAX == 8;
  const int b = nl.cum[i], e = nl.cum[i + 1];
  // Every global read of the prologue is issued in batches before its
  // values are stored, so that a wavefront waits for memory a few times
  // rather than once per element.  Waits on loads are in issue order, so in
  // the mixed form the x / z rows and the harmonic normalisation are issued
  // ahead of the A row and are in registers when its stores wait.
  // (not in the l_max 6 instances, whose 168-register budget it would
  // exceed: their X row path serves every F)
  const bool mixed_fast = !kOverlay && WM == 3 && (F & 3) == 0 && F <= 16 * kWave;
  uint2 th[4], tl[4];
  float4 zv[4];
  double n2_row = 0;
  if (WM == 3 && mixed_fast) {
    n2_row = src.norm2[row];
    const uint2 *xh = reinterpret_cast<const uint2 *>(src.xh + (int64_t)row * src.kp);
    const uint2 *xl = reinterpret_cast<const uint2 *>(src.xl + (int64_t)row * src.kp);
    const float4 *z4 = reinterpret_cast<const float4 *>(src.Z + (int64_t)row * F);
    const int F4 = F >> 2;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int q = lane + j * kWave;
      th[j] = uint2{0, 0};
      tl[j] = uint2{0, 0};
      zv[j] = float4{0, 0, 0, 0};
      if (q < F4) {
        th[j] = xh[q];
        tl[j] = xl[q];
        zv[j] = z4[q];
      }
    }
  }
  T ynv = 0;
  if (kFixed && lane < H) {
    int l = 0;
    while ((l + 1) * (l + 1) <= lane) l++;
    const int m = lane - l * l - l;
    ynv = ynorm[l * (l + 1) / 2 + (m < 0 ? -m : m)];
  }
  // the A row in one unrolled batch of KA 16-byte loads per lane, KA sized
  // for 16 channels (registers are what bounds this kernel's occupancy)
  constexpr int KA = kFixed ? (16 * (LFIX + 1) * (LFIX + 1) * (int)sizeof(T) + 16 * kWave - 1) /
                                  (16 * kWave)
                            : 1;
  const int nva = W * (int)sizeof(T) / 16;
  const bool a_batch = kFixed && !kOverlay && ((W * (int)sizeof(T)) & 15) == 0 && nva <= KA * kWave &&
                       ((reinterpret_cast<uintptr_t>(Ai) | reinterpret_cast<uintptr_t>(A)) & 15) == 0;
  uint4 ta[KA];
  if (kFixed && a_batch) {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(Ai);
#pragma unroll
    for (int j = 0; j < KA; j++) {
      const int o = lane + j * kWave;
      ta[j] = uint4{0, 0, 0, 0};
      if (o < nva) ta[j] = s4[o];
    }
  }
  // nr <= 32 here (checked at launch); the fixed-size path (8, 16 or 24
  // channels) computes its channel-pair offsets in registers
  constexpr bool kUFixed = kFixed;
  if (!kUFixed || (p.nr != 8 && p.nr != 16 && p.nr != 24))
    for (int q = lane; q < p.nr * p.nr; q += kWave) {
      const int c = q / p.nr, c2 = q - c * p.nr;
      ptab[q] = b2_pair_index(c, c2, p.nr) * (p.L + 1);
    }
  if constexpr (kOverlay) {
    // A[c][lm] into u's layout: (c / 8 H + lm) 8 + c % 8
    constexpr int HF = (LFIX + 1) * (LFIX + 1);
    for (int base = 0; base < W; base += 8 * kWave) {
      T t[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int o = base + lane + j * kWave;
        t[j] = o < W ? Ai[o] : T(0);
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int o = base + lane + j * kWave;
        if (o < W) {
          const int c = o / HF, lm = o - c * HF;
          A[((c >> 3) * HF + lm) * 8 + (c & 7)] = t[j];
        }
      }
    }
  } else if (kFixed && a_batch) {
    uint4 *d4 = reinterpret_cast<uint4 *>(A);
#pragma unroll
    for (int j = 0; j < KA; j++) {
      const int o = lane + j * kWave;
      if (o < nva) d4[o] = ta[j];
    }
  } else {
    stage_lds_vec<4>(W, Ai, A);
  }
  if (kFixed && lane < H) nrm_lm[lane] = ynv;
  if (WM == 0) {
    const T *wi = src.w + (int64_t)row * F;
    stage_lds_vec<4>(F, wi, w);
  } else if (WM == 3 && mixed_fast) {
    // mixed form, rows of up to 1,024 features: each lane keeps its 4 x 4
    // features of x (= hi + lo) and z in registers through E and w
    const double n2 = n2_row;
    const int F4 = F >> 2;
    float xv[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      xv[j][0] = __uint_as_float(th[j].x << 16) + __uint_as_float(tl[j].x << 16);
      xv[j][1] = __uint_as_float(th[j].x & 0xffff0000u) + __uint_as_float(tl[j].x & 0xffff0000u);
      xv[j][2] = __uint_as_float(th[j].y << 16) + __uint_as_float(tl[j].y << 16);
      xv[j][3] = __uint_as_float(th[j].y & 0xffff0000u) + __uint_as_float(tl[j].y & 0xffff0000u);
    }
    // four-term FP32 partial sums, accumulated in FP64
    double d = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
      d += (double)(xv[j][0] * zv[j].x + xv[j][1] * zv[j].y + xv[j][2] * zv[j].z +
                    xv[j][3] * zv[j].w);
    d = wave_sum(d);
    double E = 0, sx = 0, sz = 0;
    if (!(n2 < 1e-8)) {
      const double bn = sqrt(n2);
      E = src.normalized ? d : d / n2;
      sz = src.normalized ? 2.0 / bn : 2.0 / n2;
      sx = -E * sz;
    }
    const float szf = (float)sz, sxf = (float)sx;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int q = lane + j * kWave;
      if (q < F4) {
        w[4 * q + 0] = szf * zv[j].x + sxf * xv[j][0];
        w[4 * q + 1] = szf * zv[j].y + sxf * xv[j][1];
        w[4 * q + 2] = szf * zv[j].z + sxf * xv[j][2];
        w[4 * q + 3] = szf * zv[j].w + sxf * xv[j][3];
      }
    }
    if (lane == 0) src.E_row[row] = E;
  } else {
    const double n2 = src.norm2[row];
    // kOverlay: the X row in registers, features 2 q and 2 q + 1 for
    // q = lane + j kWave (loaded here, ahead of the z row): with 8, 16 or 24
    // channels at l_max 6, F <= 2,100 <= 2 kXPairs kWave
    constexpr int kXPairs = 17;
    constexpr bool x_regs = kOverlay;
    // BF16 rows stay packed (hi and lo of a feature pair in two registers,
    // decoded where used); T rows as values
    constexpr bool kPacked = WM == 3;
    uint32_t xh2[kOverlay && kPacked ? kXPairs : 1], xl2[kOverlay && kPacked ? kXPairs : 1];
    T xv[kOverlay && !kPacked ? 2 * kXPairs : 1];
    auto xpair = [&](int j, T &x0, T &x1) {
      if constexpr (kPacked) {
        x0 = (T)(__uint_as_float(xh2[j] << 16) + __uint_as_float(xl2[j] << 16));
        x1 = (T)(__uint_as_float(xh2[j] & 0xffff0000u) + __uint_as_float(xl2[j] & 0xffff0000u));
      } else {
        x0 = xv[2 * j];
        x1 = xv[2 * j + 1];
      }
    };
    if constexpr (x_regs) {
      if constexpr (kPacked) {
        const uint32_t *xh =
            reinterpret_cast<const uint32_t *>(src.xh + (int64_t)row * src.kp);
        const uint32_t *xl =
            reinterpret_cast<const uint32_t *>(src.xl + (int64_t)row * src.kp);
        const int F2 = (F + 1) >> 1;
#pragma unroll
        for (int j = 0; j < kXPairs; j++) {
          const int q = lane + j * kWave;
          xh2[j] = 0;
          xl2[j] = 0;
          if (q < F2) {
            xh2[j] = xh[q];
            xl2[j] = xl[q];
          }
        }
      } else {
        const T *x = src.X + (int64_t)row * F;
        if ((F & 1) == 0) {
          // even rows: each lane's feature pairs as one 2-vector load
          typedef typename std::conditional<sizeof(T) == 8, double2, float2>::type T2;
          const T2 *x2 = reinterpret_cast<const T2 *>(x);
          T2 t2[kXPairs];
#pragma unroll
          for (int j = 0; j < kXPairs; j++) {
            const int q = lane + j * kWave;
            t2[j].x = T(0);
            t2[j].y = T(0);
            if (2 * q < F) t2[j] = x2[q];
          }
#pragma unroll
          for (int j = 0; j < kXPairs; j++) {
            xv[2 * j] = t2[j].x;
            xv[2 * j + 1] = t2[j].y;
          }
        } else {
#pragma unroll
          for (int j = 0; j < kXPairs; j++) {
            const int f = 2 * (lane + j * kWave);
            xv[2 * j] = f < F ? x[f] : T(0);
            xv[2 * j + 1] = f + 1 < F ? x[f + 1] : T(0);
          }
        }
      }
    }
    if constexpr (!kOverlay) {
      if (WM == 3) {
        // pairs of BF16 values: rows are padded to an even length (kp)
        const uint32_t *xh = reinterpret_cast<const uint32_t *>(src.xh + (int64_t)row * src.kp);
        const uint32_t *xl = reinterpret_cast<const uint32_t *>(src.xl + (int64_t)row * src.kp);
        const int F2 = (F + 1) >> 1;
        for (int base = 0; base < F2; base += 8 * kWave) {
          uint32_t th[8], tl[8];
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int q = base + lane + j * kWave;
            th[j] = 0;
            tl[j] = 0;
            if (q < F2) {
              th[j] = xh[q];
              tl[j] = xl[q];
            }
          }
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int q = base + lane + j * kWave;
            if (q < F2) {
              const float v0 = __uint_as_float(th[j] << 16) + __uint_as_float(tl[j] << 16);
              const float v1 = __uint_as_float(th[j] & 0xffff0000u) +
                               __uint_as_float(tl[j] & 0xffff0000u);
              xs[2 * q] = (T)v0;
              if (2 * q + 1 < F) xs[2 * q + 1] = (T)v1;
            }
          }
        }
      } else {
        const T *x = src.X + (int64_t)row * F;
        stage_lds_vec<4>(F, x, xs);
      }
    }
    // z (= beta X, or beta_s for power 1) staged into w, turned into w below
    const T *z = WM != 2 ? src.Z + (int64_t)row * F
                         : src.beta1 + (int64_t)species[i] * F;
    stage_lds_vec<4>(F, z, w);
    __syncthreads();
    if (n2 < 1e-8) {
      for (int f = lane; f < F; f += kWave) w[f] = 0;
      if (lane == 0) src.E_row[row] = 0;
    } else {
      const double bn = sqrt(n2);
      double d = 0;
      if constexpr (kOverlay) {
#pragma unroll
        for (int j = 0; j < kXPairs; j++) {
          const int f = 2 * (lane + j * kWave);
          T x0, x1;
          xpair(j, x0, x1);
          if (f < F) d += (double)x0 * (double)w[f];
          if (f + 1 < F) d += (double)x1 * (double)w[f + 1];
        }
      } else {
        for (int f = lane; f < F; f += kWave) d += (double)xs[f] * (double)w[f];
      }
      d = wave_sum(d);
      double E, sx, sz;
      if (WM != 2) {
        E = src.normalized ? d : d / n2;
        sz = src.normalized ? 2.0 / bn : 2.0 / n2;
        sx = -E * sz;
      } else {
        E = src.normalized ? d : d / bn;
        sz = 1.0 / bn;
        sx = -E / bn * (src.normalized ? 1.0 : 1.0 / bn);
      }
      if constexpr (kOverlay) {
#pragma unroll
        for (int j = 0; j < kXPairs; j++) {
          const int f = 2 * (lane + j * kWave);
          T x0, x1;
          xpair(j, x0, x1);
          if (f < F) w[f] = (T)(sz * (double)w[f] + sx * (double)x0);
          if (f + 1 < F) w[f + 1] = (T)(sz * (double)w[f + 1] + sx * (double)x1);
        }
      } else {
        for (int f = lane; f < F; f += kWave)
          w[f] = (T)(sz * (double)w[f] + sx * (double)xs[f]);
      }
      if (lane == 0) src.E_row[row] = E;
    }
  }
  __syncthreads();
  // the first chunk's pair data, loaded here so that the u contraction's
  // matrix-core work covers its latency (loaded with the prologue's batch it
  // would hold nine registers through it, past the four-wavefront budget)
  double4 rv_next = {0, 0, 0, 0};
  int sj_next = 0;
  if (MODE == 1 && b + lane < e) {
    rv_next = reinterpret_cast<const double4 *>(nl.rel)[b + lane];
    sj_next = nl.nspec[b + lane];
  }
  bool done = false;
  if constexpr (kOverlay) {
    // u over A (p.nr is 8, 16 or 24 here)
    done = true;
    if (p.nr == 24)
      u_mfma_fixed_inplace<T, LFIX, 24>(A, w, nrm_lm);
    else if (p.nr == 16)
      u_mfma_fixed_inplace<T, LFIX, 16>(A, w, nrm_lm);
    else
      u_mfma_fixed_inplace<T, LFIX, 8>(A, w, nrm_lm);
  } else if constexpr (kUFixed) {
    done = true;
    if (p.nr == 16)
      u_mfma_fixed<T, LFIX, 16>(A, w, nrm_lm, u);
    else if (p.nr == 8)
      u_mfma_fixed<T, LFIX, 8>(A, w, nrm_lm, u);
    else if (p.nr == 24)
      u_mfma_fixed<T, LFIX, 24>(A, w, nrm_lm, u);
    else
      done = false;
  }
  if (!done) u_mfma<T, LFIX, kFixed, kFixed ? 8 : -1, MODE == 0>(p, A, w, nrm_lm, ptab, u);
  __syncthreads();
  const int si = species[i];
  if (MODE == 0) {
    // 16 lanes per pair, four pairs at a time; partial sums over the
    // pair's 3 x N x H derivative values, then a 16-lane reduction
    const int sub = lane & 15, grp = lane >> 4;
    for (int k0 = b; k0 < e; k0 += 4) {
      const int k = k0 + grp;
      double f0 = 0, f1 = 0, f2 = 0;
      bool inside = false;
      if (k < e) {
        const double *rv = nl.rel + 4 * (int64_t)k;
        const int sj = nl.nspec[k];
        inside = pair_inside(p, si, sj, rv);
        if (inside) {
          // u is stored channel-major here: u[sj] rows match dA/dr's (n, lm)
          const double *d = dA + (k - dA_p0) * 3 * NH;
          const T *ub = u + sj * NH;
          constexpr int kPer = 16;  // values per lane held in flight
          for (int o0 = sub; o0 < NH; o0 += 16 * kPer) {
            double d0[kPer], d1[kPer], d2[kPer];
#pragma unroll
            for (int t = 0; t < kPer; t++) {
              const int o = o0 + 16 * t;
              d0[t] = d1[t] = d2[t] = 0;
              if (o < NH) {
                d0[t] = d[o];
                d1[t] = d[NH + o];
                d2[t] = d[2 * NH + o];
              }
            }
#pragma unroll
            for (int t = 0; t < kPer; t++) {
              const int o = o0 + 16 * t;
              const double uo = o < NH ? (double)ub[o] : 0.0;
              f0 += d0[t] * uo;
              f1 += d1[t] * uo;
              f2 += d2[t] * uo;
            }
          }
        }
      }
#pragma unroll
      for (int off = 8; off > 0; off >>= 1) {
        f0 += __shfl_xor(f0, off, 16);
        f1 += __shfl_xor(f1, off, 16);
        f2 += __shfl_xor(f2, off, 16);
      }
      if (k < e && sub == 0) {
        pair_f[3 * (int64_t)k + 0] = f0;
        pair_f[3 * (int64_t)k + 1] = f1;
        pair_f[3 * (int64_t)k + 2] = f2;
      }
    }
  } else {
    // the first chunk's pair data was loaded before the u contraction
    for (int k = b + lane; k < e; k += kWave) {
      double4 rv = rv_next;
      int sj = sj_next;
      if (k != b + lane) {
        rv = reinterpret_cast<const double4 *>(nl.rel)[k];
        sj = nl.nspec[k];
      }
      const double rva[4] = {rv.x, rv.y, rv.z, rv.w};
      T f[3] = {0, 0, 0};
      if (pair_inside(p, si, sj, rva)) {
        T f0, f1, f2;
        if constexpr (kFixed)
          adjoint_pair_fixed<T, (LFIX >= 0 ? LFIX : 0)>(
              p, (T)rv.x, (T)rv.y, (T)rv.z, (T)rv.w,
              (T)p.cut[si * p.S + sj], u + sj * NH, f0, f1, f2);
        else
          adjoint_pair<T, NMAX, LFIX, EXACT>(p, (T)rv.x, (T)rv.y, (T)rv.z,
                                             (T)rv.w, (T)p.cut[si * p.S + sj],
                                             u + sj * NH, ynorm, f0, f1, f2);
        f[0] = f0;
        f[1] = f1;
        f[2] = f2;
      }
      pair_f[3 * (int64_t)k + 0] = f[0];
      pair_f[3 * (int64_t)k + 1] = f[1];
      pair_f[3 * (int64_t)k + 2] = f[2];
    }
  }
}

// kGatherLanes lanes per atom (kGatherBatch pairs per lane per batch).  Also
// the per-atom virial
// rows V_a[b][c] = -sum_k r_ak,b f_ak,c over the atom's own pairs (the
// pair_flare tally of (x_i - x_j) f_ij, pair_flare.cpp:160-165).
constexpr int kGatherLanes = 8, kGatherBatch = kWave / kGatherLanes;
constexpr int kGatherAtoms = 256 / kGatherLanes;  // atoms per workgroup
template <typename PF>
__global__ void __launch_bounds__(256)
gather_kernel(int n, const int *__restrict__ cum, const int *__restrict__ rev,
              const double *__restrict__ rel, const PF *__restrict__ pf,
              double *__restrict__ forces, const int *__restrict__ row_of,
              const double *__restrict__ e_row, double *__restrict__ part) {
  __shared__ double sh[kGatherAtoms][10];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = t / kGatherLanes, sub = t % kGatherLanes;
  double f0 = 0, f1 = 0, f2 = 0;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (a < n) {
    const int e = cum[a + 1];
    // kGatherBatch pairs per lane per batch, all loads issued before use
    for (int k0 = cum[a] + sub; k0 < e; k0 += kWave) {
      int r[kGatherBatch];
#pragma unroll
      for (int j = 0; j < kGatherBatch; j++) {
        const int k = k0 + kGatherLanes * j;
        r[j] = k < e ? rev[k] : -1;
      }
      double g[kGatherBatch][3], h[kGatherBatch][3];
      double4 rv[kGatherBatch];
#pragma unroll
      for (int j = 0; j < kGatherBatch; j++) {
        const int k = k0 + kGatherLanes * j;
        for (int c = 0; c < 3; c++) {
          g[j][c] = 0;
          h[j][c] = 0;
        }
        rv[j] = double4{0, 0, 0, 0};
        if (k < e) {
          for (int c = 0; c < 3; c++) g[j][c] = (double)pf[3 * (int64_t)k + c];
          rv[j] = reinterpret_cast<const double4 *>(rel)[k];
        }
        if (r[j] >= 0)
          for (int c = 0; c < 3; c++) h[j][c] = (double)pf[3 * (int64_t)r[j] + c];
      }
#pragma unroll
      for (int j = 0; j < kGatherBatch; j++) {
        f0 += g[j][0] - h[j][0];
        f1 += g[j][1] - h[j][1];
        f2 += g[j][2] - h[j][2];
        const double x[3] = {rv[j].y, rv[j].z, rv[j].w};
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
          for (int c = 0; c < 3; c++) v[3 * b + c] -= x[b] * g[j][c];
      }
    }
  }
#pragma unroll
  for (int off = kGatherLanes / 2; off > 0; off >>= 1) {
    f0 += __shfl_xor(f0, off, kGatherLanes);
    f1 += __shfl_xor(f1, off, kGatherLanes);
    f2 += __shfl_xor(f2, off, kGatherLanes);
#pragma unroll
    for (int q = 0; q < 9; q++) v[q] += __shfl_xor(v[q], off, kGatherLanes);
  }
  if (a < n && sub == 0 && forces) {
    forces[3 * (int64_t)a] = f0;
    forces[3 * (int64_t)a + 1] = f1;
    forces[3 * (int64_t)a + 2] = f2;
  }
  // per-block sums of the atoms' energies and virial rows, fixed order
  if (sub == 0) {
    const int la = threadIdx.x / kGatherLanes;
    sh[la][0] = a < n ? e_row[row_of[a]] : 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) sh[la][1 + q] = a < n ? v[q] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    double s = 0;
    for (int la = 0; la < kGatherAtoms; la++) s += sh[la][threadIdx.x];
    part[blockIdx.x * 10 + threadIdx.x] = s;
  }
}

// Column q of the gather partials summed by workgroup q (fixed thread
// partition and tree, so the result is deterministic), written straight to
// its output slots: q = 0 the energy, q = 1 + 3a + b the virial component
// V[a][b] = -sum r_a f_b (LAMMPS order xx yy zz xy xz yz) and the GP stress
// (xx xy xz yy yz zz of -(1/V) sum f_c r_c2, c2 >= c,
// normalized_dot_product.cpp:193-202).
__global__ void __launch_bounds__(256)
sums_cols_kernel(int nparts, const double *__restrict__ part, double inv_volume,
                 double *energy, double *virial6, double *stress6) {
  __shared__ double sh[256];
  const int q = blockIdx.x, t = threadIdx.x;
  double s = 0;
  constexpr int kU = 8;
  for (int b0 = t; b0 < nparts; b0 += 256 * kU) {
    double x[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int b = b0 + 256 * u;
      x[u] = b < nparts ? part[(int64_t)b * 10 + q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) s += x[u];
  }
  sh[t] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) sh[t] += sh[t + off];
    __syncthreads();
  }
  if (t != 0) return;
  const double v = sh[0];
  if (q == 0) {
    if (energy) energy[0] = v;
    return;
  }
  // virial slot of V[ab] (LAMMPS order) and stress slot (c2 >= c only)
  int vs = -1, ss = -1;
  switch (q - 1) {
    case 0: vs = 0; ss = 0; break;
    case 1: vs = 3; break;
    case 2: vs = 4; break;
    case 3: ss = 1; break;
    case 4: vs = 1; ss = 3; break;
    case 5: vs = 5; break;
    case 6: ss = 2; break;
    case 7: ss = 4; break;
    case 8: vs = 2; ss = 5; break;
  }
  if (virial6 && vs >= 0) virial6[vs] = v;
  if (stress6 && ss >= 0) stress6[ss] = v * inv_volume;
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


// ---------------------------------------------------------------------------
template <typename T, int MODE, int WM>
static void launch_forces_t(hipStream_t st, const B2Dev &p, const NListDev &nl,
                            const int *species, const int *row_of,
                            const T *ynorm, const void *A, const WSource<T> &src,
                            const double *dA, int64_t dA_p0, int a0, int a1,
                            T *pair_f) {
  // the fixed-size instances with 8, 16 or 24 channels keep the u
  // contraction's channel-pair offsets in registers (no ptab)
  const bool fixed_shape = MODE == 1 && p.N == 8 && (p.L == 3 || p.L == 4 || p.L == 6);
  const bool fixed_u = fixed_shape && (p.nr == 8 || p.nr == 16 || p.nr == 24);
  // and, at l_max 6, write u over A (LDS A/u, w, norms); l_max 6 with other
  // channel counts takes the general instance
  const bool overlay = fixed_u && p.L == 6;
  const size_t lds =
      overlay ? sizeof(T) * ((size_t)p.W + (size_t)p.F + (size_t)p.H)
              : sizeof(T) * ((size_t)p.W + (size_t)p.F + (size_t)std::max(p.W, p.F) +
                             (size_t)p.H) +
                    (fixed_u ? 0 : sizeof(int) * (size_t)p.nr * p.nr);
  FLARE_REQUIRE(lds <= 64 * 1024, "descriptor settings too large for the force kernel");
  FLARE_REQUIRE(p.nr <= 32, "at most 32 radial channels (n_species * n_max)");
#define FLARE_FORCE_LAUNCH(NM, LF, EX)                                        \
  hipLaunchKernelGGL((force_kernel<T, MODE, NM, WM, LF, EX>), dim3(a1 - a0),   \
                     dim3(kWave), lds, st, p, nl, species, row_of, ynorm,     \
                     (const T *)A, src, dA, dA_p0, a0, pair_f)
  // compile-time (n_max, l_max) variants of the common adjoint settings
  if (MODE == 1 && p.N == 8 && p.L == 3)
    FLARE_FORCE_LAUNCH(8, 3, true);
  else if (MODE == 1 && p.N == 8 && p.L == 4)
    FLARE_FORCE_LAUNCH(8, 4, true);
  else if (overlay)
    FLARE_FORCE_LAUNCH(8, 6, true);
  else if (p.N <= 8)
    FLARE_FORCE_LAUNCH(8, -1, false);
  else
    FLARE_FORCE_LAUNCH(kMaxN, -1, false);
#undef FLARE_FORCE_LAUNCH
}

template <typename T, int MODE>
static void launch_forces_w(hipStream_t st, const B2Dev &p, const NListDev &nl,
                            const int *species, const int *row_of,
                            const T *ynorm, const void *A, const WInput &wi,
                            const double *dA, int64_t dA_p0, int a0, int a1,
                            T *pair_f) {
  WSource<T> src{wi.normalized, (const T *)wi.X, (const T *)wi.Z,
                 (const T *)wi.beta1, (const T *)wi.w,
                 (const __bf16 *)wi.xh, (const __bf16 *)wi.xl, wi.kp,
                 wi.norm2, wi.E_row};
  if (wi.mode == 0)
    launch_forces_t<T, MODE, 0>(st, p, nl, species, row_of, ynorm, A, src, dA,
                                dA_p0, a0, a1, pair_f);
  else if (wi.mode == 3)
    launch_forces_t<T, MODE, 3>(st, p, nl, species, row_of, ynorm, A, src, dA,
                                dA_p0, a0, a1, pair_f);
  else if (wi.mode == 1)
    launch_forces_t<T, MODE, 1>(st, p, nl, species, row_of, ynorm, A, src, dA,
                                dA_p0, a0, a1, pair_f);
  else
    launch_forces_t<T, MODE, 2>(st, p, nl, species, row_of, ynorm, A, src, dA,
                                dA_p0, a0, a1, pair_f);
}

void launch_forces(hipStream_t st, int prec, int mode, const B2Dev &p,
                   const NListDev &nl, const int *species, const int *row_of,
                   const double *ynorm, const float *ynorm_f, const void *A,
                   const WInput &wi, const double *dA, int64_t dA_p0,
                   int a0, int a1, const PairForces &pair_f) {
  if (a1 <= a0) return;
  FLARE_REQUIRE(pair_f.f32 == (prec != kPrecF64), "pair forces in the form's precision");
  if (prec == kPrecF64) {
    double *pf = (double *)pair_f.p;
    if (mode == 0)
      launch_forces_w<double, 0>(st, p, nl, species, row_of, ynorm, A, wi, dA,
                                 dA_p0, a0, a1, pf);
    else
      launch_forces_w<double, 1>(st, p, nl, species, row_of, ynorm, A, wi, dA,
                                 dA_p0, a0, a1, pf);
  } else {
    FLARE_REQUIRE(mode == 1, "the plain (materialised) form runs in FP64 only");
    launch_forces_w<float, 1>(st, p, nl, species, row_of, ynorm_f, A, wi, dA,
                              dA_p0, a0, a1, (float *)pair_f.p);
  }
  FLARE_HIP_CHECK(hipGetLastError());
}

int gather_parts(int n) { return (n + kGatherAtoms - 1) / kGatherAtoms; }

__global__ void zero_results_kernel(double *energy, double *virial6, double *stress6) {
  const int t = threadIdx.x;
  if (energy && t == 0) energy[0] = 0.0;
  if (virial6 && t < 6) virial6[t] = 0.0;
  if (stress6 && t < 6) stress6[t] = 0.0;
}

void launch_gather(hipStream_t st, const NListDev &nl, const PairForces &pair_f,
                   double *forces, const int *row_of, const double *e_row,
                   double *part, double inv_volume, double *energy, double *virial6,
                   double *stress6) {
  if (nl.n == 0) {
    hipLaunchKernelGGL(zero_results_kernel, dim3(1), dim3(64), 0, st, energy, virial6,
                       stress6);
  } else {
    if (pair_f.f32)
      hipLaunchKernelGGL(gather_kernel<float>, dim3(gather_parts(nl.n)), dim3(256), 0, st,
                         nl.n, nl.cum, nl.rev, nl.rel, (const float *)pair_f.p, forces,
                         row_of, e_row, part);
    else
      hipLaunchKernelGGL(gather_kernel<double>, dim3(gather_parts(nl.n)), dim3(256), 0, st,
                         nl.n, nl.cum, nl.rev, nl.rel, (const double *)pair_f.p, forces,
                         row_of, e_row, part);
    hipLaunchKernelGGL(sums_cols_kernel, dim3(10), dim3(256), 0, st,
                       gather_parts(nl.n), part, inv_volume, energy, virial6, stress6);
  }
  FLARE_HIP_CHECK(hipGetLastError());
}

void launch_colsum(hipStream_t st, int n, int k, const double *x, double *out) {
  hipLaunchKernelGGL(colsum_kernel, dim3(k), dim3(1024), 0, st, n, k, x, out);
  FLARE_HIP_CHECK(hipGetLastError());
}


}  // namespace flare
