// flare_math.h -- per-pair math shared by every kernel: Chebyshev radial
// basis times cutoff, real spherical harmonics with Cartesian gradients,
// and the B2 index maps.  Host/device inline; compiles under g++ as well so
// the same code is unit-tested on the CPU.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define FLARE_HD __host__ __device__
#else
#define FLARE_HD
#endif

namespace flare {

constexpr int kMaxSpecies = 8;
constexpr int kMaxN = 16;    // radial functions per species
constexpr int kMaxL = 10;    // l_max
constexpr int kNormTableSize = (kMaxL + 1) * (kMaxL + 2) / 2;

enum CutoffKind { kCutoffQuadratic = 0, kCutoffCosine = 1 };
constexpr double kPi = 3.14159265358979323846;

// Chebyshev radial basis times the cutoff function, and its derivative in r
// (radial.cpp:138-169 chebyshev; radial.cpp:344-379 calculate_radial;
// cutoffs.cpp:44-70).  x = (r - r1) / (rcut - r1); T_0 = 1, T_1 = x,
// T_n = 2 x T_{n-1} - T_{n-2}; g_n = T_n f_c, dg_n/dr = T_n' f_c + T_n f_c'.
// The Cartesian gradient is dg_n/dr * r_vec / r.
// EXACT: N == NMAX (no per-n guard).
template <typename T, int NMAX = kMaxN, bool EXACT = false>
FLARE_HD inline void radial_basis(T r, T rcut, T r1, int N, int cutoff_kind,
                                  T *g, T *dg) {
  T fc = 0, dfc = 0;
  if (!(r > rcut)) {
    if (cutoff_kind == kCutoffQuadratic) {
      T rd = r - rcut;
      fc = rd * rd;
      dfc = 2 * rd;
    } else {
      T a = T(kPi) * r / rcut;
      fc = T(0.5) * (cos(a) + 1);
      dfc = -T(kPi) * sin(a) / (2 * rcut);
    }
  }
  T c = 1 / (rcut - r1);
  T x = (r - r1) * c;
  bool inside = !((r < r1) || (r > rcut));
  if (!inside) {
    fc = 0;
    dfc = 0;
  }
  T t2 = 0, t1 = 1, d2 = 0, d1 = 0;
  // unrolled to NMAX so that g/dg stay in registers; entries >= N untouched
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int n = 0; n < NMAX; n++) {
    if (EXACT || n < N) {
      T tv, dv;
      if (n == 0) {
        tv = 1;
        dv = 0;
      } else if (n == 1) {
        tv = x;
        dv = c;
      } else {
        tv = 2 * x * t1 - t2;
        dv = 2 * t1 * c + 2 * x * d1 - d2;
      }
      t2 = t1;
      t1 = tv;
      d2 = d1;
      d1 = dv;
      g[n] = tv * fc;
      if (dg) dg[n] = dv * fc + tv * dfc;
    }
  }
}

// Normalisation table of the real harmonics, index l(l+1)/2 + m (m >= 0):
// sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) times sqrt(2) for m > 0.
inline void ylm_norm_table(double *out) {
  for (int l = 0; l <= kMaxL; l++)
    for (int m = 0; m <= l; m++) {
      double ratio = 1;
      for (int k = l - m + 1; k <= l + m; k++) ratio /= k;
      double v = std::sqrt((2 * l + 1) / (4 * kPi) * ratio);
      if (m) v *= std::sqrt(2.0);
      out[l * (l + 1) / 2 + m] = v;
    }
}

// Real spherical harmonics in the convention of y_grad.cpp:667 (index
// l*l + l + m, m = -l..l; m > 0 ~ Re (x+iy)^m, m < 0 ~ Im (x+iy)^|m|, no
// Condon-Shortley phase): Y_lm = N_lm Q_l^|m|(u_z) {Re,Im}(u_x + i u_y)^|m|
// with Q_l^m = d^m P_l / dt^m.  Visits every (l, m) in order of m and then l
// and calls fn(index, Y, dY/dx, dY/dy, dY/dz); rinv = 1/|r|, (ux, uy, uz)
// the unit vector.  With GRAD false the gradient arguments are zero.
// LFIX >= 0 fixes l_max at compile time (the loops unroll and the recurrence
// divisions become multiplications by constants); LFIX < 0 reads Lrt.
template <typename T, bool GRAD, int LFIX = -1, typename Fn>
FLARE_HD inline void ylm_visit(T ux, T uy, T uz, T rinv, int Lrt,
                               const T *norm, Fn &&fn) {
  const int L = LFIX >= 0 ? LFIX : Lrt;
  T Cm = 1, Sm = 0;    // (ux + i uy)^m
  T Cm1 = 0, Sm1 = 0;  // (ux + i uy)^(m-1)
  T dfact = 1;         // (2m-1)!!
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int m = 0; m <= L; m++) {
    T q1 = 0, q2 = 0;  // Q_{l-1}^m, Q_{l-2}^m
    T p1 = 0, p2 = 0;  // Q_{l-1}^{m+1}, Q_{l-2}^{m+1}
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int l = m; l <= L; l++) {
      T q, p;
      if (l == m) {
        q = dfact;
        p = 0;
      } else if (l == m + 1) {
        q = T(2 * m + 1) * uz * dfact;
        p = dfact * T(2 * m + 1);  // (2(m+1)-1)!!
      } else {
        if (LFIX >= 0) {
          q = (T(2 * l - 1) * uz * q1 - T(l + m - 1) * q2) * (T(1) / T(l - m));
          if (l == m + 2)
            p = T(2 * m + 3) * uz * p1;
          else
            p = (T(2 * l - 1) * uz * p1 - T(l + m) * p2) * (T(1) / T(l - m - 1));
        } else {
          q = (T(2 * l - 1) * uz * q1 - T(l + m - 1) * q2) / T(l - m);
          if (l == m + 2)
            p = T(2 * m + 3) * uz * p1;
          else
            p = (T(2 * l - 1) * uz * p1 - T(l + m) * p2) / T(l - m - 1);
        }
      }
      q2 = q1;
      q1 = q;
      p2 = p1;
      p1 = p;
      T nrm = norm[l * (l + 1) / 2 + m];
      int base = l * l + l;
      if (m == 0) {
        T Y = nrm * q;
        if (GRAD) {
          T s = nrm * rinv * p;
          fn(base, Y, -ux * uz * s, -uy * uz * s, (1 - uz * uz) * s);
        } else {
          fn(base, Y, T(0), T(0), T(0));
        }
      } else {
        T fm = T(m);
        {  // cosine type, m > 0
          T A = Cm;
          T Y = nrm * q * A;
          if (GRAD) {
            T d0 = q * fm * Cm1, d1 = -q * fm * Sm1, d2 = p * A;
            T udg = fm * q * A + uz * p * A;
            T s = nrm * rinv;
            fn(base + m, Y, s * (d0 - ux * udg), s * (d1 - uy * udg),
               s * (d2 - uz * udg));
          } else {
            fn(base + m, Y, T(0), T(0), T(0));
          }
        }
        {  // sine type, m < 0
          T A = Sm;
          T Y = nrm * q * A;
          if (GRAD) {
            T d0 = q * fm * Sm1, d1 = q * fm * Cm1, d2 = p * A;
            T udg = fm * q * A + uz * p * A;
            T s = nrm * rinv;
            fn(base - m, Y, s * (d0 - ux * udg), s * (d1 - uy * udg),
               s * (d2 - uz * udg));
          } else {
            fn(base - m, Y, T(0), T(0), T(0));
          }
        }
      }
    }
    T cn = Cm * ux - Sm * uy;
    T sn = Sm * ux + Cm * uy;
    Cm1 = Cm;
    Sm1 = Sm;
    Cm = cn;
    Sm = sn;
    dfact *= T(2 * m + 1);
  }
}

// B2 packing (b2.cpp:209-233): channel pair c1 <= c2 of nr channels maps to
// p = c1 (2 nr - c1 + 1) / 2 + (c2 - c1); feature = p (L+1) + l.
FLARE_HD inline int b2_pair_index(int c1, int c2, int nr) {
  if (c1 > c2) {
    int t = c1;
    c1 = c2;
    c2 = t;
  }
  return c1 * (2 * nr - c1 + 1) / 2 + (c2 - c1);
}

// Feature count and single-bond width (b2.cpp:195; pair_flare.cpp:330-331).
FLARE_HD inline int b2_n_descriptors(int S, int N, int L) {
  int nr = S * N;
  return nr * (nr + 1) / 2 * (L + 1);
}

// Power-of-two scale of one row of the FP16 hi + lo operand splits: for the
// row's largest magnitude m, the exponent e with m 2^e in [0.5, 1) (in
// [1, 2) for m >= 2^127, whose inverse scale 2^128 a float cannot hold).
// A row whose m is below the smallest normal float (all zero, or FP64
// values that are denormal or zero as floats) would need 2^e beyond the
// float range: it is reported as a zero row (returns false, e = 0) and
// stored as zeros.
FLARE_HD inline bool f16_row_scale(float m, int &e) {
  e = 0;
  if (!(m >= 1.17549435e-38f)) return false;  // FLT_MIN
  std::frexp(m, &e);                          // m = f 2^e, f in [0.5, 1)
  e = e > 127 ? -127 : -e;
  return true;
}

// l of harmonic index lm = l*l + l + m
FLARE_HD inline int lm_to_l(int lm) {
  int l = 0;
  while ((l + 1) * (l + 1) <= lm) l++;
  return l;
}

}  // namespace flare
