// small_mma.h -- per-wavefront matrix-core helpers shared by the B2
// forward kernel and the force kernel: v_mfma_*_16x16x4 for FP32 / FP64 with
// their result layouts, and the Gram-matrix B2 contraction.
#pragma once

#include "common.h"

namespace flare {

typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) double f64x4;

template <typename T>
struct Mma;
template <>
struct Mma<float> {
  typedef f32x4 V;
  static __device__ inline V run(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row of result element j held by a lane of row group g = lane / 16
  static __device__ inline int row(int g, int j) { return 4 * g + j; }
};
template <>
struct Mma<double> {
  typedef f64x4 V;
  static __device__ inline V run(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // the FP64 instruction interleaves result rows across elements (gfx950)
  static __device__ inline int row(int g, int j) { return g + 4 * j; }
};

// B2 on the matrix cores: for every l the block of features is the Gram
// matrix of the single-bond rows, B_l[c1][c2] = sum_m A[c1, l^2+m] A[c2, l^2+m]
// (M = N = channels, K = 2l+1), upper triangle c1 <= c2 scattered to
// p(c1, c2) (L+1) + l in Bs.  Returns this lane's share of |B|^2.
template <typename T, int MT, int LFIX>
__device__ inline double b2_gram_mfma(const B2Dev &p, const T *Al, T *Bs) {
  typedef typename Mma<T>::V V;
  const int lane = threadIdx.x;
  const int col = lane & 15, g = lane >> 4;
  const int H = p.H, nr = p.nr;
  const int L = LFIX >= 0 ? LFIX : p.L;
  const int L1 = L + 1;
  double nrm = 0;
#pragma unroll
  for (int t1 = 0; t1 < MT; t1++)
#pragma unroll
    for (int t2 = t1; t2 < MT; t2++) {
      const int ca = t1 * 16 + col, cb = t2 * 16 + col;
      // feature offset p(c1, c2) (L+1) of this lane's four result entries
      // (upper triangle only, -1 elsewhere), fixed over l
      int fofs[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int c1 = t1 * 16 + Mma<T>::row(g, q);
        fofs[q] = (c1 <= cb && cb < nr) ? b2_pair_index(c1, cb, nr) * L1 : -1;
      }
      for (int l = 0; l <= L; l++) {
        const int K = 2 * l + 1, l2 = l * l;
        V acc = V{0, 0, 0, 0};
        for (int k0 = 0; k0 < K; k0 += 4) {
          const int k = k0 + g;
          T a = 0, b = 0;
          if (k < K) {
            if (ca < nr) a = Al[ca * H + l2 + k];
            if (cb < nr) b = Al[cb * H + l2 + k];
          }
          acc = Mma<T>::run(a, b, acc);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          if (fofs[q] >= 0) {
            const T v = acc[q];
            Bs[fofs[q] + l] = v;
            nrm += (double)v * (double)v;
          }
        }
      }
    }
  return nrm;
}

}  // namespace flare
