// mfma.hip -- split-precision matrix-core GEMM: the mapped contraction
// Z = D beta and the sparse-GP products D P_hat^T and C P_hat.
//
// Operands arrive pre-split into 16-bit hi + lo parts, K-padded with zeros
// to a multiple of 32:
//   Z (M x N, FP32) = (Ah + Al)(M x Kp) . (Bh + Bl)(N x Kp)^T,
// accumulating hi*hi + hi*lo + lo*hi (lo*lo is dropped).  BF16 parts keep
// ~16 bits (2^-16 relative per product); FP16 parts keep ~22 bits, and rows
// are pre-scaled by powers of two (exact) so that their largest element
// sits in [0.5, 1) away from the FP16 subnormal range; the store multiplies
// by the inverse row and column scales.  beta is symmetric, so its row n is
// column n and both operands are read with K contiguous.
//
// Tile 128 x 256 x 32 per workgroup of 256 threads: four wavefronts in
// 2 x 2, each 64 x 128 = 4 x 8 tiles of v_mfma_f32_16x16x32_bf16.  Global
// tiles move in 16-byte vectors and are prefetched into registers during
// the MFMAs of the current step; workgroup ids are remapped so that the
// column tiles of a row block share one XCD's L2.
//
// Operand layouts of v_mfma_f32_16x16x32_{bf16,f16} (64 lanes):
//   A[i][k]: lane = i + 16 (k / 8), element k % 8
//   B[k][j]: lane = j + 16 (k / 8), element k % 8
//   D[i][j]: lane = j + 16 (i / 4), element i % 4
#include <type_traits>

#include "common.h"
#include "flare_math.h"
#include "mfma.h"

namespace flare {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;

template <typename E>
struct Frag;
template <>
struct Frag<__bf16> {
  typedef bf16x8 V;
};
template <>
struct Frag<_Float16> {
  typedef f16x8 V;
};
__device__ inline __attribute__((ext_vector_type(4))) float mma32(bf16x8 a, bf16x8 b,
                                                                   __attribute__((ext_vector_type(4))) float c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline __attribute__((ext_vector_type(4))) float mma32(f16x8 a, f16x8 b,
                                                                   __attribute__((ext_vector_type(4))) float c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Up to kGemmMaxBatch independent products (the GemmProduct list) and the
// exclusive prefix of their tile counts.
struct GemmBatch {
  int nb;
  GemmProduct g[kGemmMaxBatch];
  int tile0[kGemmMaxBatch + 1];
};
typedef __attribute__((ext_vector_type(4))) float floatx4;

constexpr int kBM = 128, kBN = 256, kBK = 32, kLdk = kBK;
// LDS rows of 64 bytes (four 16-byte chunks, no padding); chunk c of row r
// is stored at chunk c ^ ((r >> 2) & 3).  Sixteen consecutive lanes then
// touch 16 distinct 16-byte bank groups both when a tile is stored (four
// rows written whole) and when MFMA operands are read (sixteen rows, one
// chunk each).
__device__ inline int swz(int r, int chunk) { return (chunk ^ ((r >> 2) & 3)) * 8; }
constexpr int kChunksA = kBM * kBK / 8 / 256;  // 16-byte chunks per thread
constexpr int kChunksB = kBN * kBK / 8 / 256;

struct Tile {
  uint4 a[2 * kChunksA];  // Ah, Al
  uint4 b[2 * kChunksB];  // Bh, Bl
};

template <typename E>
__device__ inline void load_tile(Tile &t, const E *__restrict__ Ah,
                                 const E *__restrict__ Al,
                                 const E *__restrict__ Bh,
                                 const E *__restrict__ Bl, int M, int N,
                                 int Kp, int m0, int n0, int k0) {
  const int tid = threadIdx.x;
  const uint4 z = {0, 0, 0, 0};
#pragma unroll
  for (int h = 0; h < kChunksA; h++) {
    const int c = tid + 256 * h;
    const int r = c >> 2, kc = (c & 3) * 8;
    const int g = m0 + r;
    const int64_t o = (int64_t)g * Kp + k0 + kc;
    t.a[h] = z;
    t.a[kChunksA + h] = z;
    if (g < M) {
      t.a[h] = *reinterpret_cast<const uint4 *>(Ah + o);
      t.a[kChunksA + h] = *reinterpret_cast<const uint4 *>(Al + o);
    }
  }
#pragma unroll
  for (int h = 0; h < kChunksB; h++) {
    const int c = tid + 256 * h;
    const int r = c >> 2, kc = (c & 3) * 8;
    const int g = n0 + r;
    const int64_t o = (int64_t)g * Kp + k0 + kc;
    t.b[h] = z;
    t.b[kChunksB + h] = z;
    if (g < N) {
      t.b[h] = *reinterpret_cast<const uint4 *>(Bh + o);
      t.b[kChunksB + h] = *reinterpret_cast<const uint4 *>(Bl + o);
    }
  }
}

template <typename E>
__device__ inline void store_tile(const Tile &t, E (*sAh)[kLdk],
                                  E (*sAl)[kLdk], E (*sBh)[kLdk],
                                  E (*sBl)[kLdk]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int h = 0; h < kChunksA; h++) {
    const int c = tid + 256 * h;
    const int r = c >> 2, kc = (c & 3) * 8;
    *reinterpret_cast<uint4 *>(&sAh[r][swz(r, kc >> 3)]) = t.a[h];
    *reinterpret_cast<uint4 *>(&sAl[r][swz(r, kc >> 3)]) = t.a[kChunksA + h];
  }
#pragma unroll
  for (int h = 0; h < kChunksB; h++) {
    const int c = tid + 256 * h;
    const int r = c >> 2, kc = (c & 3) * 8;
    *reinterpret_cast<uint4 *>(&sBh[r][swz(r, kc >> 3)]) = t.b[h];
    *reinterpret_cast<uint4 *>(&sBl[r][swz(r, kc >> 3)]) = t.b[kChunksB + h];
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap the linear
// id so that the column tiles of one row block run on the same XCD and
// share its L2 copy of the row block.
__device__ inline void tile_of_block(int n_row_tiles, int n_col_tiles, int &rt,
                                     int &ct) {
  const int nb = n_row_tiles * n_col_tiles;
  const int L = blockIdx.x;
  const int xcd = L & 7, slot = L >> 3;
  const int per = (nb + 7) >> 3;
  const int full = nb - (per - 1) * 8;  // XCDs that get `per` tiles
  int T;
  if (xcd < full)
    T = xcd * per + slot;
  else if (slot < per - 1)
    T = full * per + (xcd - full) * (per - 1) + slot;
  else
    T = nb;  // this XCD has one tile fewer: no tile for this slot
  rt = T / n_col_tiles;
  ct = T - rt * n_col_tiles;
}

template <typename E>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_x3_kernel(GemmBatch bt, const E *__restrict__ Ah0,
               const E *__restrict__ Al0, const E *__restrict__ Bh0,
               const E *__restrict__ Bl0, float *__restrict__ Z0) {
  typedef typename Frag<E>::V V8;
  __shared__ __align__(16) E sAh[kBM][kLdk], sAl[kBM][kLdk];
  __shared__ __align__(16) E sBh[kBN][kLdk], sBl[kBN][kLdk];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int wm = (wv >> 1) * 64, wn = (wv & 1) * 128;
  // global tile id over all products (column tiles of a row block adjacent)
  int T, dummy;
  tile_of_block(bt.tile0[bt.nb], 1, T, dummy);
  if (T >= bt.tile0[bt.nb]) return;
  int q = 0;
  while (T >= bt.tile0[q + 1]) q++;
  const GemmProduct &gp = bt.g[q];
  const int M = gp.M, N = gp.N, Kp = gp.K, ldz = gp.ldz;
  const int nct = (N + kBN - 1) / kBN;
  const int local = T - bt.tile0[q];
  const int rt = local / nct, ct = local - rt * nct;
  const E *Ah = Ah0 + gp.a_off, *Al = Al0 + gp.a_off;
  const E *Bh = Bh0 + gp.b_off, *Bl = Bl0 + gp.b_off;
  float *Z = Z0 + gp.z_off;
  const int m0 = rt * kBM, n0 = ct * kBN;
  floatx4 acc[4][8];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) acc[i][j] = floatx4{0, 0, 0, 0};
  const int fr = lane & 15;
  // operand rows are wm + 16 i + fr (and wn + 16 j + fr): swizzle by fr
  const int fk = swz(fr, lane >> 4);
  const int steps = Kp / kBK;
  Tile t;
  load_tile(t, Ah, Al, Bh, Bl, M, N, Kp, m0, n0, 0);
  for (int s = 0; s < steps; s++) {
    __syncthreads();
    store_tile(t, sAh, sAl, sBh, sBl);
    __syncthreads();
    if (s + 1 < steps) load_tile(t, Ah, Al, Bh, Bl, M, N, Kp, m0, n0, (s + 1) * kBK);
    V8 ah[4], al[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      ah[i] = *reinterpret_cast<const V8 *>(&sAh[wm + 16 * i + fr][fk]);
      al[i] = *reinterpret_cast<const V8 *>(&sAl[wm + 16 * i + fr][fk]);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const V8 bh = *reinterpret_cast<const V8 *>(&sBh[wn + 16 * j + fr][fk]);
      const V8 bl = *reinterpret_cast<const V8 *>(&sBl[wn + 16 * j + fr][fk]);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        acc[i][j] = mma32(al[i], bh, acc[i][j]);
        acc[i][j] = mma32(ah[i], bl, acc[i][j]);
        acc[i][j] = mma32(ah[i], bh, acc[i][j]);
      }
    }
  }
  const int dcol = lane & 15;
  const int drow = (lane >> 4) * 4;
  const bool vec4 = !gp.rs && !gp.cs && (N & 3) == 0 && (ldz & 3) == 0 && (gp.z_off & 3) == 0;
  if (vec4) {
    // 16 rows at a time through LDS (8 KB per wavefront in the B tile
    // buffers, free after the last step), then 16-byte stores: 512
    // contiguous bytes per row instead of 64
    __syncthreads();
    float *stage = reinterpret_cast<float *>(wv < 2 ? &sBh[0][0] : &sBl[0][0]) +
                   (wv & 1) * (16 * 128);
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
      for (int j = 0; j < 8; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) stage[(drow + q) * 128 + 16 * j + dcol] = acc[i][j][q];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const int f = lane + 64 * t;
        const int r = f >> 5, c = (f & 31) * 4;
        const int row = m0 + wm + 16 * i + r, col = n0 + wn + c;
        if (row < M && col < N)
          *reinterpret_cast<float4 *>(Z + (int64_t)row * ldz + col) =
              reinterpret_cast<const float4 *>(stage)[f];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  } else if (!gp.rs && !gp.cs) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int col = n0 + wn + 16 * j + dcol;
        if (col >= N) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = m0 + wm + 16 * i + drow + q;
          if (row < M) Z[(int64_t)row * ldz + col] = acc[i][j][q];
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = m0 + wm + 16 * i + drow + q;
        if (row >= M) continue;
        const float rsc = gp.rs ? gp.rs[row] : 1.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int col = n0 + wn + 16 * j + dcol;
          if (col < N)
            Z[(int64_t)row * ldz + col] = acc[i][j][q] * rsc * (gp.cs ? gp.cs[col] : 1.0f);
        }
      }
  }
}

// FP64 and FP32 Z = A B^T on the matrix cores (v_mfma_f64_16x16x4f64,
// v_mfma_f32_16x16x4f32): the mapped contraction Z = X beta of the FP64 and
// FP32 forms (beta is symmetric, so its rows are its columns).  A: M rows
// of K values, B: N rows of K values (row stride K, K a multiple of 16 bytes
// / element), Z: M x N with leading dimension ldz.
// Tile 128 x 128 per workgroup of four wavefronts (64 x 64 each, 4 x 4
// MFMA tiles), K steps of 16 double-buffered in LDS with one barrier per
// step.  Inside a step the K order is permuted the same way for both
// operands (lane group g = lane / 16 supplies k = 4 g + t in sub-step t),
// so a lane's four values of a row are contiguous: 16-byte LDS reads.
// LDS rows of 16 values are cut into 16-byte units, XOR-swizzled by row so
// that the sixteen rows a lane group reads fall in distinct bank groups.
typedef __attribute__((ext_vector_type(4))) double doublex4;
constexpr int kDM = 128, kDN = 128, kDK = 16;

template <typename T>
struct DMma;
template <>
struct DMma<double> {
  typedef doublex4 V;
  typedef double2 U;                           // one 16-byte unit
  static constexpr int kUnits = kDK / 2;       // units per LDS row
  static __device__ inline V run(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int g, int e) { return g + 4 * e; }  // gfx950 FP64
  static __device__ inline int swz(int r, int u) { return (u ^ ((r >> 1) & 7)) * 2; }
};
template <>
struct DMma<float> {
  typedef floatx4 V;
  typedef float4 U;
  static constexpr int kUnits = kDK / 4;
  static __device__ inline V run(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int g, int e) { return 4 * g + e; }
  static __device__ inline int swz(int r, int u) { return (u ^ ((r >> 2) & 3)) * 4; }
};

// Register staging of one K step: A rows (kDM) and B rows (TN, the tile's
// column count: 128, or 64 for the narrow tiles)
template <typename T, int TN>
struct DTile {
  static constexpr int kPer = kDM * DMma<T>::kUnits / 256;   // A units per thread
  static constexpr int kPerB = TN * DMma<T>::kUnits / 256;   // B units per thread
  typename DMma<T>::U a[kPer], b[kPerB];
};

template <typename T, int TN>
__device__ inline void dgemm_load(DTile<T, TN> &t, const T *__restrict__ A,
                                  const T *__restrict__ B, int M, int N, int K,
                                  int m0, int n0, int k0) {
  typedef typename DMma<T>::U U;
  constexpr int NU = DMma<T>::kUnits, EU = 16 / sizeof(T);
  const int tid = threadIdx.x;
  U z;
  memset(&z, 0, sizeof(U));
#pragma unroll
  for (int h = 0; h < DTile<T, TN>::kPer; h++) {
    const int c = tid + 256 * h;
    const int r = c / NU, k = k0 + EU * (c % NU);
    t.a[h] = z;
    if (k < K && m0 + r < M) t.a[h] = *reinterpret_cast<const U *>(A + (int64_t)(m0 + r) * K + k);
  }
#pragma unroll
  for (int h = 0; h < DTile<T, TN>::kPerB; h++) {
    const int c = tid + 256 * h;
    const int r = c / NU, k = k0 + EU * (c % NU);
    t.b[h] = z;
    if (k < K && n0 + r < N) t.b[h] = *reinterpret_cast<const U *>(B + (int64_t)(n0 + r) * K + k);
  }
}

template <typename T, int TN>
__device__ inline void dgemm_store(const DTile<T, TN> &t, T (*sA)[kDK], T (*sB)[kDK]) {
  typedef typename DMma<T>::U U;
  constexpr int NU = DMma<T>::kUnits;
  const int tid = threadIdx.x;
#pragma unroll
  for (int h = 0; h < DTile<T, TN>::kPer; h++) {
    const int c = tid + 256 * h;
    const int r = c / NU, u = c % NU;
    *reinterpret_cast<U *>(&sA[r][DMma<T>::swz(r, u)]) = t.a[h];
  }
#pragma unroll
  for (int h = 0; h < DTile<T, TN>::kPerB; h++) {
    const int c = tid + 256 * h;
    const int r = c / NU, u = c % NU;
    *reinterpret_cast<U *>(&sB[r][DMma<T>::swz(r, u)]) = t.b[h];
  }
}

// the lane's four values k = 4 g .. 4 g + 3 of LDS row r
template <typename T>
__device__ inline void dgemm_row4(const T (*s)[kDK], int r, int g, T *o) {
  if constexpr (sizeof(T) == 8) {
    const double2 lo = *reinterpret_cast<const double2 *>(&s[r][DMma<T>::swz(r, 2 * g)]);
    const double2 hi = *reinterpret_cast<const double2 *>(&s[r][DMma<T>::swz(r, 2 * g + 1)]);
    o[0] = lo.x;
    o[1] = lo.y;
    o[2] = hi.x;
    o[3] = hi.y;
  } else {
    const float4 v = *reinterpret_cast<const float4 *>(&s[r][DMma<T>::swz(r, g)]);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
}

// A batch of products with their fused epilogues (mode 0: plain store).
struct DGemmBatch {
  GemmBatch bt;
  DEpilogue ep[kGemmMaxBatch];
};

__device__ inline double powi_d(double t, double power) {
  if (power == 2.0) return t * t;
  if (power == 1.0) return t;
  return pow(t, power);
}

// TN = 128: four wavefronts of 64 x 64 (2 x 2); TN = 64 (narrow tiles, for
// products whose tile count would leave most of a second round of
// workgroups idle): four wavefronts of 32 x 64 stacked by rows, so every
// wavefront still covers one whole 64-column group of the epilogue.
template <typename T, int MODE, int TN>
__device__ inline void dgemm_body(const DGemmBatch &dbt, const T *__restrict__ A0,
                                  const T *__restrict__ B0, T *__restrict__ Z0) {
  typedef typename DMma<T>::V V;
  constexpr int IM = TN == 128 ? 4 : 2;  // 16-row blocks per wavefront
  const GemmBatch &bt = dbt.bt;
  __shared__ __align__(16) T sA[2][kDM][kDK], sB[2][TN][kDK];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int wm = TN == 128 ? (wv >> 1) * 64 : wv * 32, wn = TN == 128 ? (wv & 1) * 64 : 0;
  int Tt, dummy;
  tile_of_block(bt.tile0[bt.nb], 1, Tt, dummy);
  if (Tt >= bt.tile0[bt.nb]) return;
  int q = 0;
  while (Tt >= bt.tile0[q + 1]) q++;
  const GemmProduct &gp = bt.g[q];
  const int M = gp.M, N = gp.N, K = gp.K, ldz = gp.ldz;
  const int nct = (N + TN - 1) / TN;
  const int local = Tt - bt.tile0[q];
  const int rt = local / nct, ct = local - rt * nct;
  const T *A = A0 + gp.a_off, *B = B0 + gp.b_off;
  T *Z = Z0 + gp.z_off;
  const int m0 = rt * kDM, n0 = ct * TN;
  V acc[IM][4];
#pragma unroll
  for (int i = 0; i < IM; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = V{0, 0, 0, 0};
  const int fr = lane & 15, g = lane >> 4;
  // a lower-triangular B has nothing past the tile's last column
  const int Keff = MODE == DEpilogue::kSquares && dbt.ep[q].tri ? min(K, n0 + TN) : K;
  const int steps = (Keff + kDK - 1) / kDK;
  DTile<T, TN> t;
  dgemm_load<T, TN>(t, A, B, M, N, K, m0, n0, 0);
  dgemm_store<T, TN>(t, sA[0], sB[0]);
  if (steps > 1) dgemm_load<T, TN>(t, A, B, M, N, K, m0, n0, kDK);
  __syncthreads();
  for (int s = 0; s < steps; s++) {
    const int cur = s & 1;
    if (s + 1 < steps) {
      // step s + 1 into the other buffer (its last reader, step s - 1,
      // ended with a barrier); step s + 2 into registers
      dgemm_store<T, TN>(t, sA[cur ^ 1], sB[cur ^ 1]);
      if (s + 2 < steps) dgemm_load<T, TN>(t, A, B, M, N, K, m0, n0, (s + 2) * kDK);
    }
    T a[IM][4];
#pragma unroll
    for (int i = 0; i < IM; i++) dgemm_row4<T>(sA[cur], wm + 16 * i + fr, g, a[i]);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      T b[4];
      dgemm_row4<T>(sB[cur], wn + 16 * j + fr, g, b);
#pragma unroll
      for (int tt = 0; tt < 4; tt++)
#pragma unroll
        for (int i = 0; i < IM; i++) acc[i][j] = DMma<T>::run(a[i][tt], b[tt], acc[i][j]);
    }
    __syncthreads();
  }
  if constexpr (MODE == 0) {
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int col = n0 + wn + 16 * j + fr;
        if (col >= N) continue;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int row = m0 + wm + 16 * i + DMma<T>::row(g, e);
          if (row < M) Z[(int64_t)row * ldz + col] = acc[i][j][e];
        }
      }
  } else {
    // per row: the sum over this wavefront's 64 columns, reduced across the
    // 16 lanes (fr) that hold a row's columns
    const DEpilogue &ep = dbt.ep[q];
    double rs[IM][4];
    double inv_nb[IM][4];
    bool row_empty[IM][4];
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        rs[i][e] = 0;
        inv_nb[i][e] = 0;
        row_empty[i][e] = true;
        if constexpr (MODE == DEpilogue::kKernel) {
          const int row = m0 + wm + 16 * i + DMma<T>::row(g, e);
          if (row < M) {
            const double nb = sqrt(ep.norm2_rows[row]);
            row_empty[i][e] = nb < 1e-8;
            inv_nb[i][e] = nb;
          }
        }
      }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int col = n0 + wn + 16 * j + fr;
      double pn = 0, al = 0;
      if constexpr (MODE == DEpilogue::kKernel) {
        if (col < N) {
          pn = ep.pnorm[col];
          al = ep.alpha ? ep.alpha[col] : 0.0;
        }
      }
#pragma unroll
      for (int i = 0; i < IM; i++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int row = m0 + wm + 16 * i + DMma<T>::row(g, e);
          const double a = (double)acc[i][j][e];
          if constexpr (MODE == DEpilogue::kSquares) {
            if (col < N) rs[i][e] += a * a;
          } else {
            double kv = 0, cv = 0;
            if (col < N && !row_empty[i][e] && !(pn < 1e-8)) {
              const double tt = ep.normalized ? a : a / (inv_nb[i][e] * pn);
              const double tx = powi_d(tt, ep.power);
              kv = ep.sig2 * tx;
              cv = al * powi_d(tt, ep.power - 1.0);
              rs[i][e] += al * tx;
            }
            if (row < M && col < ep.ldk) {
              if (ep.kval) ep.kval[(int64_t)row * ep.ldk + col] = kv;
              reinterpret_cast<T *>(ep.cmat)[(int64_t)row * ep.ldk + col] = (T)cv;
            }
          }
        }
    }
#pragma unroll
    for (int i = 0; i < IM; i++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        double v = rs[i][e];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 16);
        const int row = m0 + wm + 16 * i + DMma<T>::row(g, e);
        // a 64-column group past N has no partial (its index may be nparts);
        // the row's last group also clears the groups up to nparts (the row
        // pitch may exceed this product's group count)
        if (fr == 0 && row < M && n0 + wn < N) {
          const int grp = (n0 + wn) / 64;
          double *pr = ep.part + (int64_t)row * ep.nparts;
          pr[grp] = v;
          if (grp == (N - 1) / 64)
            for (int x = grp + 1; x < ep.nparts; x++) pr[x] = 0.0;
        }
      }
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
dgemm_kernel(DGemmBatch dbt, const T *__restrict__ A0, const T *__restrict__ B0,
             T *__restrict__ Z0) {
  dgemm_body<T, MODE, 128>(dbt, A0, B0, Z0);
}

// 128 x 64 tiles, three workgroups per CU (48 KiB of LDS each in FP64)
template <typename T, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
dgemm_narrow_kernel(DGemmBatch dbt, const T *__restrict__ A0, const T *__restrict__ B0,
                    T *__restrict__ Z0) {
  dgemm_body<T, MODE, 64>(dbt, A0, B0, Z0);
}

template <typename T>
static void dgemm_batched_t(hipStream_t st, int nb, const GemmProduct *prods,
                            const DEpilogue *ep, const T *A, const T *B, T *Z) {
  FLARE_REQUIRE(nb >= 0 && nb <= kGemmMaxBatch, "gemm: at most 8 products");
  DGemmBatch dbt;
  GemmBatch &bt = dbt.bt;
  bt.nb = 0;
  int mode = 0;
  for (int q = 0; q < nb; q++) {
    const GemmProduct &g = prods[q];
    FLARE_REQUIRE((g.K * (int)sizeof(T)) % 16 == 0, "gemm: rows must be 16-byte multiples");
    FLARE_REQUIRE(!g.ld_ab, "gemm: K splits only on the wide-row products");
    if (g.M <= 0 || g.N <= 0) continue;
    if (ep) {
      const DEpilogue &e = ep[q];
      FLARE_REQUIRE(mode == 0 || e.mode == mode, "gemm: one epilogue mode per batch");
      mode = e.mode;
      FLARE_REQUIRE(e.part && e.nparts >= (g.N + 63) / 64, "gemm: partial-sum columns");
      FLARE_REQUIRE(e.mode != DEpilogue::kKernel ||
                        (e.cmat && e.ldk >= g.N && e.norm2_rows && e.pnorm),
                    "gemm: kernel-value epilogue outputs");
      dbt.ep[bt.nb] = e;
    }
    bt.g[bt.nb] = g;
    bt.nb++;
  }
  // Tile width: the 128 x 128 tiles run two workgroups per CU, the
  // 128 x 64 tiles three.  Rounds of workgroups over 256 CUs, each round of
  // narrow tiles doing 3/4 of a wide round's work per CU: narrow when that
  // is at least 1/8 less (products with a few hundred columns, whose wide
  // tiles leave most of a second round idle).
  int64_t wide = 0, narrow = 0;
  for (int q = 0; q < bt.nb; q++) {
    const int64_t rows = (bt.g[q].M + kDM - 1) / kDM;
    wide += rows * ((bt.g[q].N + kDN - 1) / kDN);
    narrow += rows * ((bt.g[q].N + 63) / 64);
  }
  const int64_t rounds_w = (wide + 511) / 512, rounds_n = (narrow + 767) / 768;
  const bool use_narrow = 8 * 3 * rounds_n <= 7 * 4 * rounds_w;
  const int TN = use_narrow ? 64 : kDN;
  int tiles = 0;
  for (int q = 0; q < bt.nb; q++) {
    bt.tile0[q] = tiles;
    tiles += (bt.g[q].M + kDM - 1) / kDM * ((bt.g[q].N + TN - 1) / TN);
  }
  bt.tile0[bt.nb] = tiles;
  if (tiles == 0) return;
  const int grid = (tiles + 7) / 8 * 8;
#define FLARE_DGEMM(KERN)                                                                \
  if (mode == 0)                                                                        \
    hipLaunchKernelGGL((KERN<T, 0>), dim3(grid), dim3(256), 0, st, dbt, A, B, Z);        \
  else if (mode == DEpilogue::kKernel)                                                  \
    hipLaunchKernelGGL((KERN<T, DEpilogue::kKernel>), dim3(grid), dim3(256), 0, st, dbt, \
                       A, B, Z);                                                        \
  else                                                                                  \
    hipLaunchKernelGGL((KERN<T, DEpilogue::kSquares>), dim3(grid), dim3(256), 0, st,     \
                       dbt, A, B, Z);
  if (use_narrow) {
    FLARE_DGEMM(dgemm_narrow_kernel)
  } else {
    FLARE_DGEMM(dgemm_kernel)
  }
#undef FLARE_DGEMM
  FLARE_HIP_CHECK(hipGetLastError());
}

void dgemm_epi_batched(hipStream_t st, int nb, const GemmProduct *prods,
                       const DEpilogue *ep, const double *A, const double *B) {
  dgemm_batched_t<double>(st, nb, prods, ep, A, B, nullptr);
}

void sgemm_epi_batched(hipStream_t st, int nb, const GemmProduct *prods,
                       const DEpilogue *ep, const float *A, const float *B) {
  dgemm_batched_t<float>(st, nb, prods, ep, A, B, nullptr);
}

void dgemm_batched(hipStream_t st, int nb, const GemmProduct *prods, const double *A,
                   const double *B, double *Z) {
  dgemm_batched_t<double>(st, nb, prods, nullptr, A, B, Z);
}

void sgemm_batched(hipStream_t st, int nb, const GemmProduct *prods, const float *A,
                   const float *B, float *Z) {
  dgemm_batched_t<float>(st, nb, prods, nullptr, A, B, Z);
}

// hi/lo split of a (rows x cols) FP64 matrix into rows of kp BF16 (zero
// padded); hi = bf16(x), lo = bf16(x - hi)
__global__ void split_bf16_kernel(int rows, int cols, int kp,
                                  const double *__restrict__ x,
                                  __bf16 *__restrict__ hi,
                                  __bf16 *__restrict__ lo) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)rows * kp) return;
  int r = (int)(q / kp), c = (int)(q - (int64_t)r * kp);
  float v = c < cols ? (float)x[(int64_t)r * cols + c] : 0.0f;
  __bf16 h = (__bf16)v;
  hi[q] = h;
  lo[q] = (__bf16)(v - (float)h);
}

int bf16_padded_k(int K) { return (K + kBK - 1) / kBK * kBK; }

template <typename E>
static void gemm_x3_batched(hipStream_t st, int nb, const GemmProduct *prods,
                            const void *Ah, const void *Al, const void *Bh,
                            const void *Bl, float *Z) {
  FLARE_REQUIRE(nb >= 0 && nb <= kGemmMaxBatch, "gemm_x3: at most 8 products");
  GemmBatch bt;
  bt.nb = 0;
  int tiles = 0;
  for (int q = 0; q < nb; q++) {
    const GemmProduct &g = prods[q];
    FLARE_REQUIRE(g.K % kBK == 0, "gemm_x3: K must be padded to 32");
    FLARE_REQUIRE(!g.ld_ab, "gemm_x3: K splits only on the wide-row products");
    if (g.M <= 0 || g.N <= 0) continue;
    bt.g[bt.nb] = g;
    bt.tile0[bt.nb] = tiles;
    tiles += (g.M + kBM - 1) / kBM * ((g.N + kBN - 1) / kBN);
    bt.nb++;
  }
  bt.tile0[bt.nb] = tiles;
  if (tiles == 0) return;
  // round up to a multiple of 8 so every XCD gets whole slots
  const int grid = (tiles + 7) / 8 * 8;
  hipLaunchKernelGGL(gemm_x3_kernel<E>, dim3(grid), dim3(256), 0, st, bt,
                     (const E *)Ah, (const E *)Al, (const E *)Bh, (const E *)Bl, Z);
  FLARE_HIP_CHECK(hipGetLastError());
}

void gemm_bf16x3_batched(hipStream_t st, int nb, const GemmProduct *prods,
                         const void *Ah, const void *Al, const void *Bh,
                         const void *Bl, float *Z) {
  gemm_x3_batched<__bf16>(st, nb, prods, Ah, Al, Bh, Bl, Z);
}

void gemm_f16x3_batched(hipStream_t st, int nb, const GemmProduct *prods,
                        const void *Ah, const void *Al, const void *Bh,
                        const void *Bl, float *Z) {
  gemm_x3_batched<_Float16>(st, nb, prods, Ah, Al, Bh, Bl, Z);
}

// ---------------------------------------------------------------------------
// Wide-row BF16 x 3 GEMM for the mapped contraction Z = D beta (N = F of a
// few hundred to a few thousand columns, M = atoms of a species).
//
// Tile 256 x 352 x 32 per workgroup of eight wavefronts in 4 x 2 (two per
// SIMD), each 64 x 176 = 4 x 11 MFMA tiles (176 FP32 accumulators per lane,
// all in AGPRs), one workgroup per CU (its 152 KB of LDS).  At F = 680 two
// column tiles cover a row (704 columns, the last 24 discarded) and the
// 32,000-atom workload is 252 workgroups: a single round on 256 CUs, where
// 128 x 256 tiles at two per CU need 1.5 rounds.
//
// Operands move global -> LDS without registers (buffer_load_dwordx4 ... lds:
// lane l of an instruction fills bytes [16 l, 16 l + 16) of a 1 KB LDS run;
// a buffer resource per operand block with the K step as the scalar offset
// keeps one 32-bit vector offset per run instead of a 64-bit address), into
// two stages of 76 KB: the loads of step s + 1 are issued between the MFMAs
// of step s's first columns; one wait and one barrier follow the first MFMA
// pass of the step's last column, after which the next step's A-lo fragments
// and first B column are read while the last column's other two passes run
// (the A-hi fragments, still in use there, are read at the top of the step).  A stage
// holds Ah, Al (256 rows each), Bh, Bl (352 rows), 64-byte rows whose four
// 16-byte chunks are stored at chunk c ^ ((r >> 2) & 3), as swz() above.
// Rows past M or N load a clamped in-range row: they only reach output
// rows / columns that are not stored.
constexpr int kWM = 256, kWN = 352, kWK = 32;
constexpr int kWaves = 8, kWaveM = 64, kWaveN = 176;
constexpr int kFragM = kWaveM / 16, kFragN = kWaveN / 16;
constexpr int kWRowBytes = kWK * 2;                            // 64
constexpr int kWStage = (2 * kWM + 2 * kWN) * kWRowBytes;      // 77,824
constexpr int kWEpiStride = kWaveN + 4;  // FP32 staging row (banks offset)
static_assert((kWM * 4) % 64 == 0 && (kWN * 4) % 64 == 0,
              "each 64-chunk load run lies inside one operand block");

// Per-lane byte offsets of the operand chunks a wavefront loads: the hi
// and lo runs of a block share them (same rows and chunks, other base), and
// the K step enters through the wave-uniform base.  Wavefront w loads A
// runs w + 8 t (t < 2) of each of Ah, Al, and B runs w + 8 t (t < 3, run
// < 22) of each of Bh, Bl.
constexpr int kWRunsA = kWM * 4 / 64, kWRunsB = kWN * 4 / 64;  // 16, 22
constexpr int kWPerA = kWRunsA / kWaves, kWPerB = (kWRunsB + kWaves - 1) / kWaves;
static_assert(kWRunsA % kWaves == 0, "A runs split evenly over the wavefronts");

struct WideOffsets {
  uint32_t a[kWPerA], b[kWPerB];
};

__device__ inline WideOffsets wide_offsets(int M, int N, int Kp, int m0, int n0, int wv,
                                           int lane) {
  WideOffsets o;
#pragma unroll
  for (int t = 0; t < kWPerA; t++) {
    const int q = (wv + kWaves * t) * 64 + lane;  // chunk within the block
    const int row = q >> 2, kc = (q & 3) ^ ((row >> 2) & 3);
    o.a[t] = (uint32_t)(min(m0 + row, M - 1) * Kp + kc * 8) * 2u;
  }
#pragma unroll
  for (int t = 0; t < kWPerB; t++) {
    const int run = min(wv + kWaves * t, kWRunsB - 1);
    const int q = run * 64 + lane;
    const int row = q >> 2, kc = (q & 3) ^ ((row >> 2) & 3);
    o.b[t] = (uint32_t)(min(n0 + row, N - 1) * Kp + kc * 8) * 2u;
  }
  return o;
}

// 16 bytes per lane from a buffer (byte offset voff + soff) into the LDS
// run at dst (wave-uniform).
__device__ inline void lds_buf16(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff,
                                 unsigned char *dst) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(
      r, (__attribute__((address_space(3))) void *)dst, 16, voff, soff, 0, 0);
}

// Raw buffer resources of the four operand blocks of one product.
struct WideRsrc {
  __amdgpu_buffer_rsrc_t ah, al, bh, bl;
};

__device__ inline __amdgpu_buffer_rsrc_t wide_rsrc(const void *base, int64_t bytes) {
  // word 3 0x00020000: the raw-buffer format of the gfx9 family
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                           (int)(uint32_t)bytes, 0x00020000);
}

// Piece j of the next stage's loads, issued between the MFMAs of column j:
// the wavefront's (up to) ten loads go out two per column over columns 0-4
// instead of in one burst at the top of the step (measured 84 vs 90 us at
// 32,000 x 680 x 704: the burst held back the operand reads of the first
// columns behind twenty queued loads per SIMD).
constexpr int kWPiecesPerCol = 2;
__device__ inline void wide_stage_piece(int j, unsigned char *stage, const WideOffsets &o,
                                        const WideRsrc &r, int kb, int wv) {
  unsigned char *sAh = stage, *sAl = stage + kWM * kWRowBytes;
  unsigned char *sBh = stage + 2 * kWM * kWRowBytes, *sBl = sBh + kWN * kWRowBytes;
#pragma unroll
  for (int u = 0; u < kWPiecesPerCol; u++) {
    const int q = j * kWPiecesPerCol + u;
    if (q < 2 * kWPerA) {
      const int t = q >> 1, run = wv + kWaves * t;
      if (q & 1)
        lds_buf16(r.al, o.a[t], kb, sAl + run * 1024);
      else
        lds_buf16(r.ah, o.a[t], kb, sAh + run * 1024);
    } else if (q < 2 * kWPerA + 2 * kWPerB) {
      const int t = (q - 2 * kWPerA) >> 1, run = wv + kWaves * t;
      if (run < kWRunsB) {
        if (q & 1)
          lds_buf16(r.bl, o.b[t], kb, sBl + run * 1024);
        else
          lds_buf16(r.bh, o.b[t], kb, sBh + run * 1024);
      }
    }
  }
}
static_assert((2 * kWPerA + 2 * kWPerB + kWPiecesPerCol - 1) / kWPiecesPerCol <= kFragN,
              "every load piece falls in a column of the step");

template <typename E>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_x3w_kernel(GemmBatch bt, const E *__restrict__ Ah0, const E *__restrict__ Al0,
                const E *__restrict__ Bh0, const E *__restrict__ Bl0,
                float *__restrict__ Z0) {
  typedef typename Frag<E>::V V8;
  __shared__ __align__(16) unsigned char smem[2 * kWStage];
  const int lane = threadIdx.x & 63;
  // wave-uniform (scalar) so the per-wavefront load runs branch on SGPRs
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = (wv >> 1) * kWaveM, wn = (wv & 1) * kWaveN;
  int T, dummy;
  tile_of_block(bt.tile0[bt.nb], 1, T, dummy);
  if (T >= bt.tile0[bt.nb]) return;
  int q = 0;
  while (T >= bt.tile0[q + 1]) q++;
  const GemmProduct &gp = bt.g[q];
  // K columns of A and B rows of stride Kp
  const int M = gp.M, N = gp.N, K = gp.K, Kp = gp.ld_ab ? gp.ld_ab : gp.K, ldz = gp.ldz;
  const int nct = (N + kWN - 1) / kWN;
  static_assert(kWaves == 2 * (kWM / kWaveM) && 2 * kWaveN == kWN, "4 x 2 wavefronts");
  const int local = T - bt.tile0[q];
  const int rt = local / nct, ct = local - rt * nct;
  const E *Ah = Ah0 + gp.a_off, *Al = Al0 + gp.a_off;
  const E *Bh = Bh0 + gp.b_off, *Bl = Bl0 + gp.b_off;
  float *Z = Z0 + gp.z_off;
  const int m0 = rt * kWM, n0 = ct * kWN;
  const int steps = K / kWK;

  floatx4 acc[kFragM][kFragN];
#pragma unroll
  for (int i = 0; i < kFragM; i++)
#pragma unroll
    for (int j = 0; j < kFragN; j++) acc[i][j] = floatx4{0, 0, 0, 0};

  // operand rows of this lane: wm + 16 i + fr (A), wn + 16 j + fr (B); the
  // lane's chunk is lane >> 4 (elements 8 (lane >> 4) .. + 7 of the K step)
  const int fr = lane & 15, gch = lane >> 4;
  const int offA = (wm + fr) * kWRowBytes + ((gch ^ (((wm + fr) >> 2) & 3)) << 4);
  const int offB = (wn + fr) * kWRowBytes + ((gch ^ (((wn + fr) >> 2) & 3)) << 4);
  // (16 i rows later the swizzle ((r >> 2) & 3) is unchanged: 16 i >> 2 is a
  // multiple of 4)

  const WideOffsets lo = wide_offsets(M, N, Kp, m0, n0, wv, lane);
  WideRsrc rs;
  const int64_t bytes_a = ((int64_t)(M - 1) * Kp + K) * 2, bytes_b = ((int64_t)(N - 1) * Kp + K) * 2;
  rs.ah = wide_rsrc(Ah, bytes_a);
  rs.al = wide_rsrc(Al, bytes_a);
  rs.bh = wide_rsrc(Bh, bytes_b);
  rs.bl = wide_rsrc(Bl, bytes_b);
#pragma unroll
  for (int j = 0; j < kFragN; j++) wide_stage_piece(j, smem, lo, rs, 0, wv);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  V8 ah[kFragM], al[kFragM], bh, bl;
  {
    const unsigned char *sAl = smem + kWM * kWRowBytes, *sBh = smem + 2 * kWM * kWRowBytes;
    const unsigned char *sBl = sBh + kWN * kWRowBytes;
#pragma unroll
    for (int i = 0; i < kFragM; i++)
      al[i] = *reinterpret_cast<const V8 *>(sAl + offA + i * 16 * kWRowBytes);
    bh = *reinterpret_cast<const V8 *>(sBh + offB);
    bl = *reinterpret_cast<const V8 *>(sBl + offB);
  }
  // step s; `more`: a step s + 1 follows (a compile-time flag, so that the
  // loop body has no branches on it)
  auto step = [&](int s, auto more_c) {
    constexpr bool more = decltype(more_c)::value;
    const unsigned char *cur = smem + (s & 1) * kWStage;
    unsigned char *nst = smem + ((s + 1) & 1) * kWStage;
    const int kb = (s + 1) * kWK * 2;  // bytes
    const unsigned char *sBh = cur + 2 * kWM * kWRowBytes;
    const unsigned char *sBl = sBh + kWN * kWRowBytes;
#pragma unroll
    for (int i = 0; i < kFragM; i++)
      ah[i] = *reinterpret_cast<const V8 *>(cur + offA + i * 16 * kWRowBytes);
#pragma unroll
    for (int j = 0; j < kFragN; j++) {
      // column j's B fragments are read while column j - 1's MFMAs run
      V8 nbh = bh, nbl = bl;
      if (more) wide_stage_piece(j, nst, lo, rs, kb, wv);
      if (j + 1 < kFragN) {
        nbh = *reinterpret_cast<const V8 *>(sBh + offB + (j + 1) * 16 * kWRowBytes);
        nbl = *reinterpret_cast<const V8 *>(sBl + offB + (j + 1) * 16 * kWRowBytes);
      }
      // three passes over the accumulators of column j, in this order
#pragma unroll
      for (int i = 0; i < kFragM; i++) acc[i][j] = mma32(al[i], bh, acc[i][j]);
      if (j == kFragN - 1) {
        // this wavefront's reads of the stage are done (its last B column is
        // in registers) and its loads of the next stage have landed
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (more) {
          const unsigned char *nAl = nst + kWM * kWRowBytes;
          const unsigned char *nBh = nst + 2 * kWM * kWRowBytes, *nBl = nBh + kWN * kWRowBytes;
#pragma unroll
          for (int i = 0; i < kFragM; i++)
            al[i] = *reinterpret_cast<const V8 *>(nAl + offA + i * 16 * kWRowBytes);
          nbh = *reinterpret_cast<const V8 *>(nBh + offB);
          nbl = *reinterpret_cast<const V8 *>(nBl + offB);
        }
      }
#pragma unroll
      for (int i = 0; i < kFragM; i++) acc[i][j] = mma32(ah[i], bl, acc[i][j]);
#pragma unroll
      for (int i = 0; i < kFragM; i++) acc[i][j] = mma32(ah[i], bh, acc[i][j]);
      bh = nbh;
      bl = nbl;
    }
  };
  for (int s = 0; s + 1 < steps; s++) step(s, std::true_type{});
  step(steps - 1, std::false_type{});

  // Z through LDS, 16 rows of the wavefront tile at a time, 16-byte stores
  float *stage = reinterpret_cast<float *>(smem) + wv * 16 * kWEpiStride;
  const int dcol = lane & 15, drow = (lane >> 4) * 4;
  const bool vec4 = (N & 3) == 0 && (ldz & 3) == 0 && (gp.z_off & 3) == 0;
  const bool scaled = gp.rs || gp.cs;
#pragma unroll
  for (int i = 0; i < kFragM; i++) {
#pragma unroll
    for (int j = 0; j < kFragN; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) stage[(drow + e) * kWEpiStride + 16 * j + dcol] = acc[i][j][e];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int kVec = kWaveN / 4;  // float4 per staged row
#pragma unroll
    for (int t = 0; t < 16 * kVec / 64; t++) {
      const int f = lane + 64 * t;
      const int r = f / kVec, c = (f - r * kVec) * 4;
      const int row = m0 + wm + 16 * i + r, col = n0 + wn + c;
      float4 v = *reinterpret_cast<const float4 *>(&stage[r * kWEpiStride + c]);
      if (row < M) {
        if (scaled) {
          // (acc x row scale) x column scale, as gemm_x3_kernel
          const float rsc = gp.rs ? gp.rs[row] : 1.0f;
          float cv[4] = {1.0f, 1.0f, 1.0f, 1.0f};
          if (gp.cs)
#pragma unroll
            for (int e = 0; e < 4; e++)
              if (col + e < N) cv[e] = gp.cs[col + e];
          v.x = v.x * rsc * cv[0];
          v.y = v.y * rsc * cv[1];
          v.z = v.z * rsc * cv[2];
          v.w = v.w * rsc * cv[3];
        }
        float *dst = Z + (int64_t)row * ldz + col;
        if (vec4) {
          if (col < N) *reinterpret_cast<float4 *>(dst) = v;
        } else {
          if (col < N) dst[0] = v.x;
          if (col + 1 < N) dst[1] = v.y;
          if (col + 2 < N) dst[2] = v.z;
          if (col + 3 < N) dst[3] = v.w;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}
static_assert(16 * (kWaveN / 4) % 64 == 0, "epilogue rows split evenly over lanes");
static_assert(kWaves * 16 * kWEpiStride * 4 <= 2 * kWStage, "epilogue staging fits the LDS");

int gemm_x3_wide_tiles(int nb, const GemmProduct *prods) {
  int tiles = 0;
  for (int q = 0; q < nb; q++)
    if (prods[q].M > 0 && prods[q].N > 0)
      tiles += (prods[q].M + kWM - 1) / kWM * ((prods[q].N + kWN - 1) / kWN);
  return tiles;
}

template <typename E>
static void gemm_x3_wide(hipStream_t st, int nb, const GemmProduct *prods, const void *Ah,
                         const void *Al, const void *Bh, const void *Bl, float *Z) {
  FLARE_REQUIRE(nb >= 0 && nb <= kGemmMaxBatch, "gemm_x3w: at most 8 products");
  GemmBatch bt;
  bt.nb = 0;
  int tiles = 0;
  for (int q = 0; q < nb; q++) {
    const GemmProduct &g = prods[q];
    FLARE_REQUIRE(g.K >= kWK && g.K % kWK == 0, "gemm_x3w: K must be padded to 32");
    const int64_t ld = g.ld_ab ? g.ld_ab : g.K;
    FLARE_REQUIRE(ld >= g.K && ld % 8 == 0 && g.a_off % 8 == 0 && g.b_off % 8 == 0,
                  "gemm_x3w: row stride below K, or operand rows not 16-byte aligned");
    // 32-bit buffer offsets and record counts
    FLARE_REQUIRE((int64_t)g.M * ld * 2 <= (int64_t)UINT32_MAX &&
                      (int64_t)g.N * ld * 2 <= (int64_t)UINT32_MAX,
                  "gemm_x3w: an operand block of one product exceeds 4 GiB");
    if (g.M <= 0 || g.N <= 0) continue;
    bt.g[bt.nb] = g;
    bt.tile0[bt.nb] = tiles;
    tiles += (g.M + kWM - 1) / kWM * ((g.N + kWN - 1) / kWN);
    bt.nb++;
  }
  bt.tile0[bt.nb] = tiles;
  if (tiles == 0) return;
  const int grid = (tiles + 7) / 8 * 8;
  hipLaunchKernelGGL(gemm_x3w_kernel<E>, dim3(grid), dim3(64 * kWaves), 0, st, bt,
                     (const E *)Ah, (const E *)Al, (const E *)Bh, (const E *)Bl, Z);
  FLARE_HIP_CHECK(hipGetLastError());
}

void gemm_bf16x3_wide(hipStream_t st, int nb, const GemmProduct *prods, const void *Ah,
                      const void *Al, const void *Bh, const void *Bl, float *Z) {
  gemm_x3_wide<__bf16>(st, nb, prods, Ah, Al, Bh, Bl, Z);
}

void gemm_f16x3_wide(hipStream_t st, int nb, const GemmProduct *prods, const void *Ah,
                     const void *Al, const void *Bh, const void *Bl, float *Z) {
  gemm_x3_wide<_Float16>(st, nb, prods, Ah, Al, Bh, Bl, Z);
}

// ---------------------------------------------------------------------------
// FP32 wide-row GEMM Z = A B^T (the FP32 form's mapped contraction Z = X beta,
// beta symmetric) on v_mfma_f32_16x16x4_f32: the wide kernel's tile (256 x
// 352, eight wavefronts of 64 x 176, 176 accumulators each) with K steps of
// 16 values (64-byte LDS rows, the same swizzle), two stages of 38 KB filled
// by buffer loads (A 16 runs, B 22, up to five loads per wavefront per step,
// issued one per column over the first five columns of the step before).
// A lane group g holds chunk g (values 4 g .. 4 g + 3) of a K step, and MFMA t
// of a block takes value 4 g + t of both operands, so each fragment is one
// 16-byte read.  Rows are K values (K a multiple of 4, row starts 16-byte
// aligned); the last step's chunks past K, which read the next row or past
// the buffer (zeros), are zeroed in both operands' fragments.
constexpr int kFK = 16;                               // K step (values)
constexpr int kFStage = (kWM + kWN) * kFK * 4;        // 38,912
constexpr int kFPerA = kWRunsA / kWaves, kFPerB = kWPerB;  // 2, 3 runs per wavefront
constexpr int kFLoads = kFPerA + kFPerB;
static_assert(kFLoads <= kFragN, "every load falls in a column of the step");

struct F32Offsets {
  uint32_t a[kFPerA], b[kFPerB];
};

__device__ inline F32Offsets f32w_offsets(int M, int N, int K, int m0, int n0, int wv,
                                          int lane) {
  F32Offsets o;
#pragma unroll
  for (int t = 0; t < kFPerA; t++) {
    const int q = (wv + kWaves * t) * 64 + lane;
    const int row = q >> 2, kc = (q & 3) ^ ((row >> 2) & 3);
    o.a[t] = (uint32_t)(min(m0 + row, M - 1) * K + kc * 4) * 4u;
  }
#pragma unroll
  for (int t = 0; t < kFPerB; t++) {
    const int run = min(wv + kWaves * t, kWRunsB - 1);
    const int q = run * 64 + lane;
    const int row = q >> 2, kc = (q & 3) ^ ((row >> 2) & 3);
    o.b[t] = (uint32_t)(min(n0 + row, N - 1) * K + kc * 4) * 4u;
  }
  return o;
}

// load q (0 .. kFLoads - 1) of this wavefront's share of a stage
__device__ inline void f32w_piece(int q, unsigned char *stage, const F32Offsets &o,
                                  __amdgpu_buffer_rsrc_t ra, __amdgpu_buffer_rsrc_t rb,
                                  int kb, int wv) {
  unsigned char *sA = stage, *sB = stage + kWM * kFK * 4;
  if (q < kFPerA) {
    const int run = wv + kWaves * q;
    lds_buf16(ra, o.a[q], kb, sA + run * 1024);
  } else if (q < kFLoads) {
    const int t = q - kFPerA, run = wv + kWaves * t;
    if (run < kWRunsB) lds_buf16(rb, o.b[t], kb, sB + run * 1024);
  }
}

__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
gemm_f32w_kernel(GemmBatch bt, const float *__restrict__ A0, const float *__restrict__ B0,
                 float *__restrict__ Z0) {
  constexpr int kEpiBytes = kWaves * 16 * kWEpiStride * 4;
  __shared__ __align__(16) unsigned char smem[2 * kFStage > kEpiBytes ? 2 * kFStage : kEpiBytes];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = (wv >> 1) * kWaveM, wn = (wv & 1) * kWaveN;
  int T, dummy;
  tile_of_block(bt.tile0[bt.nb], 1, T, dummy);
  if (T >= bt.tile0[bt.nb]) return;
  int q = 0;
  while (T >= bt.tile0[q + 1]) q++;
  const GemmProduct &gp = bt.g[q];
  const int M = gp.M, N = gp.N, K = gp.K, ldz = gp.ldz;
  const int nct = (N + kWN - 1) / kWN;
  const int local = T - bt.tile0[q];
  const int rt = local / nct, ct = local - rt * nct;
  const float *A = A0 + gp.a_off, *B = B0 + gp.b_off;
  float *Z = Z0 + gp.z_off;
  const int m0 = rt * kWM, n0 = ct * kWN;
  const int steps = (K + kFK - 1) / kFK;
  const int tail = (K % kFK) / 4;  // chunks of the last step inside K (0: all four)
  floatx4 acc[kFragM][kFragN];
#pragma unroll
  for (int i = 0; i < kFragM; i++)
#pragma unroll
    for (int j = 0; j < kFragN; j++) acc[i][j] = floatx4{0, 0, 0, 0};
  const int fr = lane & 15, g = lane >> 4;
  constexpr int kRow = kFK * 4;  // LDS row bytes (64)
  const int offA = (wm + fr) * kRow + ((g ^ (((wm + fr) >> 2) & 3)) << 4);
  const int offB = (wn + fr) * kRow + ((g ^ (((wn + fr) >> 2) & 3)) << 4);
  const F32Offsets lo = f32w_offsets(M, N, K, m0, n0, wv, lane);
  const __amdgpu_buffer_rsrc_t ra = wide_rsrc(A, (int64_t)M * K * 4);
  const __amdgpu_buffer_rsrc_t rb = wide_rsrc(B, (int64_t)N * K * 4);
#pragma unroll
  for (int p = 0; p < kFLoads; p++) f32w_piece(p, smem, lo, ra, rb, 0, wv);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int s = 0; s < steps; s++) {
    const unsigned char *cur = smem + (s & 1) * kFStage;
    unsigned char *nst = smem + ((s + 1) & 1) * kFStage;
    const bool more = s + 1 < steps;
    const int kb = (s + 1) * kRow;
    const unsigned char *sA = cur, *sB = cur + kWM * kRow;
    const bool zero = tail && s == steps - 1 && g >= tail;
    floatx4 a[kFragM];
#pragma unroll
    for (int i = 0; i < kFragM; i++) {
      a[i] = *reinterpret_cast<const floatx4 *>(sA + offA + i * 16 * kRow);
      if (zero) a[i] = floatx4{0, 0, 0, 0};
    }
    floatx4 b = *reinterpret_cast<const floatx4 *>(sB + offB);
    if (zero) b = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < kFragN; j++) {
      floatx4 nb = b;
      if (more && j < kFLoads) f32w_piece(j, nst, lo, ra, rb, kb, wv);
      if (j + 1 < kFragN) {
        nb = *reinterpret_cast<const floatx4 *>(sB + offB + (j + 1) * 16 * kRow);
        if (zero) nb = floatx4{0, 0, 0, 0};
      }
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < kFragM; i++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][t], b[t], acc[i][j], 0, 0, 0);
      b = nb;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // Z through LDS, 16 rows of the wavefront tile at a time
  float *stage = reinterpret_cast<float *>(smem) + wv * 16 * kWEpiStride;
  const int dcol = lane & 15, drow = (lane >> 4) * 4;
  const bool vec4 = (N & 3) == 0 && (ldz & 3) == 0 && (gp.z_off & 3) == 0;
#pragma unroll
  for (int i = 0; i < kFragM; i++) {
#pragma unroll
    for (int j = 0; j < kFragN; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) stage[(drow + e) * kWEpiStride + 16 * j + dcol] = acc[i][j][e];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int kVec = kWaveN / 4;
#pragma unroll
    for (int t = 0; t < 16 * kVec / 64; t++) {
      const int f = lane + 64 * t;
      const int r = f / kVec, c = (f - r * kVec) * 4;
      const int row = m0 + wm + 16 * i + r, col = n0 + wn + c;
      const float4 v = *reinterpret_cast<const float4 *>(&stage[r * kWEpiStride + c]);
      if (row < M) {
        float *dst = Z + (int64_t)row * ldz + col;
        if (vec4) {
          if (col < N) *reinterpret_cast<float4 *>(dst) = v;
        } else {
          if (col < N) dst[0] = v.x;
          if (col + 1 < N) dst[1] = v.y;
          if (col + 2 < N) dst[2] = v.z;
          if (col + 3 < N) dst[3] = v.w;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

void sgemm_wide(hipStream_t st, int nb, const GemmProduct *prods, const float *A,
                const float *B, float *Z) {
  FLARE_REQUIRE(nb >= 0 && nb <= kGemmMaxBatch, "sgemm_wide: at most 8 products");
  GemmBatch bt;
  bt.nb = 0;
  int tiles = 0;
  for (int q = 0; q < nb; q++) {
    const GemmProduct &g = prods[q];
    FLARE_REQUIRE(g.K >= 4 && g.K % 4 == 0 && g.a_off % 4 == 0 && g.b_off % 4 == 0,
                  "sgemm_wide: rows of a multiple of 4 values, 16-byte aligned");
    FLARE_REQUIRE(!g.ld_ab && !g.rs && !g.cs, "sgemm_wide: plain products only");
    FLARE_REQUIRE((int64_t)g.M * g.K * 4 <= (int64_t)UINT32_MAX &&
                      (int64_t)g.N * g.K * 4 <= (int64_t)UINT32_MAX,
                  "sgemm_wide: an operand block of one product exceeds 4 GiB");
    if (g.M <= 0 || g.N <= 0) continue;
    bt.g[bt.nb] = g;
    bt.tile0[bt.nb] = tiles;
    tiles += (g.M + kWM - 1) / kWM * ((g.N + kWN - 1) / kWN);
    bt.nb++;
  }
  bt.tile0[bt.nb] = tiles;
  if (tiles == 0) return;
  const int grid = (tiles + 7) / 8 * 8;
  hipLaunchKernelGGL(gemm_f32w_kernel, dim3(grid), dim3(64 * kWaves), 0, st, bt, A, B, Z);
  FLARE_HIP_CHECK(hipGetLastError());
}

// FP16 hi/lo split of rows with a power-of-two scale per row: the row's
// largest |x| is brought into [0.5, 1); inv_scale[r] = 2^-e.  One wavefront
// per row; X is FP64 or FP32 (row stride ldx), out rows of kp (zero padded).
template <typename T>
__global__ void __launch_bounds__(64)
split_f16_rows_kernel(int rows, int cols, int ldx, int kp, const T *__restrict__ x,
                      _Float16 *__restrict__ hi, _Float16 *__restrict__ lo,
                      float *__restrict__ inv_scale) {
  const int r = blockIdx.x, lane = threadIdx.x;
  const T *xr = x + (int64_t)r * ldx;
  float m = 0.0f;
  for (int c = lane; c < cols; c += 64) m = fmaxf(m, fabsf((float)xr[c]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  int e;
  const bool nonzero = f16_row_scale(m, e);
  const float sc = ldexpf(1.0f, e);
  for (int c = lane; c < kp; c += 64) {
    const float v = nonzero && c < cols ? (float)xr[c] * sc : 0.0f;
    const _Float16 h = (_Float16)v;
    hi[(int64_t)r * kp + c] = h;
    lo[(int64_t)r * kp + c] = (_Float16)(v - (float)h);
  }
  if (lane == 0) inv_scale[r] = ldexpf(1.0f, -e);
}

// The same for the transpose of a (rows x cols) FP64 block: output row c
// (c < cols) holds column c, ldo values (zero for r >= rows), scaled by
// its own largest element.  One wavefront per output row.
__global__ void __launch_bounds__(64)
split_f16_cols_kernel(int rows, int cols, int ldo, const double *__restrict__ x,
                      _Float16 *__restrict__ hi, _Float16 *__restrict__ lo,
                      float *__restrict__ inv_scale) {
  const int c = blockIdx.x, lane = threadIdx.x;
  float m = 0.0f;
  for (int r = lane; r < rows; r += 64) m = fmaxf(m, fabsf((float)x[(int64_t)r * cols + c]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  int e;
  const bool nonzero = f16_row_scale(m, e);
  const float sc = ldexpf(1.0f, e);
  for (int r = lane; r < ldo; r += 64) {
    const float v = nonzero && r < rows ? (float)x[(int64_t)r * cols + c] * sc : 0.0f;
    const _Float16 h = (_Float16)v;
    hi[(int64_t)c * ldo + r] = h;
    lo[(int64_t)c * ldo + r] = (_Float16)(v - (float)h);
  }
  if (lane == 0) inv_scale[c] = ldexpf(1.0f, -e);
}

void split_f16_rows(hipStream_t st, int rows, int cols, int ldx, int kp,
                    const double *x, void *hi, void *lo, float *inv_scale) {
  if (rows == 0) return;
  hipLaunchKernelGGL(split_f16_rows_kernel<double>, dim3(rows), dim3(64), 0, st,
                     rows, cols, ldx, kp, x, (_Float16 *)hi, (_Float16 *)lo,
                     inv_scale);
  FLARE_HIP_CHECK(hipGetLastError());
}

void split_f16_cols(hipStream_t st, int rows, int cols, int ldo, const double *x,
                    void *hi, void *lo, float *inv_scale) {
  if (cols == 0) return;
  hipLaunchKernelGGL(split_f16_cols_kernel, dim3(cols), dim3(64), 0, st, rows,
                     cols, ldo, x, (_Float16 *)hi, (_Float16 *)lo, inv_scale);
  FLARE_HIP_CHECK(hipGetLastError());
}

void split_bf16(hipStream_t st, int rows, int cols, int kp, const double *x,
                void *hi, void *lo) {
  if (rows == 0) return;
  int64_t n = (int64_t)rows * kp;
  hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, st, rows, cols, kp, x, (__bf16 *)hi,
                     (__bf16 *)lo);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
