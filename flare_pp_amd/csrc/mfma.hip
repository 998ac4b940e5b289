// mfma.hip -- BF16x3 matrix-core GEMM for the mapped contraction Z = D beta.
//
// Operands arrive pre-split into BF16 hi + lo parts (D by the forward
// kernel, beta once per potential), K-padded with zeros to a multiple of 32:
//   Z (M x N, FP32) = (Ah + Al)(M x Kp) . (Bh + Bl)(N x Kp)^T,
// accumulating hi*hi + hi*lo + lo*hi (lo*lo, ~2^-16 relative, is dropped).
// beta is symmetric, so its row n is column n and both operands are read
// with K contiguous.
//
// Tile 128 x 256 x 32 per workgroup of 256 threads: four wavefronts in
// 2 x 2, each 64 x 128 = 4 x 8 tiles of v_mfma_f32_16x16x32_bf16.  Global
// tiles move in 16-byte vectors and are prefetched into registers during
// the MFMAs of the current step; workgroup ids are remapped so that the
// column tiles of a row block share one XCD's L2.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (64 lanes):
//   A[i][k]: lane = i + 16 (k / 8), element k % 8
//   B[k][j]: lane = j + 16 (k / 8), element k % 8
//   D[i][j]: lane = j + 16 (i / 4), element i % 4
#include "common.h"
#include "mfma.h"

namespace flare {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;

// Up to kGemmMaxBatch independent products (the GemmProduct list) and the
// exclusive prefix of their tile counts.
struct GemmBatch {
  int nb;
  GemmProduct g[kGemmMaxBatch];
  int tile0[kGemmMaxBatch + 1];
};
typedef __attribute__((ext_vector_type(4))) float floatx4;

constexpr int kBM = 128, kBN = 256, kBK = 32, kLdk = kBK + 8;
constexpr int kChunksA = kBM * kBK / 8 / 256;  // 16-byte chunks per thread
constexpr int kChunksB = kBN * kBK / 8 / 256;

struct Tile {
  uint4 a[2 * kChunksA];  // Ah, Al
  uint4 b[2 * kChunksB];  // Bh, Bl
};

__device__ inline void load_tile(Tile &t, const __bf16 *__restrict__ Ah,
                                 const __bf16 *__restrict__ Al,
                                 const __bf16 *__restrict__ Bh,
                                 const __bf16 *__restrict__ Bl, int M, int N,
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

__device__ inline void store_tile(const Tile &t, __bf16 (*sAh)[kLdk],
                                  __bf16 (*sAl)[kLdk], __bf16 (*sBh)[kLdk],
                                  __bf16 (*sBl)[kLdk]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int h = 0; h < kChunksA; h++) {
    const int c = tid + 256 * h;
    const int r = c >> 2, kc = (c & 3) * 8;
    *reinterpret_cast<uint4 *>(&sAh[r][kc]) = t.a[h];
    *reinterpret_cast<uint4 *>(&sAl[r][kc]) = t.a[kChunksA + h];
  }
#pragma unroll
  for (int h = 0; h < kChunksB; h++) {
    const int c = tid + 256 * h;
    const int r = c >> 2, kc = (c & 3) * 8;
    *reinterpret_cast<uint4 *>(&sBh[r][kc]) = t.b[h];
    *reinterpret_cast<uint4 *>(&sBl[r][kc]) = t.b[kChunksB + h];
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

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2)))
gemm_bf16x3_kernel(GemmBatch bt, const __bf16 *__restrict__ Ah0,
                   const __bf16 *__restrict__ Al0, const __bf16 *__restrict__ Bh0,
                   const __bf16 *__restrict__ Bl0, float *__restrict__ Z0) {
  __shared__ __align__(16) __bf16 sAh[kBM][kLdk], sAl[kBM][kLdk];
  __shared__ __align__(16) __bf16 sBh[kBN][kLdk], sBl[kBN][kLdk];
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
  const __bf16 *Ah = Ah0 + gp.a_off, *Al = Al0 + gp.a_off;
  const __bf16 *Bh = Bh0 + gp.b_off, *Bl = Bl0 + gp.b_off;
  float *Z = Z0 + gp.z_off;
  const int m0 = rt * kBM, n0 = ct * kBN;
  floatx4 acc[4][8];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) acc[i][j] = floatx4{0, 0, 0, 0};
  const int fr = lane & 15;
  const int fk = (lane >> 4) * 8;
  const int steps = Kp / kBK;
  Tile t;
  load_tile(t, Ah, Al, Bh, Bl, M, N, Kp, m0, n0, 0);
  for (int s = 0; s < steps; s++) {
    __syncthreads();
    store_tile(t, sAh, sAl, sBh, sBl);
    __syncthreads();
    if (s + 1 < steps) load_tile(t, Ah, Al, Bh, Bl, M, N, Kp, m0, n0, (s + 1) * kBK);
    bf16x8 ah[4], al[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      ah[i] = *reinterpret_cast<const bf16x8 *>(&sAh[wm + 16 * i + fr][fk]);
      al[i] = *reinterpret_cast<const bf16x8 *>(&sAl[wm + 16 * i + fr][fk]);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&sBh[wn + 16 * j + fr][fk]);
      const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&sBl[wn + 16 * j + fr][fk]);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
      }
    }
  }
  const int dcol = lane & 15;
  const int drow = (lane >> 4) * 4;
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

void gemm_bf16x3_batched(hipStream_t st, int nb, const GemmProduct *prods,
                         const void *Ah, const void *Al, const void *Bh,
                         const void *Bl, float *Z) {
  FLARE_REQUIRE(nb >= 0 && nb <= kGemmMaxBatch, "gemm_bf16x3: at most 8 products");
  GemmBatch bt;
  bt.nb = 0;
  int tiles = 0;
  for (int q = 0; q < nb; q++) {
    const GemmProduct &g = prods[q];
    FLARE_REQUIRE(g.K % kBK == 0, "gemm_bf16x3: K must be padded to 32");
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
  hipLaunchKernelGGL(gemm_bf16x3_kernel, dim3(grid), dim3(256), 0, st, bt,
                     (const __bf16 *)Ah, (const __bf16 *)Al, (const __bf16 *)Bh,
                     (const __bf16 *)Bl, Z);
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
