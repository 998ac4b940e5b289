// mfma.hip -- BF16x3 matrix-core GEMM for the mapped contraction Z = D beta.
//
// Tile 128 x 128 per workgroup of 256 threads (four wavefronts in 2 x 2,
// each 64 x 64 = 4 x 4 MFMA tiles of 16 x 16), K step 32 with
// v_mfma_f32_16x16x32_bf16.  Operands arrive in FP32 and are split into
// hi = bf16(x), lo = bf16(x - hi) while staged to LDS; per K step each
// output tile takes hi*hi + hi*lo + lo*hi into one FP32 accumulator.
// beta is symmetric, so row n of beta is column n: both operands are staged
// with K contiguous.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (64 lanes):
//   A[i][k]: lane = i + 16 (k / 8), element k % 8
//   B[k][j]: lane = j + 16 (k / 8), element k % 8
//   D[i][j]: lane = j + 16 (i / 4), element i % 4
#include "common.h"
#include "mfma.h"

namespace flare {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float floatx4;

constexpr int kTileM = 128, kTileN = 128, kTileK = 32, kPadK = 40;

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }

// Stage a 128 x 32 slab of rows [r0, r0+128) x [k0, k0+32) of a row-major
// (rows x K) FP32 matrix into hi/lo LDS tiles, zero outside the matrix.
__device__ inline void stage(const float *__restrict__ src, int rows, int K,
                             int ld, int r0, int k0, __bf16 (*hi)[kPadK],
                             __bf16 (*lo)[kPadK]) {
  const int t = threadIdx.x;
  const int r = t >> 1;          // 0..127
  const int ks = (t & 1) * 16;   // 0 or 16
  const int gr = r0 + r;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    int gk = k0 + ks + q;
    float x = (gr < rows && gk < K) ? src[(int64_t)gr * ld + gk] : 0.0f;
    __bf16 h = to_bf16(x);
    __bf16 l = to_bf16(x - (float)h);
    hi[r][ks + q] = h;
    lo[r][ks + q] = l;
  }
}

__global__ void __launch_bounds__(256)
gemm_bf16x3_kernel(int M, int N, int K, const float *__restrict__ X,
                   const float *__restrict__ B, float *__restrict__ Z) {
  __shared__ __bf16 Ah[kTileM][kPadK], Al[kTileM][kPadK];
  __shared__ __bf16 Bh[kTileN][kPadK], Bl[kTileN][kPadK];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;
  const int m0 = blockIdx.y * kTileM, n0 = blockIdx.x * kTileN;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0, 0, 0, 0};
  const int fr = lane & 15;        // row (A) / column (B) within a 16 tile
  const int fk = (lane >> 4) * 8;  // k offset of this lane's 8 elements
  for (int k0 = 0; k0 < K; k0 += kTileK) {
    __syncthreads();
    stage(X, M, K, K, m0, k0, Ah, Al);
    stage(B, N, K, K, n0, k0, Bh, Bl);  // symmetric: row n of B == column n
    __syncthreads();
    bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      ah[i] = *reinterpret_cast<const bf16x8 *>(&Ah[wm + 16 * i + fr][fk]);
      al[i] = *reinterpret_cast<const bf16x8 *>(&Al[wm + 16 * i + fr][fk]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      bh[j] = *reinterpret_cast<const bf16x8 *>(&Bh[wn + 16 * j + fr][fk]);
      bl[j] = *reinterpret_cast<const bf16x8 *>(&Bl[wn + 16 * j + fr][fk]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
      }
  }
  const int dcol = lane & 15;
  const int drow = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int col = n0 + wn + 16 * j + dcol;
      if (col >= N) continue;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int row = m0 + wm + 16 * i + drow + q;
        if (row < M) Z[(int64_t)row * N + col] = acc[i][j][q];
      }
    }
}

size_t gemm_bf16x3_workspace(int, int, int) { return 0; }

void gemm_bf16x3(hipStream_t st, int M, int N, int K, const float *X,
                 const float *Bsym, float *Z, void *) {
  if (M == 0 || N == 0) return;
  FLARE_REQUIRE(N == K, "gemm_bf16x3 expects a square symmetric right operand");
  dim3 grid((N + kTileN - 1) / kTileN, (M + kTileM - 1) / kTileM);
  hipLaunchKernelGGL(gemm_bf16x3_kernel, grid, dim3(256), 0, st, M, N, K, X,
                     Bsym, Z);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
