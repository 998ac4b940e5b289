// common.h -- HIP-side helpers shared by the kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "flare_math.h"

namespace flare {

// Error reporting: every C-ABI entry returns a status and leaves a message
// retrievable with flare_last_error().
void set_error(const std::string &msg);

struct Error {
  int code;
  std::string msg;
};

#define FLARE_HIP_CHECK(expr)                                               \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess)                                                   \
      throw ::flare::Error{3, std::string(#expr) + ": " +                   \
                                  hipGetErrorString(_e)};                   \
  } while (0)

#define FLARE_REQUIRE(cond, msg)                                            \
  do {                                                                      \
    if (!(cond)) throw ::flare::Error{1, (msg)};                            \
  } while (0)

constexpr int kWave = 64;

// ---- wavefront reductions (64 lanes) ------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// ---- device parameter block of the B2 descriptor -------------------------
// Passed by value as a kernel argument.
struct B2Dev {
  int S, N, L, H, NH, W, F, nr;
  int cutoff_kind;
  int pair_pred;  // 0: r <= rc(si,sj) (b2.cpp:277); 1: r^2 < rc^2;
                  // 2: every listed pair (a caller's list, filtered on entry)
                  // (pair_flare.cpp:115)
  double r1;
  double cut[kMaxSpecies * kMaxSpecies];
};

// Neighbor list views (device pointers).  CSR over the Structure list of
// structure.cpp:52-118: pairs of atom i are [cum[i], cum[i+1]).
struct NListDev {
  int n;
  int64_t P;
  const int *cum;      // n+1
  const int *idx;      // P  structure index of the neighbor
  const int *nspec;    // P  species of the neighbor
  const double *rel;   // P*4 (r, dx, dy, dz)
  const int *rev;      // P  index of the reverse pair
  const int *center;   // P  central atom of each pair
};

__device__ inline bool pair_inside(const B2Dev &p, int si, int sj,
                                   const double *rv) {
#pragma clang fp contract(off)
  if (p.pair_pred == 2) return true;
  double rc = p.cut[si * p.S + sj];
  if (p.pair_pred == 0) return rv[0] <= rc;
  double rsq = rv[1] * rv[1] + rv[2] * rv[2] + rv[3] * rv[3];
  return rsq < rc * rc;
}

}  // namespace flare
