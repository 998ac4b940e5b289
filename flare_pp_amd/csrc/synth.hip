// synth.hip -- host-side generator of the disordered benchmark structure:
// random sequential addition (RSA) of hard points in a periodic cube.
//
// Candidates are uniform in the box (xoshiro256** stream seeded through
// splitmix64) and accepted when at least d_min from every atom accepted
// before; a cell grid of edge >= d_min bounds each check to 27 cells.
// Pure integer random bits and IEEE double arithmetic: the same seed gives
// the same positions on every machine.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.h"
#include "flare_hip.h"

namespace {

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) {
    for (int k = 0; k < 4; k++) {  // splitmix64
      seed += 0x9e3779b97f4a7c15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      s[k] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

}  // namespace

extern "C" int flare_synth_rsa(int n, double box, double d_min, uint64_t seed,
                               int64_t max_attempts, double *positions) {
  if (n < 0 || !(box > 0) || !(d_min > 0) || (n > 0 && !positions)) return FLARE_EINVAL;
  const int nc = std::max(1, (int)std::floor(box / d_min));
  const double edge = box / nc;
  if (nc < 3) return FLARE_EINVAL;  // neighbor cells must be distinct
  const double d2 = d_min * d_min;
  // cells of edge in [d_min, 2 d_min): at most 14 points fit in one with the
  // minimum distance, so a fixed 16 slots per cell (flat, no per-cell heap
  // blocks) holds every cell
  constexpr int kSlots = 16;
  const size_t n_cells = (size_t)nc * nc * nc;
  std::vector<int> slot((size_t)n_cells * kSlots);
  std::vector<unsigned char> fill(n_cells, 0);
  Xoshiro rng(seed);
  int have = 0;
  int64_t attempts = 0;
  auto wrap = [&](double d) { return d - box * std::nearbyint(d / box); };
  while (have < n) {
    if (max_attempts > 0 && attempts >= max_attempts) return FLARE_EINVAL;
    attempts++;
    const double c[3] = {rng.uniform() * box, rng.uniform() * box, rng.uniform() * box};
    int ci[3];
    for (int a = 0; a < 3; a++) ci[a] = std::min(nc - 1, (int)(c[a] / edge));
    bool ok = true;
    for (int dx = -1; dx <= 1 && ok; dx++) {
      const int x = (ci[0] + dx + nc) % nc;
      for (int dy = -1; dy <= 1 && ok; dy++) {
        const int y = (ci[1] + dy + nc) % nc;
        for (int dz = -1; dz <= 1 && ok; dz++) {
          const size_t cell = ((size_t)x * nc + y) * nc + (ci[2] + dz + nc) % nc;
          const int *js = &slot[cell * kSlots];
          for (int q = 0; q < fill[cell]; q++) {
            const double *p = positions + 3 * (size_t)js[q];
            const double u = wrap(c[0] - p[0]), v = wrap(c[1] - p[1]),
                         w = wrap(c[2] - p[2]);
            if (u * u + v * v + w * w < d2) {
              ok = false;
              break;
            }
          }
        }
      }
    }
    if (!ok) continue;
    const size_t cell = ((size_t)ci[0] * nc + ci[1]) * nc + ci[2];
    if (fill[cell] == kSlots) return FLARE_EINVAL;  // not reachable with edge < 2 d_min
    std::memcpy(positions + 3 * (size_t)have, c, sizeof(c));
    slot[cell * kSlots + fill[cell]++] = have;
    have++;
  }
  return FLARE_OK;
}
