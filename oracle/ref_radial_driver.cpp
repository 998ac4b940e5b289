// ref_radial_driver.cpp -- test infrastructure: prints the reference's own
// calculate_radial(chebyshev, quadratic|cosine) values for a fixed set of
// pair vectors, one JSON object per line.  Built by oracle/Makefile against
// radial.cpp and cutoffs.cpp from the flare_pp source tree, into oracle/_ref.
#include "cutoffs.h"
#include "radial.h"
#include <cmath>
#include <cstdio>
#include <vector>

static void emit(const char *cut, double x, double y, double z, double rcut,
                 int N) {
  std::vector<double> g(N), gx(N), gy(N), gz(N);
  double r = std::sqrt(x * x + y * y + z * z);
  std::vector<double> rh{0.0, rcut}, ch;
  if (cut[0] == 'q')
    calculate_radial(g, gx, gy, gz, chebyshev, quadratic_cutoff, x, y, z, r,
                     rcut, N, rh, ch);
  else
    calculate_radial(g, gx, gy, gz, chebyshev, cos_cutoff, x, y, z, r, rcut,
                     N, rh, ch);
  std::printf("{\"cutoff\": \"%s\", \"x\": %.17g, \"y\": %.17g, \"z\": %.17g, "
              "\"rcut\": %.17g, \"N\": %d", cut, x, y, z, rcut, N);
  const char *names[4] = {"g", "gx", "gy", "gz"};
  std::vector<double> *vals[4] = {&g, &gx, &gy, &gz};
  for (int a = 0; a < 4; a++) {
    std::printf(", \"%s\": [", names[a]);
    for (int n = 0; n < N; n++)
      std::printf("%s%.17g", n ? ", " : "", (*vals[a])[n]);
    std::printf("]");
  }
  std::printf("}\n");
}

int main() {
  // pair vectors: the reference radial test point (ctests/test_radial.cpp:12-17),
  // fcc shell vectors at a = 4.05 and a few generic directions
  const double pts[][3] = {{2.19, 1.23, -0.24}, {2.025, 2.025, 0.0},
                           {4.05, 0.0, 0.0},    {4.05, 2.025, 2.025},
                           {-1.3, 0.7, 2.9},    {0.31, -0.52, 0.11},
                           {3.1, -2.2, 1.9},    {-0.9, -3.3, -2.4}};
  const double rcuts[] = {5.0, 7.0, 4.5, 5.5};
  for (const char *cut : {"quadratic", "cosine"})
    for (auto &p : pts)
      for (double rc : rcuts) emit(cut, p[0], p[1], p[2], rc, 8);
  return 0;
}
