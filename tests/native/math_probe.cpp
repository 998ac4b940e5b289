// Host build of flare_pp_amd/csrc/flare_math.h (the per-pair math of the
// kernels) exposed to the tests through a C ABI.
#include "flare_math.h"

extern "C" {
void probe_radial(double r, double rcut, double r1, int N, int kind, double *g,
                  double *dg) {
  flare::radial_basis<double>(r, rcut, r1, N, kind, g, dg);
}
void probe_radial_f32(float r, float rcut, float r1, int N, int kind, float *g,
                      float *dg) {
  flare::radial_basis<float>(r, rcut, r1, N, kind, g, dg);
}
void probe_ylm(double x, double y, double z, int L, double *Y, double *Yx,
               double *Yy, double *Yz) {
  double norm[flare::kNormTableSize];
  flare::ylm_norm_table(norm);
  double r = std::sqrt(x * x + y * y + z * z);
  flare::ylm_visit<double, true>(x / r, y / r, z / r, 1.0 / r, L, norm,
                                 [&](int i, double v, double gx, double gy,
                                     double gz) {
                                   Y[i] = v;
                                   Yx[i] = gx;
                                   Yy[i] = gy;
                                   Yz[i] = gz;
                                 });
}
int probe_f16_row_scale(float m, int *e) { return flare::f16_row_scale(m, *e) ? 1 : 0; }
int probe_pair_index(int c1, int c2, int nr) {
  return flare::b2_pair_index(c1, c2, nr);
}
}
