// oracle.cpp -- FP64 CPU restatement of flare++'s force-evaluation path.
//
// TEST INFRASTRUCTURE ONLY: the checker for libflare_hip.so and the CPU
// baseline of the benchmark.  Nothing in the product links or calls it.
//
// Built with -ffp-contract=off so that a*b+c rounds twice, as in the
// reference's default x86-64 Release build.  Each section names the
// reference lines it restates.  How each part is pinned (tests/):
//   geometry / neighbor list   brute force IS the reference loop
//                              (structure.cpp:52-118); the cell-list build is
//                              checked against it; wrapped - raw = lattice
//                              vector (ctests/test_structure.cpp:3-22)
//   radial / cutoff            golden vectors from the reference's own
//                              radial.cpp + cutoffs.cpp compiled standalone
//                              (oracle/_ref) and finite differences
//   Y_lm                       closed forms l<=3, orthonormality by
//                              quadrature, addition theorem, finite
//                              differences (ctests/test_y_grad.cpp:141-276)
//   single bond / B2           brute-force triple loops, rotation and
//                              translation invariance, finite differences of
//                              dB/dr
//   envs_envs / envs_struc     sum rules (ctests/test_kernels.cpp:35-75),
//                              finite differences of energy vs force/stress
//                              columns, diag = sigma^2
//   Cholesky / variance        L L^T = K + jitter I, L^-1 L = I, variance
//                              >= 0, diag identity
//   mapping / pair_flare       mapped E/F/S == GP mean for power 1 and 2
//                              (tests/test_lammps.py:97-102), FD forces,
//                              virial == -V * GP stress
//   varmap / std/atom          std/atom from the L^-1 + sparse-descriptor
//                              files and from the variance map equal the
//                              signed square roots of the GP local variances
//                              over sigma^2 (tests/test_lammps.py:207-237)
// Bit-level Eigen reduction orders in GEMV/GEMM/LLT are not restated
// (summations run in index order); those agree to rounding only.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.14159265358979323846;  // cutoffs.cpp:4

// ---------------------------------------------------------------------------
// 3x3 helpers following Eigen's dynamic-size code paths.
// Products: element (i,j) = ((a_i0 b_0j + a_i1 b_1j) + a_i2 b_2j).

void matmul3(const double *a, const double *b, double *c) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = a[i * 3 + 0] * b[0 * 3 + j];
      s = s + a[i * 3 + 1] * b[1 * 3 + j];
      s = s + a[i * 3 + 2] * b[2 * 3 + j];
      t[i * 3 + j] = s;
    }
  std::memcpy(c, t, sizeof(t));
}

void transpose3(const double *a, double *t) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[j * 3 + i] = a[i * 3 + j];
}

// In-place partial-pivot LU (Eigen PartialPivLU, unblocked path for
// size <= 16).  Returns number of transpositions; trans[k] = pivot row.
int plu3(double lu[3][3], int trans[3]) {
  int ntrans = 0;
  for (int k = 0; k < 3; k++) {
    int p = k;
    double best = std::fabs(lu[k][k]);
    for (int i = k + 1; i < 3; i++) {
      if (std::fabs(lu[i][k]) > best) {
        best = std::fabs(lu[i][k]);
        p = i;
      }
    }
    trans[k] = p;
    if (best != 0) {
      if (p != k) {
        for (int j = 0; j < 3; j++) std::swap(lu[k][j], lu[p][j]);
        ntrans++;
      }
      for (int i = k + 1; i < 3; i++) lu[i][k] = lu[i][k] / lu[k][k];
    }
    for (int i = k + 1; i < 3; i++)
      for (int j = k + 1; j < 3; j++) lu[i][j] = lu[i][j] - lu[i][k] * lu[k][j];
  }
  return ntrans;
}

}  // namespace

extern "C" void orc_inverse3(const double a[9], double out[9]) {
  double lu[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) lu[i][j] = a[i * 3 + j];
  int trans[3];
  plu3(lu, trans);
  double x[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int k = 0; k < 3; k++)
    if (trans[k] != k)
      for (int j = 0; j < 3; j++) std::swap(x[k][j], x[trans[k]][j]);
  // unit-lower forward solve, column oriented (TRSM small-panel path)
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      double b = x[i][j];
      for (int r = i + 1; r < 3; r++) x[r][j] = x[r][j] - b * lu[r][i];
    }
  // upper backward solve: multiply by the reciprocal of the pivot
  for (int j = 0; j < 3; j++)
    for (int i = 2; i >= 0; i--) {
      double a_inv = 1.0 / lu[i][i];
      double b = x[i][j] * a_inv;
      x[i][j] = b;
      for (int r = 0; r < i; r++) x[r][j] = x[r][j] - b * lu[r][i];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) out[i * 3 + j] = x[i][j];
}

extern "C" double orc_volume(const double cell[9]) {
  double lu[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) lu[i][j] = cell[i * 3 + j];
  int trans[3];
  int nt = plu3(lu, trans);
  double prod = lu[0][0] * lu[1][1];
  prod = prod * lu[2][2];
  double det = (nt % 2 ? -1.0 : 1.0) * prod;
  return std::fabs(det);
}

// structure.cpp:135-165
extern "C" double orc_single_sweep_cutoff(const double cell[9]) {
  const double *v1 = cell, *v2 = cell + 3, *v3 = cell + 6;
  double a_dot_b = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
  double a_dot_c = v1[0] * v3[0] + v1[1] * v3[1] + v1[2] * v3[2];
  double b_dot_c = v2[0] * v3[0] + v2[1] * v3[1] + v2[2] * v3[2];
  double a = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  double b = std::sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  double c = std::sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
  double m[6];
  m[0] = a * std::sqrt(1 - std::pow(a_dot_b / (a * b), 2));
  m[1] = b * std::sqrt(1 - std::pow(a_dot_b / (a * b), 2));
  m[2] = a * std::sqrt(1 - std::pow(a_dot_c / (a * c), 2));
  m[3] = c * std::sqrt(1 - std::pow(a_dot_c / (a * c), 2));
  m[4] = b * std::sqrt(1 - std::pow(b_dot_c / (b * c), 2));
  m[5] = c * std::sqrt(1 - std::pow(b_dot_c / (b * c), 2));
  double s = m[0];
  for (int i = 0; i < 6; i++)
    if (m[i] < s) s = m[i];
  return s;
}

namespace {
struct CellMats {
  double cell[9], ct[9], cti[9], cd[9], cdi[9];
};
void cell_mats(const double cell[9], CellMats &m) {
  std::memcpy(m.cell, cell, sizeof(m.cell));
  transpose3(cell, m.ct);
  orc_inverse3(m.ct, m.cti);
  matmul3(cell, m.ct, m.cd);
  orc_inverse3(m.cd, m.cdi);
}
// row vector (1x3) times 3x3, left to right
inline void rowmul(const double *v, const double *b, double *out) {
  for (int j = 0; j < 3; j++) {
    double s = v[0] * b[j];
    s = s + v[1] * b[3 + j];
    s = s + v[2] * b[6 + j];
    out[j] = s;
  }
}
// structure.cpp:120-133; also returns the wrapped fractional coordinates
void wrap_one(const CellMats &m, const double *p, double *w, double *frac) {
  double t[3], rel[3];
  rowmul(p, m.ct, t);
  rowmul(t, m.cdi, rel);
  double rw[3];
  for (int c = 0; c < 3; c++) rw[c] = rel[c] - std::floor(rel[c]);
  double t2[3];
  rowmul(rw, m.cd, t2);
  rowmul(t2, m.cti, w);
  if (frac)
    for (int c = 0; c < 3; c++) frac[c] = rw[c];
}
}  // namespace

extern "C" void orc_wrap_positions(const double cell[9], int n,
                                   const double *pos, double *wrapped) {
  CellMats m;
  cell_mats(cell, m);
  for (int i = 0; i < n; i++) wrap_one(m, pos + 3 * i, wrapped + 3 * i, nullptr);
}

// ---------------------------------------------------------------------------
// Neighbor list (structure.cpp:52-118)

struct orc_nlist {
  int n = 0;
  int sweep = 0;
  double cutoff = 0, volume = 0, single_sweep = 0;
  double cell[9];
  std::vector<double> wrapped;
  std::vector<int> count, cum, idx, nspec, shift;
  std::vector<double> rel;
};

namespace {

// relative image vector and the reference predicate (structure.cpp:68-77)
inline bool image_test(const double *wi, const double *wj, const double *cell,
                       int s1, int s2, int s3, double cutoff, double *out4) {
  double im[3];
  for (int c = 0; c < 3; c++) {
    double d = wj[c] - wi[c];
    d = d + (double)s1 * cell[0 + c];
    d = d + (double)s2 * cell[3 + c];
    d = d + (double)s3 * cell[6 + c];
    im[c] = d;
  }
  double dist = std::sqrt(im[0] * im[0] + im[1] * im[1] + im[2] * im[2]);
  if ((dist < cutoff) && (dist != 0)) {
    out4[0] = dist;
    out4[1] = im[0];
    out4[2] = im[1];
    out4[3] = im[2];
    return true;
  }
  return false;
}

struct Cand {
  int64_t key;
  int j, s1, s2, s3;
  double v[4];
};

void prepare(orc_nlist *h, const double cell[9], int n, const double *pos,
             double cutoff, std::vector<double> *frac) {
  h->n = n;
  h->cutoff = cutoff;
  std::memcpy(h->cell, cell, sizeof(h->cell));
  h->single_sweep = orc_single_sweep_cutoff(cell);
  h->volume = orc_volume(cell);
  h->sweep = (int)std::ceil(cutoff / h->single_sweep);
  CellMats m;
  cell_mats(cell, m);
  h->wrapped.resize((size_t)3 * n);
  if (frac) frac->resize((size_t)3 * n);
  for (int i = 0; i < n; i++)
    wrap_one(m, pos + 3 * i, &h->wrapped[3 * i], frac ? &(*frac)[3 * i] : nullptr);
}

void finalize(orc_nlist *h, std::vector<std::vector<Cand>> &per_atom,
              const int *species) {
  int n = (int)per_atom.size();
  h->count.assign(n, 0);
  h->cum.assign(n + 1, 0);
  for (int i = 0; i < n; i++) h->count[i] = (int)per_atom[i].size();
  for (int i = 1; i <= n; i++) h->cum[i] = h->cum[i - 1] + h->count[i - 1];
  int64_t P = h->cum[n];
  h->idx.resize(P);
  h->nspec.resize(P);
  h->rel.resize(P * 4);
  h->shift.resize(P * 3);
  for (int i = 0; i < n; i++) {
    int64_t o = h->cum[i];
    for (size_t k = 0; k < per_atom[i].size(); k++) {
      const Cand &c = per_atom[i][k];
      h->idx[o + k] = c.j;
      h->nspec[o + k] = species[c.j];
      for (int q = 0; q < 4; q++) h->rel[(o + k) * 4 + q] = c.v[q];
      h->shift[(o + k) * 3 + 0] = c.s1;
      h->shift[(o + k) * 3 + 1] = c.s2;
      h->shift[(o + k) * 3 + 2] = c.s3;
    }
  }
}

inline int64_t cand_key(int j, int s1, int s2, int s3, int sweep) {
  int64_t u = 2 * sweep + 1;
  return (((int64_t)j * u + (s1 + sweep)) * u + (s2 + sweep)) * u + (s3 + sweep);
}

void brute_atom(const orc_nlist *h, int i, std::vector<Cand> &out) {
  const int sw = h->sweep;
  const double *wi = &h->wrapped[3 * i];
  for (int j = 0; j < h->n; j++) {
    const double *wj = &h->wrapped[3 * j];
    for (int s1 = -sw; s1 < sw + 1; s1++)
      for (int s2 = -sw; s2 < sw + 1; s2++)
        for (int s3 = -sw; s3 < sw + 1; s3++) {
          Cand c;
          if (image_test(wi, wj, h->cell, s1, s2, s3, h->cutoff, c.v)) {
            c.j = j;
            c.s1 = s1;
            c.s2 = s2;
            c.s3 = s3;
            c.key = cand_key(j, s1, s2, s3, sw);
            out.push_back(c);
          }
        }
  }
}

}  // namespace

extern "C" orc_nlist *orc_nlist_brute(const double cell[9], int n,
                                      const double *pos, const int *species,
                                      double cutoff) {
  orc_nlist *h = new orc_nlist;
  prepare(h, cell, n, pos, cutoff, nullptr);
  std::vector<std::vector<Cand>> per(n);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n; i++) brute_atom(h, i, per[i]);
  finalize(h, per, species);
  return h;
}

extern "C" orc_nlist *orc_nlist_brute_subset(const double cell[9], int n,
                                             const double *pos,
                                             const int *species, double cutoff,
                                             int n_sub, const int *atoms) {
  orc_nlist *h = new orc_nlist;
  prepare(h, cell, n, pos, cutoff, nullptr);
  std::vector<std::vector<Cand>> per(n_sub);
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < n_sub; a++) brute_atom(h, atoms[a], per[a]);
  finalize(h, per, species);
  return h;
}

// Cell-list build producing the same set and order as the brute force.
// Bins in fractional coordinates; along an axis with >= 3 bins only the
// adjacent bins can hold an image within the cutoff (bin width >= cutoff /
// perpendicular width); along a thinner axis every bin and every shift in
// [-sweep, sweep] is tried.  Candidates go through the reference predicate
// and are sorted by (j, s1, s2, s3).
extern "C" orc_nlist *orc_nlist_cells(const double cell[9], int n,
                                      const double *pos, const int *species,
                                      double cutoff, int nthreads) {
  orc_nlist *h = new orc_nlist;
  std::vector<double> frac;
  prepare(h, cell, n, pos, cutoff, &frac);
  const int sw = h->sweep;
  // perpendicular widths: V / |a_j x a_k|
  const double *a0 = cell, *a1 = cell + 3, *a2 = cell + 6;
  auto cross_norm = [](const double *u, const double *v) {
    double c0 = u[1] * v[2] - u[2] * v[1];
    double c1 = u[2] * v[0] - u[0] * v[2];
    double c2 = u[0] * v[1] - u[1] * v[0];
    return std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  };
  double width[3] = {h->volume / cross_norm(a1, a2),
                     h->volume / cross_norm(a2, a0),
                     h->volume / cross_norm(a0, a1)};
  int nb[3];
  for (int k = 0; k < 3; k++) {
    int b = (int)std::floor(width[k] / (cutoff * (1.0 + 1e-9)));
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    nb[k] = b;
  }
  int nbins = nb[0] * nb[1] * nb[2];
  std::vector<int> bin_of(n), bin_count(nbins + 1, 0);
  for (int i = 0; i < n; i++) {
    int b[3];
    for (int k = 0; k < 3; k++) {
      int v = (int)std::floor(frac[3 * i + k] * nb[k]);
      b[k] = std::min(std::max(v, 0), nb[k] - 1);
    }
    bin_of[i] = (b[0] * nb[1] + b[1]) * nb[2] + b[2];
    bin_count[bin_of[i] + 1]++;
  }
  for (int b = 0; b < nbins; b++) bin_count[b + 1] += bin_count[b];
  std::vector<int> bin_atoms(n), fillp(bin_count.begin(), bin_count.end() - 1);
  for (int i = 0; i < n; i++) bin_atoms[fillp[bin_of[i]]++] = i;

  // per-axis (bin offset target, shift) candidate lists for a home bin
  auto axis_list = [&](int k, int home, std::vector<std::pair<int, int>> &out) {
    out.clear();
    if (nb[k] >= 3) {
      for (int o = -1; o <= 1; o++) {
        int t = home + o, s = 0;
        if (t < 0) {
          t += nb[k];
          s = -1;
        } else if (t >= nb[k]) {
          t -= nb[k];
          s = 1;
        }
        out.push_back({t, s});
      }
    } else {
      for (int t = 0; t < nb[k]; t++)
        for (int s = -sw; s <= sw; s++) out.push_back({t, s});
    }
  };

  std::vector<std::vector<Cand>> per(n);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<std::pair<int, int>> lx, ly, lz;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
      int hb = bin_of[i];
      int hx = hb / (nb[1] * nb[2]), hy = (hb / nb[2]) % nb[1], hz = hb % nb[2];
      axis_list(0, hx, lx);
      axis_list(1, hy, ly);
      axis_list(2, hz, lz);
      const double *wi = &h->wrapped[3 * i];
      std::vector<Cand> &out = per[i];
      for (auto &px : lx)
        for (auto &py : ly)
          for (auto &pz : lz) {
            if (std::abs(px.second) > sw || std::abs(py.second) > sw ||
                std::abs(pz.second) > sw)
              continue;
            int b = (px.first * nb[1] + py.first) * nb[2] + pz.first;
            for (int q = bin_count[b]; q < bin_count[b + 1]; q++) {
              int j = bin_atoms[q];
              Cand c;
              if (image_test(wi, &h->wrapped[3 * j], h->cell, px.second,
                             py.second, pz.second, cutoff, c.v)) {
                c.j = j;
                c.s1 = px.second;
                c.s2 = py.second;
                c.s3 = pz.second;
                c.key = cand_key(j, c.s1, c.s2, c.s3, sw);
                out.push_back(c);
              }
            }
          }
      std::sort(out.begin(), out.end(),
                [](const Cand &a, const Cand &b) { return a.key < b.key; });
      // the same (j, shift) can be reached twice only if an axis has < 3
      // bins and offsets alias; drop duplicates
      out.erase(std::unique(out.begin(), out.end(),
                            [](const Cand &a, const Cand &b) {
                              return a.key == b.key;
                            }),
                out.end());
    }
  }
  finalize(h, per, species);
  return h;
}

extern "C" void orc_nlist_free(orc_nlist *h) { delete h; }
extern "C" int64_t orc_nlist_n_pairs(const orc_nlist *h) {
  return h->cum.empty() ? 0 : h->cum.back();
}
extern "C" int orc_nlist_n_atoms(const orc_nlist *h) {
  return (int)h->count.size();
}
extern "C" int orc_nlist_sweep(const orc_nlist *h) { return h->sweep; }
extern "C" void orc_nlist_get(const orc_nlist *h, int *count, int *cum,
                              int *idx, int *nspec, double *rel, int *shift,
                              double *wrapped) {
  size_t n = h->count.size();
  size_t P = h->idx.size();
  if (count) std::copy(h->count.begin(), h->count.end(), count);
  if (cum) std::copy(h->cum.begin(), h->cum.end(), cum);
  if (idx) std::copy(h->idx.begin(), h->idx.end(), idx);
  if (nspec) std::copy(h->nspec.begin(), h->nspec.end(), nspec);
  if (rel) std::copy(h->rel.begin(), h->rel.end(), rel);
  if (shift) std::copy(h->shift.begin(), h->shift.end(), shift);
  if (wrapped) std::copy(h->wrapped.begin(), h->wrapped.begin() + 3 * h->n, wrapped);
  (void)n;
  (void)P;
}

// ---------------------------------------------------------------------------
// Radial basis (radial.cpp:138-169, 344-379) and cutoffs (cutoffs.cpp:44-70)

extern "C" void orc_radial(double x, double y, double z, double r,
                           double rcut, int N, double r1, int cutoff_kind,
                           double *g, double *gx, double *gy, double *gz) {
  double rc0 = 0, rc1 = 0;
  if (!(r > rcut)) {
    if (cutoff_kind == ORC_CUTOFF_QUADRATIC) {
      double rdiff = r - rcut;
      rc0 = rdiff * rdiff;
      rc1 = 2 * rdiff;
    } else {
      rc0 = (1. / 2.) * (std::cos(kPi * r / rcut) + 1);
      rc1 = -kPi * std::sin(kPi * r / rcut) / (2 * rcut);
    }
  }
  std::vector<double> bv(N, 0), bd(N, 0);
  double r2 = rcut;
  if (!((r < r1) || (r > r2))) {
    double c = 1 / (r2 - r1);
    double xx = (r - r1) * c;
    for (int n = 0; n < N; n++) {
      if (n == 0) {
        bv[n] = 1;
        bd[n] = 0;
      } else if (n == 1) {
        bv[n] = xx;
        bd[n] = c;
      } else {
        bv[n] = 2 * xx * bv[n - 1] - bv[n - 2];
        bd[n] = 2 * bv[n - 1] * c + 2 * xx * bd[n - 1] - bd[n - 2];
      }
    }
  }
  double xrel = x / r, yrel = y / r, zrel = z / r;
  for (int n = 0; n < N; n++) {
    g[n] = bv[n] * rc0;
    gx[n] = bd[n] * xrel * rc0 + bv[n] * xrel * rc1;
    gy[n] = bd[n] * yrel * rc0 + bv[n] * yrel * rc1;
    gz[n] = bd[n] * zrel * rc0 + bv[n] * zrel * rc1;
  }
}

// ---------------------------------------------------------------------------
// Real spherical harmonics.  Convention of y_grad.cpp:667-2754 (index
// l*l + l + m, m = -l..l; m < 0 ~ sin, m > 0 ~ cos, no Condon-Shortley
// phase).  Written from the definition
//   Y_lm = N_lm Q_l^|m|(u_z) {Re, Im}(u_x + i u_y)^|m|,  Q_l^m = d^m P_l/dt^m,
// with Cartesian gradients from the chain rule through u = r/|r|.

extern "C" void orc_ylm(double x, double y, double z, int L, double *Y,
                        double *Yx, double *Yy, double *Yz) {
  double r = std::sqrt(x * x + y * y + z * z);
  double u[3] = {x / r, y / r, z / r};
  std::vector<double> C(L + 2), S(L + 2);
  C[0] = 1;
  S[0] = 0;
  for (int m = 1; m <= L + 1; m++) {
    C[m] = C[m - 1] * u[0] - S[m - 1] * u[1];
    S[m] = S[m - 1] * u[0] + C[m - 1] * u[1];
  }
  // Q[l][m] for 0 <= m <= l+1 (Q[l][l+1] = 0)
  std::vector<double> Q((L + 1) * (L + 2), 0.0);
  auto q = [&](int l, int m) -> double & { return Q[l * (L + 2) + m]; };
  double t = u[2];
  for (int m = 0; m <= L; m++) {
    double dfact = 1;  // (2m-1)!!
    for (int k = 1; k <= 2 * m - 1; k += 2) dfact *= k;
    q(m, m) = dfact;
    if (m + 1 <= L) q(m + 1, m) = (2 * m + 1) * t * dfact;
    for (int l = m + 2; l <= L; l++)
      q(l, m) = ((2 * l - 1) * t * q(l - 1, m) - (l + m - 1) * q(l - 2, m)) /
                (l - m);
  }
  for (int l = 0; l <= L; l++) {
    for (int m = -l; m <= l; m++) {
      int mm = m < 0 ? -m : m;
      double ratio = 1;  // (l-mm)!/(l+mm)!
      for (int k = l - mm + 1; k <= l + mm; k++) ratio /= k;
      double norm = std::sqrt((2 * l + 1) / (4 * kPi) * ratio);
      if (m != 0) norm *= std::sqrt(2.0);
      double Qv = q(l, mm);
      double Qd = (mm + 1 <= l) ? q(l, mm + 1) : 0.0;
      double A, dA0, dA1;  // angular factor and its d/du_x, d/du_y
      if (m == 0) {
        A = 1;
        dA0 = 0;
        dA1 = 0;
      } else if (m > 0) {
        A = C[mm];
        dA0 = mm * C[mm - 1];
        dA1 = -mm * S[mm - 1];
      } else {
        A = S[mm];
        dA0 = mm * S[mm - 1];
        dA1 = mm * C[mm - 1];
      }
      int idx = l * l + l + m;
      Y[idx] = norm * Qv * A;
      // gradient of G(u) = Q(u_z) A(u_x, u_y) projected: (dG - u (u.dG)) / r
      double dG[3] = {Qv * dA0, Qv * dA1, Qd * A};
      double udg = mm * Qv * A + u[2] * Qd * A;
      Yx[idx] = norm * (dG[0] - u[0] * udg) / r;
      Yy[idx] = norm * (dG[1] - u[1] * udg) / r;
      Yz[idx] = norm * (dG[2] - u[2] * udg) / r;
    }
  }
}

// ---------------------------------------------------------------------------
// B2 descriptor (b2.cpp:76-399)

struct orc_b2 {
  int n = 0, S = 0, N = 0, L = 0, H = 0, W = 0, F = 0;
  double volume = 0;
  bool dervs = false;
  std::vector<int> species;
  std::vector<int> count, cum, nidx;  // filtered CSR, atom order
  std::vector<double> coords;         // P' x 3
  std::vector<double> A, dA;          // n x W, 3P' x W
  std::vector<double> B, dB;          // n x F, 3P' x F
  std::vector<double> norms, fdots;   // n, 3P'
};

extern "C" orc_b2 *orc_b2_compute(const orc_nlist *nl, const int *species,
                                  int S, int N, int L, double r1,
                                  int cutoff_kind, const double *cutoffs,
                                  int want_dervs) {
  orc_b2 *h = new orc_b2;
  const int n = nl->n;
  h->n = n;
  h->S = S;
  h->N = N;
  h->L = L;
  h->H = (L + 1) * (L + 1);
  h->W = S * N * h->H;
  int nr = S * N;
  h->F = nr * (nr + 1) / 2 * (L + 1);
  h->volume = nl->volume;
  h->dervs = want_dervs != 0;
  h->species.assign(species, species + n);
  const int H = h->H, W = h->W, F = h->F;

  // single_bond_multiple_cutoffs: filter r <= rcut(s_i, s_j) (b2.cpp:264-281)
  h->count.assign(n, 0);
  for (int i = 0; i < n; i++)
    for (int k = nl->cum[i]; k < nl->cum[i + 1]; k++) {
      double rcut = cutoffs[species[i] * S + nl->nspec[k]];
      if (nl->rel[(size_t)k * 4] <= rcut) h->count[i]++;
    }
  h->cum.assign(n + 1, 0);
  for (int i = 1; i <= n; i++) h->cum[i] = h->cum[i - 1] + h->count[i - 1];
  int64_t P2 = h->cum[n];
  h->nidx.resize(P2);
  h->coords.assign(P2 * 3, 0.0);
  h->A.assign((size_t)n * W, 0.0);
  if (h->dervs) h->dA.assign((size_t)P2 * 3 * W, 0.0);

#pragma omp parallel
  {
    std::vector<double> g(N), gx(N), gy(N), gz(N), Y(H), Yx(H), Yy(H), Yz(H);
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
      int64_t ni = h->cum[i];
      int si = species[i];
      for (int k = nl->cum[i]; k < nl->cum[i + 1]; k++) {
        int sj = nl->nspec[k];
        double rcut = cutoffs[si * S + sj];
        const double *rv = &nl->rel[(size_t)k * 4];
        double r = rv[0];
        if (r > rcut) continue;
        double x = rv[1], y = rv[2], z = rv[3];
        h->nidx[ni] = nl->idx[k];
        h->coords[ni * 3 + 0] = x;
        h->coords[ni * 3 + 1] = y;
        h->coords[ni * 3 + 2] = z;
        orc_radial(x, y, z, r, rcut, N, r1, cutoff_kind, g.data(), gx.data(),
                   gy.data(), gz.data());
        orc_ylm(x, y, z, L, Y.data(), Yx.data(), Yy.data(), Yz.data());
        int dc = sj * N * H;
        for (int rn = 0; rn < N; rn++) {
          for (int a = 0; a < H; a++) {
            double bond = g[rn] * Y[a];
            h->A[(size_t)i * W + dc] += bond;
            if (h->dervs) {
              double bx = gx[rn] * Y[a] + g[rn] * Yx[a];
              double by = gy[rn] * Y[a] + g[rn] * Yy[a];
              double bz = gz[rn] * Y[a] + g[rn] * Yz[a];
              h->dA[(size_t)(ni * 3 + 0) * W + dc] += bx;
              h->dA[(size_t)(ni * 3 + 1) * W + dc] += by;
              h->dA[(size_t)(ni * 3 + 2) * W + dc] += bz;
            }
            dc++;
          }
        }
        ni++;
      }
    }
  }

  // compute_b2 (b2.cpp:181-242)
  h->B.assign((size_t)n * F, 0.0);
  h->norms.assign(n, 0.0);
  if (h->dervs) {
    h->dB.assign((size_t)P2 * 3 * F, 0.0);
    h->fdots.assign(P2 * 3, 0.0);
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int at = 0; at < n; at++) {
    int nn = h->count[at];
    int64_t fs = h->cum[at] * 3;
    const double *Aa = &h->A[(size_t)at * W];
    double *Ba = &h->B[(size_t)at * F];
    int counter = 0;
    for (int n1 = 0; n1 < nr; n1++)
      for (int n2 = n1; n2 < nr; n2++)
        for (int l = 0; l < L + 1; l++) {
          for (int m = 0; m < 2 * l + 1; m++) {
            int n1_l = n1 * H + (l * l + m);
            int n2_l = n2 * H + (l * l + m);
            Ba[counter] += Aa[n1_l] * Aa[n2_l];
            if (h->dervs) {
              for (int q = 0; q < nn * 3; q++) {
                int64_t ind = fs + q;
                h->dB[(size_t)ind * F + counter] +=
                    Aa[n1_l] * h->dA[(size_t)ind * W + n2_l] +
                    h->dA[(size_t)ind * W + n1_l] * Aa[n2_l];
              }
            }
          }
          counter++;
        }
    double s = 0;
    for (int c = 0; c < F; c++) s += Ba[c] * Ba[c];
    h->norms[at] = std::sqrt(s);
    if (h->dervs)
      for (int q = 0; q < nn * 3; q++) {
        int64_t ind = fs + q;
        double d = 0;
        for (int c = 0; c < F; c++) d += h->dB[(size_t)ind * F + c] * Ba[c];
        h->fdots[ind] = d;
      }
  }
  return h;
}

extern "C" void orc_b2_free(orc_b2 *h) { delete h; }
extern "C" int orc_b2_n_descriptors(const orc_b2 *h) { return h->F; }
extern "C" int64_t orc_b2_n_pairs(const orc_b2 *h) { return h->cum[h->n]; }

extern "C" void orc_b2_get_atom_order(const orc_b2 *h, double *single_bond,
                                      double *b2, double *norms) {
  if (single_bond) std::copy(h->A.begin(), h->A.end(), single_bond);
  if (b2) std::copy(h->B.begin(), h->B.end(), b2);
  if (norms) std::copy(h->norms.begin(), h->norms.end(), norms);
}

extern "C" void orc_b2_get_single_bond_dervs(const orc_b2 *h, double *dervs) {
  if (dervs && h->dervs) std::copy(h->dA.begin(), h->dA.end(), dervs);
}

// per-species regroup (b2.cpp:104-176)
extern "C" void orc_b2_get_grouped(const orc_b2 *h, int *n_by_type,
                                   int *neigh_by_type, double *descriptors,
                                   double *norms, double *force_dervs,
                                   double *force_dots, double *neighbor_coords,
                                   int *neighbor_counts,
                                   int *cum_neighbor_counts, int *atom_indices,
                                   int *neighbor_indices) {
  const int S = h->S, F = h->F, n = h->n;
  std::vector<int> ns(S, 0), nn(S, 0);
  for (int i = 0; i < n; i++) {
    ns[h->species[i]]++;
    nn[h->species[i]] += h->count[i];
  }
  std::vector<int64_t> row0(S, 0), pair0(S, 0);
  for (int s = 1; s < S; s++) {
    row0[s] = row0[s - 1] + ns[s - 1];
    pair0[s] = pair0[s - 1] + nn[s - 1];
  }
  if (n_by_type) std::copy(ns.begin(), ns.end(), n_by_type);
  if (neigh_by_type) std::copy(nn.begin(), nn.end(), neigh_by_type);
  std::vector<int> sc(S, 0), nc(S, 0);
  for (int i = 0; i < n; i++) {
    int s = h->species[i];
    int64_t row = row0[s] + sc[s];
    int cnt = h->count[i];
    int64_t prow = pair0[s] + nc[s];
    int64_t src = h->cum[i];
    if (descriptors)
      std::copy(&h->B[(size_t)i * F], &h->B[(size_t)i * F] + F,
                descriptors + row * F);
    if (norms) norms[row] = h->norms[i];
    if (force_dervs && h->dervs)
      std::copy(&h->dB[(size_t)src * 3 * F], &h->dB[(size_t)(src + cnt) * 3 * F],
                force_dervs + prow * 3 * F);
    if (force_dots && h->dervs)
      std::copy(&h->fdots[src * 3], &h->fdots[(src + cnt) * 3],
                force_dots + prow * 3);
    if (neighbor_coords)
      std::copy(&h->coords[src * 3], &h->coords[(src + cnt) * 3],
                neighbor_coords + prow * 3);
    if (neighbor_counts) neighbor_counts[row] = cnt;
    if (cum_neighbor_counts) cum_neighbor_counts[row] = nc[s];
    if (atom_indices) atom_indices[row] = i;
    if (neighbor_indices)
      std::copy(&h->nidx[src], &h->nidx[src + cnt], neighbor_indices + prow);
    sc[s]++;
    nc[s] += cnt;
  }
}

// ---------------------------------------------------------------------------
// NormalizedDotProduct (normalized_dot_product.cpp:27-210)

namespace {
inline double dotn(const double *a, const double *b, int F) {
  double s = 0;
  for (int c = 0; c < F; c++) s += a[c] * b[c];
  return s;
}
}  // namespace

extern "C" void orc_envs_envs(int S, int F, const int *n1_by_type,
                              const double *d1, const double *norms1,
                              const int *n2_by_type, const double *d2,
                              const double *norms2, double sigma, double power,
                              double *out) {
  double sig_sq = sigma * sigma;
  const double empty_thresh = 1e-8;
  int n1 = 0, n2 = 0;
  for (int s = 0; s < S; s++) {
    n1 += n1_by_type[s];
    n2 += n2_by_type[s];
  }
  std::fill(out, out + (size_t)n1 * n2, 0.0);
  int c1 = 0, c2 = 0;
  for (int s = 0; s < S; s++) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n1_by_type[s]; i++) {
      double norm_i = norms1[c1 + i];
      if (norm_i < empty_thresh) continue;
      for (int j = 0; j < n2_by_type[s]; j++) {
        double norm_j = norms2[c2 + j];
        double norm_ij = norm_i * norm_j;
        if (norm_j < empty_thresh) continue;
        double dv = dotn(d1 + (size_t)(c1 + i) * F, d2 + (size_t)(c2 + j) * F, F);
        double norm_dot = dv / norm_ij;
        out[(size_t)(c1 + i) * n2 + (c2 + j)] += sig_sq * std::pow(norm_dot, power);
      }
    }
    c1 += n1_by_type[s];
    c2 += n2_by_type[s];
  }
}

namespace {
// grouped view of a B2 result (descriptor-values layout)
struct Grouped {
  std::vector<int> ns, nn, counts, cumc, atoms, nidx;
  std::vector<double> d, norms, fd, fdots, coords;
};
void grouped(const orc_b2 *h, Grouped &g, bool with_dervs) {
  int S = h->S, F = h->F, n = h->n;
  int64_t P2 = h->cum[n];
  g.ns.resize(S);
  g.nn.resize(S);
  g.d.resize((size_t)n * F);
  g.norms.resize(n);
  g.counts.resize(n);
  g.cumc.resize(n);
  g.atoms.resize(n);
  g.nidx.resize(P2);
  g.coords.resize(P2 * 3);
  if (with_dervs) {
    g.fd.resize((size_t)P2 * 3 * F);
    g.fdots.resize(P2 * 3);
  }
  orc_b2_get_grouped(h, g.ns.data(), g.nn.data(), g.d.data(), g.norms.data(),
                     with_dervs ? g.fd.data() : nullptr,
                     with_dervs ? g.fdots.data() : nullptr, g.coords.data(),
                     g.counts.data(), g.cumc.data(), g.atoms.data(),
                     g.nidx.data());
}
}  // namespace

extern "C" void orc_envs_struc(const orc_b2 *struc, const int *m_by_type,
                               const double *sparse,
                               const double *sparse_norms, double sigma,
                               double power, double volume, double *out) {
  const int S = struc->S, F = struc->F, n = struc->n;
  Grouped g;
  grouped(struc, g, true);
  const double sig_sq = sigma * sigma;
  const double vol_inv = 1 / volume;
  const double empty_thresh = 1e-8;
  int M = 0;
  for (int s = 0; s < S; s++) M += m_by_type[s];
  const int n_out = 1 + 3 * n + 6;
  std::fill(out, out + (size_t)M * n_out, 0.0);
  int c_sparse = 0, row0 = 0;
  int64_t pair0 = 0;
  for (int s = 0; s < S; s++) {
    int n_sparse = m_by_type[s];
    int n_struc = g.ns[s];
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n_sparse; i++) {
      double norm_i = sparse_norms[c_sparse + i];
      if (norm_i < empty_thresh) continue;
      int sparse_index = c_sparse + i;
      const double *p = sparse + (size_t)sparse_index * F;
      double *krow = out + (size_t)sparse_index * n_out;
      for (int j = 0; j < n_struc; j++) {
        double norm_j = g.norms[row0 + j];
        double norm_ij = norm_i * norm_j;
        double norm_ij3 = norm_ij * norm_j * norm_j;
        if (norm_j < empty_thresh) continue;
        double dot_val = dotn(p, &g.d[(size_t)(row0 + j) * F], F);
        double norm_dot = dot_val / norm_ij;
        double dval = power * std::pow(norm_dot, power - 1);
        krow[0] += sig_sq * std::pow(norm_dot, power);
        int n_neigh = g.counts[row0 + j];
        int c_neigh = g.cumc[row0 + j];
        int atom_index = g.atoms[row0 + j];
        for (int k = 0; k < n_neigh; k++) {
          int64_t ind = pair0 + c_neigh + k;
          int neighbor_index = g.nidx[ind];
          int stress_counter = 0;
          for (int comp = 0; comp < 3; comp++) {
            int64_t force_index = 3 * ind + comp;
            double fdot = dotn(p, &g.fd[(size_t)force_index * F], F);
            double f1 = fdot / norm_ij;
            double f2 = dot_val * g.fdots[force_index] / norm_ij3;
            double f3 = f1 - f2;
            double force_kern_val = sig_sq * dval * f3;
            krow[1 + 3 * neighbor_index + comp] -= force_kern_val;
            krow[1 + 3 * atom_index + comp] += force_kern_val;
            for (int comp2 = comp; comp2 < 3; comp2++) {
              double coord = g.coords[ind * 3 + comp2];
              krow[1 + 3 * n + stress_counter] -= force_kern_val * coord * vol_inv;
              stress_counter++;
            }
          }
        }
      }
    }
    c_sparse += n_sparse;
    row0 += n_struc;
    pair0 += g.nn[s];
  }
}

// ---------------------------------------------------------------------------
// self_kernel_struc (normalized_dot_product.cpp:446-578): k(o, o) of every
// output o = [E, F (3n), S (6)] of a structure with itself, summed over atom
// pairs i <= j of one species (factor 2 when i != j) and their neighbor
// pairs.  Serial restatement of the reference loops.

extern "C" void orc_self_kernel_struc(const orc_b2 *struc, double sigma,
                                      double power, double volume,
                                      double *out) {
  const int S = struc->S, F = struc->F, n = struc->n;
  Grouped g;
  grouped(struc, g, true);
  const double sig_sq = sigma * sigma;
  const double vol_inv = 1 / volume;
  const double vol_inv_sq = vol_inv * vol_inv;
  const double empty_thresh = 1e-8;
  const int n_out = 1 + 3 * n + 6;
  std::fill(out, out + n_out, 0.0);
  int row0 = 0;
  int64_t pair0 = 0;
  for (int s = 0; s < S; s++) {
    const int n_struc = g.ns[s];
    const int64_t np3 = 3 * (int64_t)g.nn[s];
    // dot_vals, force_dot (3P_s x n_s) and force_force (3P_s x 3P_s)
    auto D = [&](int i) { return &g.d[(size_t)(row0 + i) * F]; };
    auto dB = [&](int64_t f) { return &g.fd[(size_t)(3 * pair0 + f) * F]; };
    std::vector<double> force_dot((size_t)np3 * n_struc);
    for (int64_t f = 0; f < np3; f++)
      for (int j = 0; j < n_struc; j++)
        force_dot[(size_t)f * n_struc + j] = dotn(dB(f), D(j), F);
    std::vector<double> par((size_t)n_struc * n_out, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n_struc; i++) {
      double norm_i = g.norms[row0 + i];
      if (norm_i < empty_thresh) continue;
      double norm_i2 = norm_i * norm_i;
      double norm_i3 = norm_i2 * norm_i;
      double *pm = &par[(size_t)i * n_out];
      for (int j = i; j < n_struc; j++) {
        double norm_j = g.norms[row0 + j];
        if (norm_j < empty_thresh) continue;
        double mult_fac = i == j ? 1 : 2;
        double norm_j2 = norm_j * norm_j;
        double norm_j3 = norm_j2 * norm_j;
        double norm_ij = norm_i * norm_j;
        double norm_dot = dotn(D(i), D(j), F) / norm_ij;
        double c1 = (power - 1) * power * std::pow(norm_dot, power - 2);
        double c2 = power * std::pow(norm_dot, power - 1);
        pm[0] += sig_sq * mult_fac * std::pow(norm_dot, power);
        int n_neigh_1 = g.counts[row0 + i], c_neigh_1 = g.cumc[row0 + i];
        int c_ind_1 = g.atoms[row0 + i];
        int n_neigh_2 = g.counts[row0 + j], c_neigh_2 = g.cumc[row0 + j];
        int c_ind_2 = g.atoms[row0 + j];
        for (int k = 0; k < n_neigh_1; k++) {
          int ind1 = c_neigh_1 + k;
          int n_ind_1 = g.nidx[pair0 + ind1];
          for (int l = 0; l < n_neigh_2; l++) {
            int ind2 = c_neigh_2 + l;
            int n_ind_2 = g.nidx[pair0 + ind2];
            int stress_counter = 0;
            for (int m = 0; m < 3; m++) {
              int64_t f1 = 3 * (int64_t)ind1 + m, f2 = 3 * (int64_t)ind2 + m;
              double sfd1 = g.fdots[3 * pair0 + f1], sfd2 = g.fdots[3 * pair0 + f2];
              double fdj = force_dot[(size_t)f1 * n_struc + j];
              double fdi = force_dot[(size_t)f2 * n_struc + i];
              double v1 = fdj / norm_ij - norm_dot * sfd1 / norm_i2;
              double v2 = fdi / norm_ij - norm_dot * sfd2 / norm_j2;
              double v3 = dotn(dB(f1), dB(f2), F) / norm_ij;
              double v4 = sfd1 * fdi / (norm_i3 * norm_j);
              double v5 = sfd2 * fdj / (norm_i * norm_j3);
              double v6 = sfd1 * sfd2 * norm_dot / (norm_i2 * norm_j2);
              double kern_val =
                  sig_sq * mult_fac * (c1 * v1 * v2 + c2 * (v3 - v4 - v5 + v6));
              if (c_ind_1 == c_ind_2) pm[1 + c_ind_1 * 3 + m] += kern_val;
              if (c_ind_1 == n_ind_2) pm[1 + c_ind_1 * 3 + m] -= kern_val;
              if (n_ind_1 == c_ind_2) pm[1 + n_ind_1 * 3 + m] -= kern_val;
              if (n_ind_1 == n_ind_2) pm[1 + n_ind_1 * 3 + m] += kern_val;
              for (int nn = m; nn < 3; nn++) {
                double coord1 = g.coords[(pair0 + ind1) * 3 + nn];
                double coord2 = g.coords[(pair0 + ind2) * 3 + nn];
                pm[1 + 3 * n + stress_counter] += kern_val * coord1 * coord2 * vol_inv_sq;
                stress_counter++;
              }
            }
          }
        }
      }
    }
    for (int i = 0; i < n_struc; i++)
      for (int c = 0; c < n_out; c++) out[c] += par[(size_t)i * n_out + c];
    row0 += n_struc;
    pair0 += g.nn[s];
  }
}

// ---------------------------------------------------------------------------
// SparseGP pieces (sparse_gp.cpp:93-133, 576-628, 682-702)

extern "C" int orc_cholesky_inverse(int M, const double *K, double jitter,
                                    double *Linv) {
  std::vector<double> Lm((size_t)M * M, 0.0);
  for (int j = 0; j < M; j++) {
    double d = K[(size_t)j * M + j] + jitter;
    for (int k = 0; k < j; k++) d -= Lm[(size_t)j * M + k] * Lm[(size_t)j * M + k];
    if (!(d > 0)) return 1;
    double ljj = std::sqrt(d);
    Lm[(size_t)j * M + j] = ljj;
    for (int i = j + 1; i < M; i++) {
      double v = K[(size_t)i * M + j];
      for (int k = 0; k < j; k++) v -= Lm[(size_t)i * M + k] * Lm[(size_t)j * M + k];
      Lm[(size_t)i * M + j] = v / ljj;
    }
  }
  // L^-1 by forward substitution, one column of the identity at a time
  std::fill(Linv, Linv + (size_t)M * M, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int c = 0; c < M; c++) {
    for (int i = c; i < M; i++) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; k++) v -= Lm[(size_t)i * M + k] * Linv[(size_t)k * M + c];
      Linv[(size_t)i * M + c] = v / Lm[(size_t)i * M + i];
    }
  }
  return 0;
}

extern "C" void orc_predict_mean(const orc_b2 *struc, const int *m_by_type,
                                 const double *sparse,
                                 const double *sparse_norms, double sigma,
                                 double power, double volume,
                                 const double *alpha, double *mean_efs) {
  int M = 0;
  for (int s = 0; s < struc->S; s++) M += m_by_type[s];
  const int n_out = 1 + 3 * struc->n + 6;
  std::vector<double> K((size_t)M * n_out);
  orc_envs_struc(struc, m_by_type, sparse, sparse_norms, sigma, power, volume,
                 K.data());
  for (int c = 0; c < n_out; c++) {
    double v = 0;
    for (int u = 0; u < M; u++) v += K[(size_t)u * n_out + c] * alpha[u];
    mean_efs[c] = v;
  }
}

// compute_cluster_uncertainties (sparse_gp.cpp:93-133) for the grouped rows
// rows[0..nrows) (ascending): each variance k(d, d) - |L^-1 k|^2 depends on
// its own descriptor only, so a subset is evaluated on its own.
extern "C" void orc_cluster_uncertainties_rows(const orc_b2 *struc,
                                               const int *m_by_type,
                                               const double *sparse,
                                               const double *sparse_norms,
                                               double sigma, double power,
                                               const double *Linv, int nrows,
                                               const int *rows, double *variances) {
  const int S = struc->S, F = struc->F;
  Grouped g;
  grouped(struc, g, false);
  int M = 0;
  for (int s = 0; s < S; s++) M += m_by_type[s];
  // the selected rows and their per-species counts (rows are species-grouped)
  std::vector<int> off(S + 1, 0), ns(S, 0);
  for (int s = 0; s < S; s++) off[s + 1] = off[s] + g.ns[s];
  std::vector<double> d((size_t)nrows * F), nrm(nrows);
  for (int q = 0; q < nrows; q++) {
    const int r = rows[q];
    for (int s = 0; s < S; s++)
      if (r >= off[s] && r < off[s + 1]) ns[s]++;
    std::copy(&g.d[(size_t)r * F], &g.d[(size_t)r * F] + F, &d[(size_t)q * F]);
    nrm[q] = g.norms[r];
  }
  // K_self = diag(envs_envs(clusters, clusters))
  std::vector<double> kself(nrows, 0.0);
  const double empty_thresh = 1e-8;
  for (int q = 0; q < nrows; q++) {
    if (nrm[q] < empty_thresh) continue;
    double dv = dotn(&d[(size_t)q * F], &d[(size_t)q * F], F);
    kself[q] = sigma * sigma * std::pow(dv / (nrm[q] * nrm[q]), power);
  }
  std::vector<double> kmat((size_t)nrows * M);
  orc_envs_envs(S, F, ns.data(), d.data(), nrm.data(), m_by_type, sparse,
                sparse_norms, sigma, power, kmat.data());
#pragma omp parallel for schedule(static)
  for (int q = 0; q < nrows; q++) {
    double qs = 0;
    for (int u = 0; u < M; u++) {
      double t = 0;
      for (int v = 0; v <= u; v++) t += Linv[(size_t)u * M + v] * kmat[(size_t)q * M + v];
      qs += t * t;
    }
    variances[q] = kself[q] - qs;
  }
}

extern "C" void orc_cluster_uncertainties(const orc_b2 *struc,
                                          const int *m_by_type,
                                          const double *sparse,
                                          const double *sparse_norms,
                                          double sigma, double power,
                                          const double *Linv,
                                          double *variances) {
  std::vector<int> rows(struc->n);
  for (int r = 0; r < struc->n; r++) rows[r] = r;
  orc_cluster_uncertainties_rows(struc, m_by_type, sparse, sparse_norms, sigma, power,
                                 Linv, struc->n, rows.data(), variances);
}

// ---------------------------------------------------------------------------
// Mapping coefficients (normalized_dot_product.cpp:643-759)

extern "C" int orc_mapping_coefficients(int S, int F, const int *m_by_type,
                                        const double *sparse,
                                        const double *sparse_norms,
                                        double sigma, double power,
                                        const double *alpha, double *out) {
  const double sig2 = sigma * sigma;
  const double empty_thresh = 1e-8;
  int ipow = (int)power;
  if (ipow != 1 && ipow != 2) return 1;
  size_t beta_size = ipow == 1 ? (size_t)F : (size_t)F * (F + 1) / 2;
  std::fill(out, out + beta_size * S, 0.0);
  int c = 0;
  for (int s = 0; s < S; s++) {
    double *row = out + beta_size * s;
    for (int j = 0; j < m_by_type[s]; j++) {
      const double *p = sparse + (size_t)(c + j) * F;
      double p_norm = sparse_norms[c + j];
      if (p_norm < empty_thresh) continue;
      double alpha_val = alpha[c + j];
      size_t bc = 0;
      for (int k = 0; k < F; k++) {
        double p_ik = p[k] / p_norm;
        if (ipow == 1) {
          row[bc] += sig2 * p_ik * alpha_val;
          bc++;
          continue;
        }
        for (int l = k; l < F; l++) {
          double p_il = p[l] / p_norm;
          double beta_val = sig2 * p_ik * p_il * alpha_val;
          if (k != l)
            row[bc] += 2 * beta_val;
          else
            row[bc] += beta_val;
          bc++;
        }
      }
    }
    c += m_by_type[s];
  }
  return 0;
}

// pair_flare.cpp:363-401
extern "C" void orc_unpack_beta(int S, int F, int power, const double *packed,
                                double *full) {
  size_t bc = 0;
  if (power == 1) {
    for (int s = 0; s < S; s++)
      for (int i = 0; i < F; i++) full[(size_t)s * F + i] = packed[bc++];
    return;
  }
  for (int s = 0; s < S; s++) {
    double *b = full + (size_t)s * F * F;
    for (int i = 0; i < F; i++)
      for (int j = i; j < F; j++) {
        if (i == j)
          b[(size_t)i * F + j] = packed[bc];
        else {
          double v = packed[bc] / 2;
          b[(size_t)i * F + j] = v;
          b[(size_t)j * F + i] = v;
        }
        bc++;
      }
  }
}

// ---------------------------------------------------------------------------
// Mapped pair style (pair_flare.cpp:68-184; lammps_descriptor.cpp:7-107,
// 209-304).  Thread-private force and virial buffers, summed in thread
// order after the atom loop (the Kokkos-OpenMP threading model).

extern "C" void orc_pair_compute_full(const orc_nlist *nl, const int *species,
                                      int S, int N, int L, int power,
                                      const double *cutoffs, int cutoff_kind,
                                      const double *beta_full, int nthreads,
                                      double *energy, double *forces,
                                      double *virial, double *eatom) {
  const int n = nl->n;
  const int H = (L + 1) * (L + 1);
  const int nr = S * N;
  const int W = nr * H;
  const int F = nr * (nr + 1) / 2 * (L + 1);
  const double empty_thresh = 1e-8;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  int nt = 1;
#endif
  std::vector<std::vector<double>> fbuf(nt);
  std::vector<double> vbuf((size_t)nt * 6, 0.0), ebuf(nt, 0.0);
  if (eatom) std::fill(eatom, eatom + n, 0.0);

#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    std::vector<double> &fl = fbuf[tid];
    fl.assign((size_t)n * 3, 0.0);
    double *vl = &vbuf[(size_t)tid * 6];
    std::vector<double> g(N), gx(N), gy(N), gz(N), Y(H), Yx(H), Yy(H), Yz(H);
    std::vector<double> A(W), B(F), Z, w(F), u(W), dA;
    std::vector<int> inner;
#pragma omp for schedule(static)
    for (int i = 0; i < n; i++) {
      int si = species[i];
      inner.clear();
      for (int k = nl->cum[i]; k < nl->cum[i + 1]; k++) {
        const double *rv = &nl->rel[(size_t)k * 4];
        double rsq = rv[1] * rv[1] + rv[2] * rv[2] + rv[3] * rv[3];
        double rc = cutoffs[si * S + nl->nspec[k]];
        if (rsq < rc * rc) inner.push_back(k);
      }
      int n_inner = (int)inner.size();
      std::fill(A.begin(), A.end(), 0.0);
      dA.assign((size_t)n_inner * 3 * W, 0.0);
      for (int q = 0; q < n_inner; q++) {
        int k = inner[q];
        const double *rv = &nl->rel[(size_t)k * 4];
        double x = rv[1], y = rv[2], z = rv[3];
        double rsq = x * x + y * y + z * z;
        double r = std::sqrt(rsq);
        int s = nl->nspec[k];
        double rc = cutoffs[si * S + s];
        orc_radial(x, y, z, r, rc, N, 0.0, cutoff_kind, g.data(), gx.data(),
                   gy.data(), gz.data());
        orc_ylm(x, y, z, L, Y.data(), Yx.data(), Yy.data(), Yz.data());
        int dc = s * N * H;
        for (int rn = 0; rn < N; rn++)
          for (int a = 0; a < H; a++) {
            A[dc] += g[rn] * Y[a];
            dA[(size_t)(q * 3 + 0) * W + dc] += gx[rn] * Y[a] + g[rn] * Yx[a];
            dA[(size_t)(q * 3 + 1) * W + dc] += gy[rn] * Y[a] + g[rn] * Yy[a];
            dA[(size_t)(q * 3 + 2) * W + dc] += gz[rn] * Y[a] + g[rn] * Yz[a];
            dc++;
          }
      }
      // B2_descriptor (lammps_descriptor.cpp:209-248)
      std::fill(B.begin(), B.end(), 0.0);
      for (int n1 = nr - 1; n1 >= 0; n1--) {
        int n1c = (n1 * (2 * nr - n1 + 1)) / 2;
        for (int n2 = n1; n2 < nr; n2++) {
          int n2c = n2 - n1;
          for (int l = 0; l < L + 1; l++) {
            int counter = l + (n1c + n2c) * (L + 1);
            for (int m = 0; m < 2 * l + 1; m++)
              B[counter] += A[n1 * H + l * l + m] * A[n2 * H + l * l + m];
          }
        }
      }
      double norm2 = 0;
      for (int c = 0; c < F; c++) norm2 += B[c] * B[c];
      if (norm2 < empty_thresh) continue;
      // compute_energy_and_u (lammps_descriptor.cpp:250-304)
      const double *beta = beta_full + (size_t)si * (power == 1 ? F : (size_t)F * F);
      double evdwl;
      if (power == 1) {
        double bn = std::pow(norm2, 0.5);
        evdwl = dotn(B.data(), beta, F) / bn;
        for (int c = 0; c < F; c++) w[c] = beta[c] / bn - evdwl * B[c] / norm2;
      } else {
        Z.assign(F, 0.0);
        for (int j = 0; j < F; j++) {
          const double *col = beta + (size_t)j * F;
          double bj = B[j];
          for (int c = 0; c < F; c++) Z[c] += col[c] * bj;
        }
        evdwl = dotn(B.data(), Z.data(), F) / norm2;
        for (int c = 0; c < F; c++) w[c] = 2 * (Z[c] - evdwl * B[c]) / norm2;
      }
      std::fill(u.begin(), u.end(), 0.0);
      for (int n1 = nr - 1; n1 >= 0; n1--)
        for (int n2 = 0; n2 < nr; n2++) {
          int n1c, n2c;
          double factor;
          if (n1 == n2) {
            n1c = (n1 * (2 * nr - n1 + 1)) / 2;
            n2c = n2 - n1;
            factor = 1.0;
          } else if (n1 < n2) {
            n1c = (n1 * (2 * nr - n1 + 1)) / 2;
            n2c = n2 - n1;
            factor = 0.5;
          } else {
            n1c = (n2 * (2 * nr - n2 + 1)) / 2;
            n2c = n1 - n2;
            factor = 0.5;
          }
          for (int l = 0; l < L + 1; l++) {
            int counter = l + (n1c + n2c) * (L + 1);
            for (int m = 0; m < 2 * l + 1; m++)
              u[n1 * H + l * l + m] += w[counter] * A[n2 * H + l * l + m] * factor;
          }
        }
      for (int c = 0; c < W; c++) u[c] *= 2;
      // pair forces (pair_flare.cpp:141-174)
      for (int q = 0; q < n_inner; q++) {
        int k = inner[q];
        const double *rv = &nl->rel[(size_t)k * 4];
        double delx = -rv[1], dely = -rv[2], delz = -rv[3];
        double fx = dotn(&dA[(size_t)(q * 3 + 0) * W], u.data(), W);
        double fy = dotn(&dA[(size_t)(q * 3 + 1) * W], u.data(), W);
        double fz = dotn(&dA[(size_t)(q * 3 + 2) * W], u.data(), W);
        int j = nl->idx[k];
        fl[3 * i + 0] += fx;
        fl[3 * i + 1] += fy;
        fl[3 * i + 2] += fz;
        fl[3 * j + 0] -= fx;
        fl[3 * j + 1] -= fy;
        fl[3 * j + 2] -= fz;
        vl[0] += delx * fx;
        vl[1] += dely * fy;
        vl[2] += delz * fz;
        vl[3] += delx * fy;
        vl[4] += delx * fz;
        vl[5] += dely * fz;
      }
      ebuf[tid] += evdwl;
      if (eatom) eatom[i] = evdwl;
    }
  }
  double e = 0;
  for (int t = 0; t < nt; t++) e += ebuf[t];
  *energy = e;
  std::fill(forces, forces + (size_t)n * 3, 0.0);
  for (int t = 0; t < nt; t++) {
    if (fbuf[t].empty()) continue;
    for (size_t q = 0; q < (size_t)n * 3; q++) forces[q] += fbuf[t][q];
  }
  for (int c = 0; c < 6; c++) {
    double v = 0;
    for (int t = 0; t < nt; t++) v += vbuf[(size_t)t * 6 + c];
    virial[c] = v;
  }
}

// ---------------------------------------------------------------------------
// Variance map (normalized_dot_product.cpp:761-852): power-1 kernels only.
// beta_s[k][l] = sum over non-empty i, j of species s of
//   sig^4 p_ik p_jl (-Kinv_ij / |p_i| / |p_j|), plus sig^2 on the diagonal,
// added inside the (i = 0, j = 0) iteration, i.e. only when the species'
// first sparse environment is non-empty.

extern "C" int orc_varmap_coefficients(int S, int F, const int *m_by_type,
                                       const double *sparse,
                                       const double *sparse_norms,
                                       double sigma, const double *Kinv,
                                       double *out) {
  const double sig2 = sigma * sigma;
  const double empty_thresh = 1e-8;
  int M = 0;
  for (int s = 0; s < S; s++) M += m_by_type[s];
  const size_t FF = (size_t)F * F;
  std::fill(out, out + FF * S, 0.0);
  int K_ind = 0;
  for (int s = 0; s < S; s++) {
    double *row = out + FF * s;
    const int nt = m_by_type[s];
    for (int i = 0; i < nt; i++) {
      const double *pi = sparse + (size_t)(K_ind + i) * F;
      double pi_norm = sparse_norms[K_ind + i];
      if (pi_norm < empty_thresh) continue;
      for (int j = 0; j < nt; j++) {
        const double *pj = sparse + (size_t)(K_ind + j) * F;
        double pj_norm = sparse_norms[K_ind + j];
        if (pj_norm < empty_thresh) continue;
        double kij = Kinv[(size_t)(K_ind + i) * M + (K_ind + j)];
        double kn = kij / pi_norm / pj_norm;
        size_t bc = 0;
        for (int k = 0; k < F; k++) {
          double p_ik = pi[k];
          for (int l = 0; l < F; l++) {
            double p_jl = pj[l];
            row[bc] += sig2 * sig2 * p_ik * p_jl * (-kn);
            if (k == l && i == 0 && j == 0) row[bc] += sig2;
            bc++;
          }
        }
      }
    }
    K_ind += nt;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// compute flare/std/atom (compute_flare_std_atom.cpp:99-213).  Per atom: the
// pair-style descriptor (rsq < rc(si,sj)^2, lammps_descriptor.cpp:7-107,
// 209-248); empty environments (|B|^2 < 1e-8) keep std 0.
//   use_map: variance = B^T beta B / |B|^2 / sig^2 with beta * B taken as
//            Eigen's matrix-vector product of the file's row-major matrix
//            (compute_flare_std_atom.cpp:186-191, 427-450;
//            lammps_descriptor.cpp:265-268).
//   else:    k_s = (P_s d)^power over the species' normalised sparse rows,
//            zero elsewhere; variance = 1 - sig^2 |L^-1 k|^2 with the full
//            lower-triangular L^-1 (compute_flare_std_atom.cpp:192-206).
// std = sqrt(variance), negative square root of |variance| when negative.

extern "C" void orc_std_atom(const orc_nlist *nl, const int *species, int S,
                             int N, int L, const double *cutoffs,
                             int cutoff_kind, double sigma, int use_map,
                             const double *beta_var, int power,
                             int n_clusters, const int *n_by_type,
                             const double *sparse_normed, const double *Linv,
                             int nthreads, double *stds) {
  const int n = nl->n;
  const int H = (L + 1) * (L + 1);
  const int nr = S * N;
  const int W = nr * H;
  const int F = nr * (nr + 1) / 2 * (L + 1);
  const double empty_thresh = 1e-8;
  const double sig2 = sigma * sigma;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  int nt = 1;
  (void)nthreads;
#endif
  std::fill(stds, stds + n, 0.0);
  std::vector<int> cum_types(S + 1, 0);
  if (!use_map)
    for (int s = 0; s < S; s++) cum_types[s + 1] = cum_types[s] + n_by_type[s];

#pragma omp parallel num_threads(nt)
  {
    std::vector<double> g(N), gx(N), gy(N), gz(N), Y(H), Yx(H), Yy(H), Yz(H);
    std::vector<double> A(W), B(F), Z(F), dn(F), kv, q;
#pragma omp for schedule(static)
    for (int i = 0; i < n; i++) {
      int si = species[i];
      std::fill(A.begin(), A.end(), 0.0);
      for (int k = nl->cum[i]; k < nl->cum[i + 1]; k++) {
        const double *rv = &nl->rel[(size_t)k * 4];
        double x = rv[1], y = rv[2], z = rv[3];
        double rsq = x * x + y * y + z * z;
        int s = nl->nspec[k];
        double rc = cutoffs[si * S + s];
        if (!(rsq < rc * rc)) continue;
        double r = std::sqrt(rsq);
        orc_radial(x, y, z, r, rc, N, 0.0, cutoff_kind, g.data(), gx.data(),
                   gy.data(), gz.data());
        orc_ylm(x, y, z, L, Y.data(), Yx.data(), Yy.data(), Yz.data());
        int dc = s * N * H;
        for (int rn = 0; rn < N; rn++)
          for (int a = 0; a < H; a++) A[dc++] += g[rn] * Y[a];
      }
      std::fill(B.begin(), B.end(), 0.0);
      for (int n1 = nr - 1; n1 >= 0; n1--) {
        int n1c = (n1 * (2 * nr - n1 + 1)) / 2;
        for (int n2 = n1; n2 < nr; n2++) {
          int n2c = n2 - n1;
          for (int l = 0; l < L + 1; l++) {
            int counter = l + (n1c + n2c) * (L + 1);
            for (int m = 0; m < 2 * l + 1; m++)
              B[counter] += A[n1 * H + l * l + m] * A[n2 * H + l * l + m];
          }
        }
      }
      double norm2 = 0;
      for (int c = 0; c < F; c++) norm2 += B[c] * B[c];
      if (norm2 < empty_thresh) continue;
      double variance;
      if (use_map) {
        const double *beta = beta_var + (size_t)si * F * F;
        std::fill(Z.begin(), Z.end(), 0.0);
        for (int j = 0; j < F; j++) {
          double bj = B[j];
          for (int c = 0; c < F; c++) Z[c] += beta[(size_t)c * F + j] * bj;
        }
        variance = dotn(B.data(), Z.data(), F) / norm2;
        variance /= sig2;
      } else {
        double bn = std::pow(norm2, 0.5);
        for (int c = 0; c < F; c++) dn[c] = B[c] / bn;
        kv.assign(n_clusters, 0.0);
        for (int u = 0; u < n_by_type[si]; u++) {
          int row = cum_types[si] + u;
          double t = dotn(sparse_normed + (size_t)row * F, dn.data(), F);
          kv[row] = std::pow(t, (double)power);
        }
        double qs = 0;
        for (int u = 0; u < n_clusters; u++) {
          double v = 0;
          for (int w = 0; w <= u; w++) v += Linv[(size_t)u * n_clusters + w] * kv[w];
          qs += v * v;
        }
        variance = 1.0 - sig2 * qs;
      }
      stds[i] = variance >= 0.0 ? std::pow(variance, 0.5)
                                : -std::pow(std::fabs(variance), 0.5);
    }
  }
}

// ---------------------------------------------------------------------------
// PairFLARE::compute on a LAMMPS full neighbor list with ghost atoms
// (pair_flare.cpp:68-184; lammps_descriptor.cpp:7-107, 209-304), serial:
// forces on all n_all atoms accumulate as f[i] += f_ij, f[j] -= f_ij in list
// order; virial by ev_tally_xyz with del = x_i - x_j; energy by
// ev_tally_full(i, 2 evdwl) (eatom[i] += evdwl).  firstneigh is given as CSR
// (offsets[i] .. offsets[i] + numneigh[i]) indexed by atom i.

extern "C" void orc_pair_compute_lmp(int n_all, const double *x,
                                     const int *type, int inum,
                                     const int *ilist, const int *numneigh,
                                     const int *offsets, const int *neigh,
                                     int S, int N, int L, int power,
                                     const double *cutoffs, int cutoff_kind,
                                     const double *beta_full, double *energy,
                                     double *forces, double *virial,
                                     double *eatom) {
  const int H = (L + 1) * (L + 1);
  const int nr = S * N;
  const int W = nr * H;
  const int F = nr * (nr + 1) / 2 * (L + 1);
  const double empty_thresh = 1e-8;
  std::vector<double> g(N), gx(N), gy(N), gz(N), Y(H), Yx(H), Yy(H), Yz(H);
  std::vector<double> A(W), B(F), Z, w(F), u(W), dA;
  double e = 0;
  for (int c = 0; c < 6; c++) virial[c] = 0;
  for (int ii = 0; ii < inum; ii++) {
    const int i = ilist[ii];
    const int itype = type[i];
    const int jnum = numneigh[i];
    const int *jlist = neigh + offsets[i];
    const double xtmp = x[3 * i], ytmp = x[3 * i + 1], ztmp = x[3 * i + 2];
    const int si = itype - 1;
    int n_inner = 0;
    for (int jj = 0; jj < jnum; jj++) {
      int j = jlist[jj];
      double cv = cutoffs[si * S + type[j] - 1];
      double dx = x[3 * j] - xtmp, dy = x[3 * j + 1] - ytmp, dz = x[3 * j + 2] - ztmp;
      double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < cv * cv) n_inner++;
    }
    std::fill(A.begin(), A.end(), 0.0);
    dA.assign((size_t)n_inner * 3 * W, 0.0);
    int q = 0;
    for (int jj = 0; jj < jnum; jj++) {
      int j = jlist[jj];
      double dx = x[3 * j] - xtmp, dy = x[3 * j + 1] - ytmp, dz = x[3 * j + 2] - ztmp;
      double rsq = dx * dx + dy * dy + dz * dz;
      double r = std::sqrt(rsq);
      int s = type[j] - 1;
      double rc = cutoffs[si * S + s];
      double cutforcesq = rc * rc;
      if (!(rsq < cutforcesq)) continue;
      orc_radial(dx, dy, dz, r, rc, N, 0.0, cutoff_kind, g.data(), gx.data(),
                 gy.data(), gz.data());
      orc_ylm(dx, dy, dz, L, Y.data(), Yx.data(), Yy.data(), Yz.data());
      int dc = s * N * H;
      for (int rn = 0; rn < N; rn++)
        for (int a = 0; a < H; a++) {
          A[dc] += g[rn] * Y[a];
          dA[(size_t)(q * 3 + 0) * W + dc] += gx[rn] * Y[a] + g[rn] * Yx[a];
          dA[(size_t)(q * 3 + 1) * W + dc] += gy[rn] * Y[a] + g[rn] * Yy[a];
          dA[(size_t)(q * 3 + 2) * W + dc] += gz[rn] * Y[a] + g[rn] * Yz[a];
          dc++;
        }
      q++;
    }
    std::fill(B.begin(), B.end(), 0.0);
    for (int n1 = nr - 1; n1 >= 0; n1--) {
      int n1c = (n1 * (2 * nr - n1 + 1)) / 2;
      for (int n2 = n1; n2 < nr; n2++) {
        int n2c = n2 - n1;
        for (int l = 0; l < L + 1; l++) {
          int counter = l + (n1c + n2c) * (L + 1);
          for (int m = 0; m < 2 * l + 1; m++)
            B[counter] += A[n1 * H + l * l + m] * A[n2 * H + l * l + m];
        }
      }
    }
    double norm2 = 0;
    for (int c = 0; c < F; c++) norm2 += B[c] * B[c];
    if (norm2 < empty_thresh) continue;
    const double *beta = beta_full + (size_t)si * (power == 1 ? F : (size_t)F * F);
    double evdwl;
    if (power == 1) {
      double bn = std::pow(norm2, 0.5);
      evdwl = dotn(B.data(), beta, F) / bn;
      for (int c = 0; c < F; c++) w[c] = beta[c] / bn - evdwl * B[c] / norm2;
    } else {
      Z.assign(F, 0.0);
      for (int j = 0; j < F; j++) {
        const double *col = beta + (size_t)j * F;
        double bj = B[j];
        for (int c = 0; c < F; c++) Z[c] += col[c] * bj;
      }
      evdwl = dotn(B.data(), Z.data(), F) / norm2;
      for (int c = 0; c < F; c++) w[c] = 2 * (Z[c] - evdwl * B[c]) / norm2;
    }
    std::fill(u.begin(), u.end(), 0.0);
    for (int n1 = nr - 1; n1 >= 0; n1--)
      for (int n2 = 0; n2 < nr; n2++) {
        int n1c, n2c;
        double factor;
        if (n1 == n2) {
          n1c = (n1 * (2 * nr - n1 + 1)) / 2;
          n2c = n2 - n1;
          factor = 1.0;
        } else if (n1 < n2) {
          n1c = (n1 * (2 * nr - n1 + 1)) / 2;
          n2c = n2 - n1;
          factor = 0.5;
        } else {
          n1c = (n2 * (2 * nr - n2 + 1)) / 2;
          n2c = n1 - n2;
          factor = 0.5;
        }
        for (int l = 0; l < L + 1; l++) {
          int counter = l + (n1c + n2c) * (L + 1);
          for (int m = 0; m < 2 * l + 1; m++)
            u[n1 * H + l * l + m] += w[counter] * A[n2 * H + l * l + m] * factor;
        }
      }
    for (int c = 0; c < W; c++) u[c] *= 2;
    q = 0;
    for (int jj = 0; jj < jnum; jj++) {
      int j = jlist[jj];
      double rc = cutoffs[si * S + type[j] - 1];
      double delx = xtmp - x[3 * j], dely = ytmp - x[3 * j + 1], delz = ztmp - x[3 * j + 2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (!(rsq < rc * rc)) continue;
      double fx = dotn(&dA[(size_t)(q * 3 + 0) * W], u.data(), W);
      double fy = dotn(&dA[(size_t)(q * 3 + 1) * W], u.data(), W);
      double fz = dotn(&dA[(size_t)(q * 3 + 2) * W], u.data(), W);
      forces[3 * i] += fx;
      forces[3 * i + 1] += fy;
      forces[3 * i + 2] += fz;
      forces[3 * j] -= fx;
      forces[3 * j + 1] -= fy;
      forces[3 * j + 2] -= fz;
      virial[0] += delx * fx;
      virial[1] += dely * fy;
      virial[2] += delz * fz;
      virial[3] += delx * fy;
      virial[4] += delx * fz;
      virial[5] += dely * fz;
      q++;
    }
    e += evdwl;  // ev_tally_full(i, 2 evdwl): eng_vdwl += 0.5 * 2 evdwl
    if (eatom) eatom[i] += evdwl;
  }
  *energy = e;
}

extern "C" void orc_pair_compute(const orc_nlist *nl, const int *species,
                                 int S, int N, int L, int power,
                                 const double *cutoffs, int cutoff_kind,
                                 const double *beta_packed, int nthreads,
                                 double *energy, double *forces,
                                 double *virial, double *eatom) {
  const int nr = S * N;
  const int F = nr * (nr + 1) / 2 * (L + 1);
  std::vector<double> full(power == 1 ? (size_t)S * F : (size_t)S * F * F);
  orc_unpack_beta(S, F, power, beta_packed, full.data());
  orc_pair_compute_full(nl, species, S, N, L, power, cutoffs, cutoff_kind,
                        full.data(), nthreads, energy, forces, virial, eatom);
}
