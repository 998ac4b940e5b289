/*
 * oracle.h -- FP64 CPU restatement of flare++'s force-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  The product (libflare_hip.so) never links it and never falls back to
 * it.
 *
 * Every function restates the reference algorithm it cites (paths are
 * relative to the flare_pp source tree).  Arithmetic is plain IEEE FP64 with
 * floating-point contraction disabled at build time, i.e. the operation
 * order of the reference's default x86-64 Release build (no FMA).
 *
 * Layouts are row-major and flat:
 *   cell[9]            rows are the lattice vectors (structure.h:95-97)
 *   positions[n*3]
 *   relative[P*4]      (r, dx, dy, dz) per neighbor (structure.cpp:80-86)
 *   single bond[W]     W = S*N*(L+1)^2, index (s*N + n)*(L+1)^2 + l*l + l + m
 *   B2[F]              F = nr(nr+1)/2*(L+1), nr = S*N, index p(n1,n2)*(L+1)+l
 *
 * Parity status of each part is documented in the header comment of
 * oracle.cpp.
 */
#ifndef FLARE_ORACLE_H
#define FLARE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cutoff function codes (cutoffs.cpp:44-70) */
#define ORC_CUTOFF_QUADRATIC 0
#define ORC_CUTOFF_COSINE 1

/* ---- geometry: structure.cpp:7-25, 120-165 ---------------------------- */
double orc_single_sweep_cutoff(const double cell[9]);
double orc_volume(const double cell[9]);
/* Eigen PartialPivLU inverse of a 3x3 (dynamic-size Eigen path). */
void orc_inverse3(const double a[9], double out[9]);
void orc_wrap_positions(const double cell[9], int n, const double *pos,
                        double *wrapped);

/* ---- neighbor list: structure.cpp:52-118 ------------------------------- */
typedef struct orc_nlist orc_nlist;
/* brute force O(n^2 * (2 sweep+1)^3), exactly the reference loop */
orc_nlist *orc_nlist_brute(const double cell[9], int n, const double *pos,
                           const int *species, double cutoff);
/* same list, built with a periodic cell list (O(n)) */
orc_nlist *orc_nlist_cells(const double cell[9], int n, const double *pos,
                           const int *species, double cutoff, int nthreads);
/* brute-force neighbors of a subset of atoms only (for spot checks of the
 * cell-list build at large n); writes counts[n_sub] and returns a list whose
 * rows follow the order of `atoms`. */
orc_nlist *orc_nlist_brute_subset(const double cell[9], int n,
                                  const double *pos, const int *species,
                                  double cutoff, int n_sub, const int *atoms);
void orc_nlist_free(orc_nlist *h);
int64_t orc_nlist_n_pairs(const orc_nlist *h);
int orc_nlist_n_atoms(const orc_nlist *h);
int orc_nlist_sweep(const orc_nlist *h);
/* copies out: count[n], cum[n+1], idx[P], nspec[P], rel[P*4], shift[P*3],
 * wrapped[n*3]; any pointer may be NULL */
void orc_nlist_get(const orc_nlist *h, int *count, int *cum, int *idx,
                   int *nspec, double *rel, int *shift, double *wrapped);

/* ---- radial basis and harmonics: radial.cpp:138-169, 344-379;
 *      cutoffs.cpp:44-70; y_grad.cpp:667 (convention) ------------------ */
void orc_radial(double x, double y, double z, double r, double rcut, int N,
                double r1, int cutoff_kind, double *g, double *gx, double *gy,
                double *gz);
/* real orthonormal Y_lm and Cartesian gradients, (L+1)^2 values each */
void orc_ylm(double x, double y, double z, int L, double *Y, double *Yx,
             double *Yy, double *Yz);

/* ---- B2 descriptor: b2.cpp:76-399 -------------------------------------- */
typedef struct orc_b2 orc_b2;
/* cutoffs: S*S matrix, row = central species.  want_dervs: materialise
 * single-bond and B2 force derivatives exactly as b2.cpp does. */
orc_b2 *orc_b2_compute(const orc_nlist *nl, const int *species, int S, int N,
                       int L, double r1, int cutoff_kind,
                       const double *cutoffs, int want_dervs);
void orc_b2_free(orc_b2 *h);
int orc_b2_n_descriptors(const orc_b2 *h);
int64_t orc_b2_n_pairs(const orc_b2 *h); /* pairs inside the B2 cutoff */
/* atom-order outputs (before the per-species regroup) */
void orc_b2_get_atom_order(const orc_b2 *h, double *single_bond, double *b2,
                           double *norms);
/* DescriptorValues layout (b2.cpp:104-176): species blocks concatenated.
 * n_by_type[S], neigh_by_type[S], descriptors[n*F], norms[n],
 * force_dervs[3P'*F], force_dots[3P'], neighbor_coords[P'*3],
 * neighbor_counts[n], cum_neighbor_counts[n] (offset within species block),
 * atom_indices[n], neighbor_indices[P'] */
void orc_b2_get_grouped(const orc_b2 *h, int *n_by_type, int *neigh_by_type,
                        double *descriptors, double *norms,
                        double *force_dervs, double *force_dots,
                        double *neighbor_coords, int *neighbor_counts,
                        int *cum_neighbor_counts, int *atom_indices,
                        int *neighbor_indices);
/* per-pair single-bond derivatives dA/dr (3P' x W), atom order */
void orc_b2_get_single_bond_dervs(const orc_b2 *h, double *dervs);

/* ---- NormalizedDotProduct: normalized_dot_product.cpp:27-210 ----------- */
/* Sparse set: per species s, rows sparse[off_s .. off_s + m_by_type[s]) of
 * an (M x F) matrix with norms. envs_envs of two such sets. */
void orc_envs_envs(int S, int F, const int *n1_by_type, const double *d1,
                   const double *norms1, const int *n2_by_type,
                   const double *d2, const double *norms2, double sigma,
                   double power, double *out /* n1 x n2 */);
/* kernel matrix rows [E, F(3n), S(6)] for each sparse environment */
void orc_envs_struc(const orc_b2 *struc, const int *m_by_type,
                    const double *sparse, const double *sparse_norms,
                    double sigma, double power, double volume,
                    double *out /* M x (1+3n+6) */);

/* self_kernel_struc (normalized_dot_product.cpp:446-578): out[1+3n+6] */
void orc_self_kernel_struc(const orc_b2 *struc, double sigma, double power,
                           double volume, double *out);

/* ---- SparseGP: sparse_gp.cpp:93-133, 576-628, 682-702 ------------------ */
/* L = chol(K + jitter I), Linv = L^-1 (lower); returns 0 on success */
int orc_cholesky_inverse(int M, const double *K, double jitter, double *Linv);
/* mean_efs = envs_struc^T alpha */
void orc_predict_mean(const orc_b2 *struc, const int *m_by_type,
                      const double *sparse, const double *sparse_norms,
                      double sigma, double power, double volume,
                      const double *alpha, double *mean_efs);
/* local variances in cluster (species-grouped) order */
void orc_cluster_uncertainties(const orc_b2 *struc, const int *m_by_type,
                               const double *sparse,
                               const double *sparse_norms, double sigma,
                               double power, const double *Linv,
                               double *variances);
/* the same for the grouped rows rows[0..nrows) (ascending) only */
void orc_cluster_uncertainties_rows(const orc_b2 *struc, const int *m_by_type,
                                    const double *sparse,
                                    const double *sparse_norms, double sigma,
                                    double power, const double *Linv,
                                    int nrows, const int *rows,
                                    double *variances);

/* ---- mapping coefficients: normalized_dot_product.cpp:643-759 ---------- */
/* out: S x beta_size, beta_size = F (power 1) or F(F+1)/2 (power 2) */
int orc_mapping_coefficients(int S, int F, const int *m_by_type,
                             const double *sparse, const double *sparse_norms,
                             double sigma, double power, const double *alpha,
                             double *out);

/* ---- mapped pair style: pair_flare.cpp:68-184, 378-395;
 *      lammps_descriptor.cpp:7-107, 209-304 ------------------------------ */
/* beta_packed: S x beta_size as stored in the coefficient file.
 * Neighbors come from `nl` (Structure order); the pair predicate is
 * rsq < rc(si,sj)^2 as in pair_flare.  virial order: xx yy zz xy xz yz.
 * eatom may be NULL.  nthreads <= 0 uses the OpenMP default. */
void orc_pair_compute(const orc_nlist *nl, const int *species, int S, int N,
                      int L, int power, const double *cutoffs,
                      int cutoff_kind, const double *beta_packed,
                      int nthreads, double *energy, double *forces,
                      double *virial, double *eatom);
/* as above, but each beta is given unpacked (S x F x F, symmetric); skips
 * the unpacking so it can be timed alone */
void orc_pair_compute_full(const orc_nlist *nl, const int *species, int S,
                           int N, int L, int power, const double *cutoffs,
                           int cutoff_kind, const double *beta_full,
                           int nthreads, double *energy, double *forces,
                           double *virial, double *eatom);
/* ---- variance path: normalized_dot_product.cpp:761-852;
 *      compute_flare_std_atom.cpp:99-213 ----------------------------------- */
/* variance map of a power-1 kernel: out S x F x F (row-major).  Kinv is
 * K_uu^-1 = L^-T L^-1 (M x M, sparse_gp.cpp:588). */
int orc_varmap_coefficients(int S, int F, const int *m_by_type,
                            const double *sparse, const double *sparse_norms,
                            double sigma, const double *Kinv, double *out);
/* per-atom signed stds (atom order).  use_map != 0: beta_var S x F x F as
 * read from the variance-map file.  Otherwise: kernel power, n_clusters,
 * n_by_type[S], normalised sparse rows (n_clusters x F, species-grouped) and
 * L^-1 (n_clusters x n_clusters row-major, lower). */
void orc_std_atom(const orc_nlist *nl, const int *species, int S, int N,
                  int L, const double *cutoffs, int cutoff_kind, double sigma,
                  int use_map, const double *beta_var, int power,
                  int n_clusters, const int *n_by_type,
                  const double *sparse_normed, const double *Linv,
                  int nthreads, double *stds);

/* PairFLARE::compute on a LAMMPS full list with ghosts (serial): x n_all x 3,
 * LAMMPS types (1-based), centres ilist[inum], neighbors of atom i at
 * neigh[offsets[i] .. offsets[i] + numneigh[i]).  forces (n_all x 3) and
 * eatom (n_all, may be NULL) accumulate; energy and virial are written. */
void orc_pair_compute_lmp(int n_all, const double *x, const int *type,
                          int inum, const int *ilist, const int *numneigh,
                          const int *offsets, const int *neigh, int S, int N,
                          int L, int power, const double *cutoffs,
                          int cutoff_kind, const double *beta_full,
                          double *energy, double *forces, double *virial,
                          double *eatom);

/* unpack file beta (pair_flare.cpp:363-401): packed -> S x F x F */
void orc_unpack_beta(int S, int F, int power, const double *packed,
                     double *full);

#ifdef __cplusplus
}
#endif
#endif
