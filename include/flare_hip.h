/*
 * flare_hip.h -- C ABI of the MI355X (gfx950) force-evaluation path of
 * flare++.
 *
 * Every entry point takes plain pointers and sizes and returns FLARE_OK or
 * an error code; flare_last_error() gives the message of the last failure on
 * the calling thread.  Arrays named d_* are device (HBM) pointers, all other
 * pointers are host memory.  Work is enqueued on the context's stream; the
 * call returns when its device results are ready unless stated otherwise.
 *
 * Row-major layouts throughout.  Reference citations are relative to the
 * flare_pp source tree.
 */
#ifndef FLARE_HIP_H
#define FLARE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------*/
#define FLARE_OK 0
#define FLARE_EINVAL 1  /* bad argument or unsupported setting */
#define FLARE_ENOMEM 2  /* device allocation failed */
#define FLARE_EHIP 3    /* HIP / rocBLAS runtime error */
#define FLARE_EIO 4     /* file cannot be opened or written */
#define FLARE_EFORMAT 5 /* malformed coefficient file */

/* ---- settings ------------------------------------------------------------*/
#define FLARE_MAX_SPECIES 8
#define FLARE_RADIAL_CHEBYSHEV 0 /* radial.cpp:138-169 */
#define FLARE_CUTOFF_QUADRATIC 0 /* cutoffs.cpp:44-56 */
#define FLARE_CUTOFF_COSINE 1    /* cutoffs.cpp:58-70 */

/* Arithmetic forms of a force evaluation.
 * PLAIN_FP64    FP64 end to end; per-pair dA/dr materialised as pair_flare
 *               does (lammps_descriptor.cpp:44, 97-99).  The drop-in.
 * ADJOINT_FP64  reverse-mode kernel: dE/dB -> u -> per-pair recomputation of
 *               the radial basis and Y_lm; no per-pair derivative tensor.
 * ADJOINT_FP32  adjoint, per-pair work and contractions in FP32 (Kahan
 *               single-bond sums), norms and reductions in FP64.
 * ADJOINT_MIXED adjoint; dense contractions on BF16 MFMA with FP32
 *               accumulation and hi+lo splitting; the rest as ADJOINT_FP32,
 *               with uncompensated single-bond sums. */
#define FLARE_FORM_PLAIN_FP64 0
#define FLARE_FORM_ADJOINT_FP64 1
#define FLARE_FORM_ADJOINT_FP32 2
#define FLARE_FORM_ADJOINT_MIXED 3

typedef struct flare_ctx flare_ctx;
typedef struct flare_nlist flare_nlist;
typedef struct flare_pair flare_pair;
typedef struct flare_sgp flare_sgp;

/* B2 descriptor settings: B2(radial_basis, cutoff_function, radial_hyps,
 * cutoff_hyps, {n_species, n_max, l_max}, cutoffs) (b2.h:32-46).
 * cutoffs: host, n_species x n_species, row = central species. */
typedef struct flare_b2_params {
  int n_species, n_max, l_max;
  int radial_basis;    /* FLARE_RADIAL_CHEBYSHEV */
  int cutoff_function; /* FLARE_CUTOFF_* */
  double radial_r1;    /* radial_hyps[0]; radial_hyps[1] is the pair cutoff */
  const double *cutoffs;
} flare_b2_params;

/* Read-only views of a neighbor list (valid until the next build). */
typedef struct flare_nlist_view {
  int n_atoms;
  int64_t n_pairs;
  int sweep;
  double volume, single_sweep_cutoff, cutoff;
  int species_count[FLARE_MAX_SPECIES];
  const int *d_neighbor_count;            /* n */
  const int *d_cumulative_neighbor_count; /* n+1 */
  const int *d_structure_indices;         /* P */
  const int *d_neighbor_species;          /* P */
  const double *d_relative_positions;     /* P x 4: r, dx, dy, dz */
  const double *d_wrapped_positions;      /* n x 3 */
  const int *d_reverse_pair;              /* P */
  const int *d_atom_indices;              /* n: atom of species-grouped row */
} flare_nlist_view;

typedef struct flare_pair_info {
  int power, n_species, n_max, l_max, n_descriptors;
  int64_t beta_size;
  int cutoff_function;
  int legacy; /* 1 if the file had no power line (power taken as 2) */
  double cutoff; /* largest pair cutoff */
  char radial_basis[32];
  char cutoff_name[32];
} flare_pair_info;

/* ---- library ------------------------------------------------------------*/
const char *flare_version(void);
const char *flare_last_error(void);

/* ---- context: device, stream and BLAS handle ------------------------------*/
int flare_ctx_create(int device, flare_ctx **out);
int flare_ctx_destroy(flare_ctx *ctx);
/* hipStream_t to enqueue on (NULL: the null stream).  A neighbor list built
 * on one stream and first used after the stream was changed is waited for
 * (host synchronisation of the build's stream) before its use. */
int flare_ctx_set_stream(flare_ctx *ctx, void *stream);
int flare_ctx_synchronize(flare_ctx *ctx);
/* blocking = 0: compute entry points return once their work is enqueued
 * (flare_ctx_synchronize waits); 1 (default): they return with results
 * ready.  No counterpart in the reference (its calls are synchronous). */
int flare_ctx_set_blocking(flare_ctx *ctx, int blocking);
/* Memory budget (bytes, default 8 GiB) of the per-pair derivative tensor
 * that the plain form materialises: larger lists are processed in ranges of
 * atoms.  Plays the role of pair_flare_kokkos' batch size
 * (pair_flare_kokkos.cpp:117-124); results do not depend on it. */
int flare_ctx_set_plain_chunk(flare_ctx *ctx, int64_t bytes);
/* Stage timing: with profiling enabled every stage (nlist, forward, gemm,
 * pair_dA, forces_*, gather_reduce; in the sparse GP gp_K, gp_epilogue,
 * gp_G, gp_w, gp_kself, gp_Q, gp_var) is bracketed by HIP events on
 * the context stream; enabling or disabling resets the totals.  Stage
 * `index` (0-based, in order of first appearance) reports its name, total
 * milliseconds and count.  No counterpart in the reference. */
int flare_ctx_profile(flare_ctx *ctx, int enable);
int flare_ctx_profile_get(flare_ctx *ctx, int index, char *name, int name_len,
                          double *total_ms, int64_t *count);

/* ---- Structure neighbor list ---------------------------------------------
 * Replaces Structure::Structure(cell, species, positions, cutoff, ...) and
 * Structure::compute_neighbors (structure.cpp:27-43, 52-118): wrapped
 * positions, neighbor counts, cumulative counts, structure indices,
 * neighbor species and relative positions, in the reference's order. */
int flare_nlist_create(flare_ctx *ctx, flare_nlist **out);
int flare_nlist_destroy(flare_nlist *nl);
/* cell: host 3x3, rows are lattice vectors. d_species in [0, 8).  With
 * blocking off the build is only enqueued; its summary (species counts,
 * pair total) is read on the list's first use.  d_positions is read by the
 * build's own kernels only (a regrouping of over-full bins works from the
 * list's wrapped positions); d_species must stay valid until that first use
 * (the regrouping and the rebuild of rows over 128 neighbors read it). */
int flare_nlist_build(flare_ctx *ctx, flare_nlist *nl, int n_atoms,
                      const double *cell, const double *d_positions,
                      const int *d_species, double cutoff);
int flare_nlist_get_view(const flare_nlist *nl, flare_nlist_view *out);

/* ---- B2 descriptor ---------------------------------------------------------
 * Replaces B2::compute_struc (b2.cpp:76-179).  Outputs follow the
 * DescriptorValues layout (descriptor.h:31-57) with the per-species blocks
 * concatenated (species 0 rows first; original atom order within a
 * species). */
int flare_b2_n_descriptors(const flare_b2_params *p);
/* descriptors n x F, norms n, single-bond values n x W (atom order, W =
 * S*N*(L+1)^2, may be NULL).  Any output may be NULL. */
int flare_b2_compute(flare_ctx *ctx, const flare_b2_params *p,
                     const flare_nlist *nl, const int *d_species,
                     double *d_descriptors, double *d_norms,
                     double *d_single_bond);
/* Number of pairs inside the per-species B2 cutoffs (r <= rc(si,sj),
 * b2.cpp:277) and per-species split; n_pairs_by_type may be NULL. */
int flare_b2_count_pairs(flare_ctx *ctx, const flare_b2_params *p,
                         const flare_nlist *nl, const int *d_species,
                         int64_t *n_pairs, int64_t *n_pairs_by_type);
/* Force derivatives of the DescriptorValues layout (b2.cpp:199-240):
 * force_dervs 3P' x F, force_dots 3P', neighbor_coordinates P' x 3,
 * neighbor_counts n, cumulative_neighbor_counts n (offset within the
 * species block), neighbor_indices P'.  Sized by flare_b2_count_pairs. */
int flare_b2_compute_dervs(flare_ctx *ctx, const flare_b2_params *p,
                           const flare_nlist *nl, const int *d_species,
                           double *d_force_dervs, double *d_force_dots,
                           double *d_neighbor_coordinates,
                           int *d_neighbor_counts,
                           int *d_cumulative_neighbor_counts,
                           int *d_neighbor_indices);

/* ---- mapped pair style: pair_style flare ---------------------------------*/
/* PairFLARE::read_file (pair_flare.cpp:263-402): current layout and the
 * legacy layout without the power line (one cutoff for every pair). */
int flare_pair_load(flare_ctx *ctx, const char *path, flare_pair **out);
/* The same parser on the host only (no device work): fills info and, when
 * non-NULL, cutoffs (S x S) and the packed values (S x beta_size). */
int flare_coeff_read(const char *path, flare_pair_info *info, double *cutoffs,
                     double *beta_packed);
/* From memory: beta_packed host, n_species x beta_size in the file order. */
int flare_pair_create(flare_ctx *ctx, int power, int n_species, int n_max,
                      int l_max, int cutoff_function, const double *cutoffs,
                      const double *beta_packed, flare_pair **out);
int flare_pair_destroy(flare_pair *pr);
int flare_pair_get_info(const flare_pair *pr, flare_pair_info *out);
/* host copies: cutoffs S x S; beta unpacked S x F x F (power 2) or S x F */
int flare_pair_get_cutoffs(const flare_pair *pr, double *cutoffs);
int flare_pair_get_beta(flare_ctx *ctx, const flare_pair *pr, double *beta);
/* PairFLARE::compute (pair_flare.cpp:68-184): energy (1), forces n x 3,
 * virial (6: xx yy zz xy xz yz, sum of (x_i - x_j) f_ij) and optionally
 * per-atom energies (n).  d_beta: optional device override of the unpacked
 * beta (S x F x F or S x F); NULL uses the loaded coefficients.  A power-2
 * override must be symmetric, as the file's beta is by construction: the
 * forces use w = 2 (beta B - E B) / |B|^2. */
int flare_pair_compute(flare_ctx *ctx, flare_pair *pr, const flare_nlist *nl,
                       const int *d_species, int form, const double *d_beta,
                       double *d_energy, double *d_forces, double *d_virial,
                       double *d_eatom);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while they run (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): flare_pair_compute with the same
 * arguments and results, whose workspace is carved out of a static device
 * array of FLARE_NOALLOC_SCRATCH_BYTES instead of the context's growable
 * buffers.  Each GPU the library runs on holds its own array until the
 * process ends.  The call makes no hipMalloc / hipFree and no rocBLAS call;
 * it returns FLARE_EINVAL, before any device work, for
 *  - FP64 / FP32 power-2 forms whose contraction would need rocBLAS
 *    (FP64 with F > 1024 or F odd; FP32 with F not a multiple of 4);
 *  - a pending neighbor list with atoms over the one-pass row capacity
 *    (complete it with flare_nlist_get_view first);
 * and, once the workspace outgrows the array, for inputs that are too large
 * (the mixed form at F = 680 needs about 7.5 KB per atom: 32,000 atoms fit;
 * the plain form's per-pair derivatives are processed in atom ranges that
 * fit).  Every call on a GPU shares that GPU's array and overwrites it:
 * calls that overlap on one GPU (other streams, threads or contexts) race.
 * Stage profiling is off during the call. */
#define FLARE_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int flare_pair_compute_noalloc(flare_ctx *ctx, flare_pair *pr,
                               const flare_nlist *nl, const int *d_species,
                               int form, const double *d_beta,
                               double *d_energy, double *d_forces,
                               double *d_virial, double *d_eatom);
/* PairFLARE::compute on the caller's own full neighbor list over local and
 * ghost atoms (pair_flare.cpp:68-184; the Kokkos contract of
 * pair_flare_kokkos.cpp:83-124, 633-662): what a LAMMPS pair_style
 * flare/hip passes every step.
 *   d_x       n_all x 3 positions, local atoms first, then ghosts (atom->x)
 *   d_type    n_all LAMMPS types, 1-based (atom->type)
 *   d_ilist   inum centre atoms (list->ilist)
 *   d_numneigh, d_neighbors: neighbor jj of atom i is
 *             d_neighbors[i * stride_i + jj * stride_j] for jj <
 *             d_numneigh[i] (LAMMPS' special-bond bits are masked off), i.e.
 *             Kokkos' d_neighbors(i, jj) in either layout; a CPU adaptor
 *             packs firstneigh into such an array.
 * Pairs with r^2 < rc(ti,tj)^2 are used in list order.  Outputs: d_forces
 * (n_all x 3) and d_eatom (n_all, may be NULL) are ACCUMULATED (+=) like
 * LAMMPS' f and eatom, ghost rows included (newton on: LAMMPS' reverse
 * communication sums them onto their owners); d_energy (1) and d_virial (6:
 * xx yy zz xy xz yz) are written with this call's totals. */
int flare_pair_compute_neigh(flare_ctx *ctx, flare_pair *pr, int n_all,
                             int inum, const int *d_ilist,
                             const int *d_numneigh, const int *d_neighbors,
                             int64_t stride_i, int64_t stride_j,
                             const double *d_x, const int *d_type, int form,
                             double *d_energy, double *d_forces,
                             double *d_virial, double *d_eatom);
/* dE/dbeta of the last flare_pair_compute: sum over atoms of species s of
 * d_i d_i^T, d = B/|B| (power 2, S x F x F) or d_i (power 1, S x F). */
int flare_pair_beta_grad(flare_ctx *ctx, flare_pair *pr, double *d_grad);
/* Coefficient file writer (sparse_gp.cpp:1142-1206 layout). */
int flare_write_coefficients(const char *path, const char *contributor,
                             int power, int n_species, int n_max, int l_max,
                             const char *radial_basis,
                             const char *cutoff_function,
                             const double *cutoffs, int64_t beta_size,
                             const double *beta_packed);

/* ---- sparse GP -------------------------------------------------------------
 * SparseGP with one NormalizedDotProduct kernel over one B2 descriptor
 * (sparse_gp.h:17-31, normalized_dot_product.h:11-73).  The sparse set is
 * species-grouped (ClusterDescriptor layout, descriptor.h:61-90):
 * d_sparse M x F, n_sparse_by_type host S.  On creation K_uu + jitter I,
 * its Cholesky factor and L^-1 are formed on the device in FP64
 * (sparse_gp.cpp:576-587). */
int flare_sgp_create(flare_ctx *ctx, int n_species, int n_descriptors,
                     const int *n_sparse_by_type, const double *d_sparse,
                     double sigma, double power, double jitter,
                     flare_sgp **out);
int flare_sgp_destroy(flare_sgp *gp);
int flare_sgp_set_alpha(flare_ctx *ctx, flare_sgp *gp, const double *d_alpha);
/* L^-1 (M x M, row-major lower; block diagonal by species) */
int flare_sgp_get_L_inverse(flare_ctx *ctx, const flare_sgp *gp,
                            double *d_out);
/* K_uu (M x M, without jitter) */
int flare_sgp_get_Kuu(flare_ctx *ctx, const flare_sgp *gp, double *d_out);
/* predict_mean (sparse_gp.cpp:612-628): mean_efs = [E, F (3n), S (6)] with
 * S = (xx, xy, xz, yy, yz, zz) as in envs_struc. */
int flare_sgp_predict_mean(flare_ctx *ctx, flare_sgp *gp,
                           const flare_b2_params *p, const flare_nlist *nl,
                           const int *d_species, int form,
                           double *d_mean_efs);
/* predict_local_uncertainties (sparse_gp.cpp:682-702): mean_efs and local
 * variances in species-grouped (cluster) order (n). */
int flare_sgp_predict_local_uncertainties(flare_ctx *ctx, flare_sgp *gp,
                                          const flare_b2_params *p,
                                          const flare_nlist *nl,
                                          const int *d_species, int form,
                                          double *d_mean_efs,
                                          double *d_variances);
/* envs_struc (normalized_dot_product.cpp:115-210): the kernel matrix of the
 * sparse set against every output [E, F (3n), S (6)] of a structure, d_K
 * M x (1+3n+6), row = sparse environment.  Materialises the structure's dB/dr
 * (3 P' x F) and an M_s x 3P'_s block per species. */
int flare_sgp_envs_struc(flare_ctx *ctx, flare_sgp *gp,
                         const flare_b2_params *p, const flare_nlist *nl,
                         const int *d_species, double *d_K);
/* self_kernel_struc (normalized_dot_product.cpp:446-578): k(o, o) of every
 * output o of a structure, d_out (1+3n+6). */
int flare_sgp_self_kernel_struc(flare_ctx *ctx, flare_sgp *gp,
                                const flare_b2_params *p, const flare_nlist *nl,
                                const int *d_species, double *d_out);
/* R^-1 of the training QR (sparse_gp.cpp:600-607), M x M row-major, device.
 * Training is not part of this library: the caller supplies it, as alpha. */
int flare_sgp_set_R_inv(flare_ctx *ctx, flare_sgp *gp, const double *d_R_inv);
/* predict_SOR (sparse_gp.cpp:630-650): mean_efs = K^T alpha, variance_efs =
 * diag(K^T R^-1 R^-T K).  predict_DTC (:652-680): variance_efs = self kernel
 * - diag(K^T K_uu^-1 K) + diag(K^T Sigma K), Sigma = R^-1 R^-T.  Both
 * (1+3n+6); they need alpha and R_inv. */
int flare_sgp_predict_SOR(flare_ctx *ctx, flare_sgp *gp,
                          const flare_b2_params *p, const flare_nlist *nl,
                          const int *d_species, double *d_mean_efs,
                          double *d_variance_efs);
int flare_sgp_predict_DTC(flare_ctx *ctx, flare_sgp *gp,
                          const flare_b2_params *p, const flare_nlist *nl,
                          const int *d_species, double *d_mean_efs,
                          double *d_variance_efs);
/* envs_envs (normalized_dot_product.cpp:27-87) of two descriptor sets
 * (species-grouped, n1 x F and n2 x F) into d_out n1 x n2. */
int flare_envs_envs(flare_ctx *ctx, int n_species, int n_descriptors,
                    const int *n1_by_type, const double *d1,
                    const int *n2_by_type, const double *d2, double sigma,
                    double power, double *d_out);
/* mapping coefficients (normalized_dot_product.cpp:643-759), packed file
 * order, S x beta_size */
int flare_sgp_mapping_coefficients(flare_ctx *ctx, const flare_sgp *gp,
                                   double *d_out);

/* ---- variance path: compute flare/std/atom ---------------------------------
 * ComputeFlareStdAtom (compute_flare_std_atom.cpp:31-664) on the GPU.  Two
 * model sources, as in ComputeFlareStdAtom::coeff (:297-314):
 *  - a variance-map file (write_varmap_coefficients, sparse_gp.cpp:
 *    1208-1281; reader :349-451): variance = B^T beta_s B / |B|^2 / sigma^2;
 *  - an L^-1 file plus a sparse-descriptor file (write_L_inverse /
 *    write_sparse_descriptors, sparse_gp.cpp:1283-1396; readers :453-650):
 *    variance = 1 - sigma^2 |L^-1 k|^2, k = (P_hat_s B/|B|)^power on the
 *    atom's species, zero elsewhere.
 * Per-atom result: sqrt(variance), or -sqrt(|variance|) when negative; 0 for
 * empty environments (|B|^2 < 1e-8).  FP64 throughout: both forms subtract
 * nearly equal terms. */
typedef struct flare_std flare_std;
typedef struct flare_std_info {
  int use_map;  /* 1: variance map; 0: L^-1 + sparse descriptors */
  int power;    /* kernel power (L^-1 file); 2 for the variance map */
  int n_hyps;
  double hyps[4]; /* sigma, energy noise, force noise, stress noise */
  int n_species, n_max, l_max, n_descriptors;
  int64_t beta_size; /* variance map: F*F */
  int n_kernels;     /* L^-1 and sparse-descriptor files (1 supported) */
  int n_clusters;
  int n_clusters_by_type[FLARE_MAX_SPECIES];
  int cutoff_function;
  double cutoff; /* largest pair cutoff */
  char radial_basis[32];
  char cutoff_name[32];
} flare_std_info;

/* Host-only parser of the files above.  path2 == NULL reads path1 as a
 * variance map, otherwise path1 is the L^-1 file and path2 the sparse
 * descriptors.  Fills info and, where non-NULL, cutoffs (S x S), values1
 * (variance map: S x F x F row-major as stored; L^-1: n_clusters x
 * n_clusters row-major, upper triangle zero) and values2 (sparse rows
 * n_clusters x F, species-grouped).  Call with NULL arrays to get sizes. */
int flare_std_read(const char *path1, const char *path2, flare_std_info *info,
                   double *cutoffs, double *values1, double *values2);
/* ComputeFlareStdAtom::coeff: load the model onto the device. */
int flare_std_load(flare_ctx *ctx, const char *path1, const char *path2,
                   flare_std **out);
/* The same from memory: info fields use_map, power, n_hyps, hyps,
 * n_species, n_max, l_max, cutoff_function, n_clusters and
 * n_clusters_by_type are read; arrays as in flare_std_read (host). */
int flare_std_create(flare_ctx *ctx, const flare_std_info *info,
                     const double *cutoffs, const double *values1,
                     const double *values2, flare_std **out);
int flare_std_destroy(flare_std *sd);
int flare_std_get_info(const flare_std *sd, flare_std_info *out);
/* ComputeFlareStdAtom::compute_peratom (:99-213): d_stds n (atom order). */
int flare_std_compute(flare_ctx *ctx, flare_std *sd, const flare_nlist *nl,
                      const int *d_species, double *d_stds);

/* Variance map of a power-1 sparse GP (normalized_dot_product.cpp:761-852):
 * d_out S x F x F, beta_s = sigma^2 I - sigma^4 P_hat_s^T K_uu^-1 P_hat_s
 * over non-empty sparse rows; the sigma^2 I term only when the species'
 * first sparse environment is non-empty (:829-842).  FLARE_EINVAL for other
 * powers (the reference prints a message and returns an empty matrix). */
int flare_sgp_varmap_coefficients(flare_ctx *ctx, const flare_sgp *gp,
                                  double *d_out);
/* Writers, host arrays (sparse_gp.cpp:1208-1396 layouts; the DATE line uses
 * the local time).  hyps: n_hyps values (sigma first). */
int flare_write_varmap_coefficients(const char *path, const char *contributor,
                                    int n_hyps, const double *hyps,
                                    int n_species, int n_max, int l_max,
                                    const char *radial_basis,
                                    const char *cutoff_function,
                                    const double *cutoffs, int64_t beta_size,
                                    const double *beta);
/* L_inv: n_clusters x n_clusters row-major; its lower triangle is written */
int flare_write_L_inverse(const char *path, const char *contributor,
                          double power, int n_hyps, const double *hyps,
                          int n_species, int n_max, int l_max,
                          const char *radial_basis,
                          const char *cutoff_function, const double *cutoffs,
                          int n_clusters, const double *L_inv);
/* descriptors n_clusters x F (species-grouped, raw), norms n_clusters;
 * rows are written divided by their norm, zero when the norm < 1e-8 */
int flare_write_sparse_descriptors(const char *path, const char *contributor,
                                   int n_species, int n_descriptors,
                                   const int *n_clusters_by_type,
                                   const double *descriptors,
                                   const double *norms);

/* ---- synthetic structures ------------------------------------------------
 * No counterpart in the reference: generator of the disordered benchmark
 * box.  Random sequential addition of n points in a periodic cube of edge
 * box with minimum distance d_min (host only, deterministic for a seed);
 * positions: host n x 3.  max_attempts <= 0: unlimited.  FLARE_EINVAL when
 * the box is thinner than three grid cells or the attempts run out. */
int flare_synth_rsa(int n, double box, double d_min, uint64_t seed,
                    int64_t max_attempts, double *positions);

#ifdef __cplusplus
}
#endif
#endif /* FLARE_HIP_H */
