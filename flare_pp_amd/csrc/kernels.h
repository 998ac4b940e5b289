// kernels.h -- device buffers and the host-side launchers of every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.h"

namespace flare {

enum Prec { kPrecF64 = 0, kPrecF32 = 1 };

// Growable device allocation owned by a handle.
struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  void ensure(size_t bytes) {
    if (bytes <= cap && ptr) return;
    if (ptr) {
      FLARE_HIP_CHECK(hipFree(ptr));
      ptr = nullptr;
      cap = 0;
    }
    size_t b = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
      ptr = nullptr;
      throw Error{2, std::string("hipMalloc of ") + std::to_string(b) +
                         " bytes failed: " + hipGetErrorString(e)};
    }
    cap = b;
  }
  template <typename T>
  T *as() const {
    return reinterpret_cast<T *>(ptr);
  }
  ~DevBuf() {
    if (ptr) (void)hipFree(ptr);
  }
};

// Cell matrices of the wrap (structure.cpp:120-133) and the cell-list grid.
// Per-pair forces f_ik (P x 3): FP64, or FP32 in the FP32 and mixed forms,
// whose pair forces are computed in FP32 (the values are the same either
// way; FP32 halves their traffic).
struct PairForces {
  void *p = nullptr;
  bool f32 = false;
};

struct WrapHost {
  double ct[9], cdi[9], cd[9], cti[9], cell[9];
  int nb[3];
  int sweep;
  double cutoff, volume, single_sweep;
};

struct NListBuffers {
  DevBuf wrapped, bin_of, bin_count, bin_start, bin_fill, bin_atoms, bin_pos, count,
      cum, keys_tmp, rel_tmp, keys_pad, keys, rel, idx, nspec, center,
      rev;
  DevBuf row_of, atom_of_row, species_of_row, species_count, species_blk,
      scan_tmp, blk_sum;
  int P = 0;
  // counters left zeroed by the last build (bin and block counts, buffers)
  // and the summary block the next build uses
  bool zero_ready = false;
  int zr_bins = -1, zr_ncb = -1;
  const void *zr_ptr[4] = {nullptr, nullptr, nullptr, nullptr};
  int parity = 0;
  // the last build grouped atoms in fixed-capacity bins (else scanned), and
  // its summary block on the device
  bool fcb = false;
  int *counts_block = nullptr;
};

// neighbors.hip
// Ints of the host-side build summary: species counts (kMaxSpecies), a
// bad-species flag, the largest over-capacity neighbor count, the pair
// total, a fixed-capacity bin overflow flag.
constexpr int kNlistStats = kMaxSpecies + 4;
// The host-side summary block (pinned, coherent): the kNlistStats ints,
// then at kNlistSeq the sequence number of the build they belong to, stored
// last with system scope by the kernel that completes the summary.
constexpr int kNlistSeq = 15, kNlistHostInts = 16;
static_assert(kNlistStats <= kNlistSeq, "summary ints before the sequence slot");
// Neighbors per atom held in LDS by the fused enumerate + sort pass; rows
// with more go to the two-pass build.
constexpr int kNbrCap = 128;
// Enqueues the neighbor list (count, cum, idx, nspec, rel, center, rev,
// keys), wrapped positions and the species-grouped row map row_of[i],
// atom_of_row[row], species_of_row[row] without waiting; the summary is
// written by the device into the host block at stats_dev (the device address
// of a kNlistHostInts pinned block) with sequence number seq.
void nlist_build_async(hipStream_t st, NListBuffers &b, int n, const double *d_pos,
                       const int *d_species, const WrapHost &wh, int *stats_dev,
                       unsigned seq);
// Waits on the host until the summary of build seq is in stats_host; an
// error if stream st (where the build was enqueued) ends or fails first.
void nlist_wait_stats(hipStream_t st, const volatile int *stats_host, unsigned seq);
// After nlist_wait_stats: validates the species codes, regroups the atoms
// in the scanned bin layout and enumerates again when a fixed-capacity bin
// overflowed (a new summary with sequence number ++*seq), sets b.P,
// completes the list with the two-pass path when an atom had more neighbors
// than the fused path holds, copies the species counts to host_counts and
// returns the number of species present (max species + 1).
int nlist_resolve(hipStream_t st, NListBuffers &b, int n, const int *d_species,
                  const WrapHost &wh, const volatile int *stats_host, int *stats_dev,
                  unsigned *seq, int *host_counts);

// A caller's neighbor list over local + ghost atoms (LAMMPS: x, type,
// ilist, numneigh, firstneigh as a 2-D array; pair_flare.cpp:80-118,
// pair_flare_kokkos.cpp:633-662).  Neighbor jj of atom i is
// neighbors[i * stride_i + jj * stride_j] with LAMMPS' special-bond bits
// masked off; types are LAMMPS' 1-based codes.
struct ExtList {
  int n_all, inum;
  const int *ilist, *numneigh, *neighbors;
  int64_t stride_i, stride_j;
  const double *x;
  const int *type;
};
// Builds the CSR list of the inum centres (pairs with r^2 < rc(si,sj)^2 in
// list order, relative vectors x_j - x_i, neighbor index j in [0, n_all)),
// center = centre index, rev = -1, the species-grouped row map of the
// centres and their 0-based species (cspec, inum).  Waits for the pair
// count; host_counts receives the centres' species counts.  Returns P.
int64_t nlist_from_external(hipStream_t st, NListBuffers &b, const ExtList &e,
                            const double *cut_host, int S, int *cspec,
                            int *host_counts);
// Target-sorted view of the pairs for the force assembly: in_cum (n_all+1)
// and the pair indices sorted by neighbor (stable, ascending pair index).
void ext_incoming(hipStream_t st, NListBuffers &b, int n_all, int64_t P,
                  int *in_cum, int *in_pairs);
// Pairs grouped by a key in [0, n_all) (stable, ascending pair index):
// in_cum (n_all + 1) and in_pairs (P).  work, cnt_buf, scan_tmp: scratch.
void index_by_target(hipStream_t st, DevBuf &work, DevBuf &cnt_buf, DevBuf &scan_tmp,
                     const int *keys, int64_t P, int n_all, int *in_cum,
                     int *in_pairs);
// forces[a] -= sum of pf over pairs whose neighbor is a (a < n_all), then
// forces[ilist[c]] += own[c] (c < inum); eatom[ilist[c]] += E_row[row_of[c]]
// when eatom is non-null.
void ext_assemble(hipStream_t st, int n_all, int inum, const int *ilist,
                  const int *in_cum, const int *in_pairs, const PairForces &pf,
                  const double *own, const int *row_of, const double *E_row,
                  double *forces, double *eatom);

// descriptors.hip
// A (atom order), GEMM rows (species-grouped: B when out_mode 0, B/|B| when
// 1, plus the hi/lo split of the latter into rows of kp when xh/xl are
// non-null: BF16, or FP16 with one power-of-two scale per row when
// f16_inv_scale is non-null, which receives the rows' 2^-e) and |B|^2.
void launch_b2_forward(hipStream_t st, const B2Dev &p, const NListDev &nl,
                       const int *species, const int *row_of,
                       const double *ynorm, const float *ynorm_f,
                       const int *b2_idx, int prec, int out_mode, int kp,
                       void *xh, void *xl, void *A_out, void *B_out,
                       double *norm2_out, float *f16_inv_scale = nullptr);
// dA/dr of pairs [p_begin, p_end) (plain form)
void launch_pair_dA(hipStream_t st, const B2Dev &p, const NListDev &nl,
                    const int *species, const double *ynorm, int64_t p_begin,
                    int64_t p_end, double *out);
void launch_b2_dervs(hipStream_t st, const B2Dev &p, const NListDev &nl,
                     const int *species, const int *row_of,
                     const double *ynorm, const int *b2_idx, int64_t nq,
                     const int64_t *fpair, const double *A, const double *B,
                     double *dB, double *fdots, double *coords, int *nidx);

// forces.hip
// Source of w = dE_i/dB_i for the force kernel.  mode 0: precomputed w
// rows; 1: mapped power 2 from GEMM rows X and Z = beta X; 2: mapped power 1
// from X and beta1 (one F-vector per species); 3: as 1 with X rebuilt from
// BF16 hi + lo rows (xh, xl, row stride kp).  Modes 1-3 also write the
// per-row energies.  X is B, or B/|B| when `normalized`.
struct WInput {
  int mode, normalized;
  const void *X, *Z, *beta1, *w;
  const double *norm2;
  double *E_row;
  const void *xh = nullptr, *xl = nullptr;  // mode 3: X as BF16 hi + lo
  int kp = 0;
};
// Per-atom u = dE/dA and pair forces f_ik for atoms [a_begin, a_end).
// mode 0: plain (reads dA of pair k at dA[(k - dA_first_pair) * 3 NH]),
// mode 1: adjoint (recomputes the pair basis).
void launch_forces(hipStream_t st, int prec, int mode, const B2Dev &p,
                   const NListDev &nl, const int *species, const int *row_of,
                   const double *ynorm, const float *ynorm_f, const void *A,
                   const WInput &wi, const double *dA, int64_t dA_first_pair,
                   int a_begin, int a_end, const PairForces &pair_f);
// F_a = sum_k (f_ak - f_rev(ak)) into forces (stride 3; may be NULL; pairs
// without a reverse contribute only their own term), and the energy (sum of
// e_row[row_of[a]]), virial (xx yy zz xy xz yz of V[a][b] = -sum_k r_a f_b)
// and GP stress (xx xy xz yy yz zz, times inv_volume); any of the three may
// be NULL.  Per workgroup of 16 atoms a partial row in part (gather_parts(n)
// x 10), then one workgroup per column sums them in a fixed order, so the
// result is deterministic.
int gather_parts(int n);
void launch_gather(hipStream_t st, const NListDev &nl, const PairForces &pair_f,
                   double *forces, const int *row_of, const double *e_row,
                   double *part, double inv_volume, double *energy, double *virial6,
                   double *stress6);
// column sums of an (n x k) row-major matrix into out[k], deterministic
void launch_colsum(hipStream_t st, int n, int k, const double *x, double *out);

// gp.hip
// GP kernel epilogue on raw dot products K = X P^T (rows x M), per row:
//   t = K/(|B||P|), k = sig2 t^xi (stored in kval), c = alpha t^(xi-1),
//   e = sum alpha t^xi; empty rows/columns give zeros.
void launch_gp_epilogue(hipStream_t st, int rows, int M, int normalized,
                        const double *K, const double *norm2_rows,
                        const double *pnorm, const double *alpha, double sig2,
                        double power, double *kval, double *cmat,
                        double *e_row, int ld = 0);
// w = sig2 xi / |B| (G - e B / |B|) (Appendix-A chain rule); e of row r is
// the sum of e_part[r * nparts .. + nparts)
void launch_gp_w(hipStream_t st, int rows, int F, const double *G,
                 const double *e_part, int nparts, const double *B, const double *norm2,
                 double sig2, double power, double *w, double *E_row);
void launch_row_sqsum_sub(hipStream_t st, int rows, int M, const double *Q,
                          const double *kself, double *var);
// var[r] = kself[r] - sum of part[r * nparts .. + nparts)
void launch_var_from_parts(hipStream_t st, int rows, int nparts, const double *part,
                           const double *kself, double *var);
void launch_kself(hipStream_t st, int rows, int F, const double *B,
                  const double *norm2, double sig2, double power,
                  double *kself);
void launch_row_norms(hipStream_t st, int rows, int F, const double *X,
                      double *norm, double *xhat);
// panel Cholesky of an nb x nb diagonal block (column-major, leading dim ld)
void launch_chol_panel(hipStream_t st, int nb, double *A, int ld, int *info);
void launch_zero_upper(hipStream_t st, int M, double *A, int ld);
void launch_scale_rows(hipStream_t st, int rows, int F, const double *X,
                       const double *s, double *out);
void launch_pack_beta(hipStream_t st, int F, int power, const double *full,
                      double *packed);
void launch_unpack_beta(hipStream_t st, int F, int power, const double *packed,
                        double *full);
void launch_to_f32(hipStream_t st, int64_t n, const double *x, float *y);
// compute flare/std/atom tails (compute_flare_std_atom.cpp:172-212), one row
// per species-grouped descriptor row, written to stds[atom_of_row[r]]:
//  map:  v = (B.Z / |B|^2) / sig2, Z = beta B
//  kvec: T <- (T / |B|)^power in place (rows x M)
//  linv: v = 1 - sig2 sum_u Q[r][u]^2
// std = sqrt(v), -sqrt(|v|) when v < 0, 0 when |B|^2 < 1e-8.
void launch_std_map(hipStream_t st, int rows, int F, const double *B,
                    const double *Z, const double *norm2, double sig2,
                    const int *atom_of_row, double *stds);
void launch_std_kvec(hipStream_t st, int rows, int M, double *T,
                     const double *norm2, double power);
void launch_std_linv(hipStream_t st, int rows, int M, const double *Q,
                     const double *norm2, double sig2, const int *atom_of_row,
                     double *stds);
// gp_struc.hip: explicit kernel matrices of a structure (envs_struc,
// self_kernel_struc) for the SOR / DTC variances.
// Per-pair values from FD (in place) plus the energy and stress entries of K
// rows (one species; rows u < Ms; K row stride n_out).
void launch_envs_struc_epilogue(hipStream_t st, int Ms, int ns, int64_t Q,
                                const double *DOT, double *FD, const double *pnorm,
                                const double *rnorm2, const double *fdots,
                                const double *coords, const int *pair_row,
                                double sig2, double power, double vol_inv, double *K,
                                int64_t n_out, int n_atoms);
// K[u][1 + 3a + c] += own-pair sums - incoming-pair sums of VAL
void launch_envs_struc_gather(hipStream_t st, int Ms, int n_atoms, int64_t Q,
                              const double *VAL, const int *own0, const int *ownc,
                              const int *in_cum, const int *in_pairs, double *K,
                              int64_t n_out);
void launch_self_g(hipStream_t st, int ns, int F, int m, int nn, const int64_t *q0,
                   const int *cnt, const double *dB, const double *fdots,
                   const double *coords, double *G, double *g);
void launch_self_rows(hipStream_t st, int ns, const double *T, const double *U,
                      const double *V, const double *g, const double *rnorm2,
                      double sig2, double power, int energy, double *row_out);
// entries per atom are limited to this many touching pairs
constexpr int kSelfMaxPairs = 512;
void launch_self_force(hipStream_t st, int n_atoms, int F, const int *row_of,
                       const int64_t *q0, const int *cnt, const int *in_cum,
                       const int *in_pairs, const int *pair_rowg,
                       const int *row_species, const double *D, const double *rnorm2,
                       const double *dB, const double *fdots, double sig2,
                       double power, double *out, int *overflow);
void launch_col_sqsum(hipStream_t st, int rows, int64_t cols, const double *X,
                      double sign, int accumulate, double *out);
// out[i][i] += v for i < F (row-major F x F)
void launch_add_diag_rm(hipStream_t st, int F, double *out, double v);
// Low-precision GP forms: epilogue from t (FP32, rows x M) into kval (FP64,
// rows of ldk >= M, zero past M), c as FP32 rows of ldk (cf) or as FP16 hi/lo rows
// of Mp scaled per row (ch, cl, inverse scales c_is), e (FP64)
void launch_gp_epilogue_lp(hipStream_t st, int rows, int M, int Mp,
                           const float *K, const double *norm2_rows,
                           const double *pnorm, const double *alpha, double sig2,
                           double power, double *kval, int ldk, float *cf, void *ch,
                           void *cl, float *c_is, double *e_row);
// The same epilogue for every species in one launch: species q covers rows
// row0 .. row0 + rows of a packed row space (species without rows or
// sparse environments left out), four rows per workgroup, one wavefront per
// row with the row's values in registers (single pass) when max(M, ldk, Mp)
// <= 1,024; false (nothing launched) otherwise.
struct GpEpiLp {
  int row0, rows, M, Mp, ldk;
  const float *K, *K2;  // t = K (+ K2: the second half of a K split)
  const double *norm2_rows, *pnorm, *alpha;
  double *kval;
  float *cf;
  void *ch, *cl;
  float *c_is;
  double *e_row;
};
bool launch_gp_epilogue_lp_rows(hipStream_t st, int nsp, const GpEpiLp *sp, double sig2,
                                double power);
// w (FP32) = sig2 xi / |B| (G - e d), d the FP32 normalised rows; e as in
// launch_gp_w
void launch_gp_w_lp(hipStream_t st, int rows, int F, const float *G,
                    const double *e_part, int nparts, const float *d, const double *norm2,
                    double sig2, double power, float *w, double *E_row);
void launch_kself_lp(hipStream_t st, int rows, int F, const float *d,
                     const double *norm2, double sig2, double power,
                     double *kself);


}  // namespace flare
