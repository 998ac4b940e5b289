// api.hip -- the C ABI of include/flare_hip.h: handles, host-side setup,
// coefficient-file I/O and the orchestration of the HIP kernels and rocBLAS
// contractions.
#include <rocblas/rocblas.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"
#include "flare_hip.h"
#include "kernels.h"
#include "mfma.h"
#include "util.h"

namespace flare {
thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
}  // namespace flare

using flare::DevBuf;
using flare::Error;

#define BLAS_CHECK(expr)                                                     \
  do {                                                                       \
    rocblas_status _s = (expr);                                              \
    if (_s != rocblas_status_success)                                        \
      throw ::flare::Error{FLARE_EHIP, std::string(#expr) + ": " +           \
                                           rocblas_status_to_string(_s)};    \
  } while (0)

#define REQUIRE(cond, msg) FLARE_REQUIRE(cond, msg)

template <class Fn>
static int guarded(Fn &&fn) {
  try {
    fn();
    return FLARE_OK;
  } catch (const Error &e) {
    flare::set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    flare::set_error("host allocation failed");
    return FLARE_ENOMEM;
  } catch (const std::exception &e) {
    flare::set_error(e.what());
    return FLARE_EINVAL;
  }
}

// ---------------------------------------------------------------------------
// Static scratch of flare_pair_compute_noalloc.  Each GPU holds its own copy
// of the array at its own address; host code uses only the address that
// hipGetSymbolAddress gives for the device (the array's name is a host-side
// stand-in).  The address is looked up once per device, with that device
// current, when the first context on it is created.

static __device__ __align__(256) unsigned char scratch[FLARE_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_ptr;       // device -> address of its array
std::map<int, uint64_t> g_scratch_gen;     // device -> last evaluation in it
std::atomic<uint64_t> g_eval_gen{0};       // mapped evaluations, numbered

void scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return;
  char *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw Error{FLARE_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                hipGetErrorString(e)};
  g_scratch_ptr[device] = ptr;
  g_scratch_gen[device] = 0;
}

char *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  return it == g_scratch_ptr.end() ? nullptr : it->second;
}
}  // namespace

// ---------------------------------------------------------------------------
// context

struct flare_ctx {
  int device = 0;
  int n_cu = 0;  // compute units of the device
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  int blocking = 1;
  int64_t plain_chunk_bytes = int64_t(8) << 30;  // plain-form dA/dr budget
  // stage timing with HIP events on the context stream (when enabled)
  int profiling = 0;
  struct Stage {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Stage> pending;
  std::vector<std::string> stage_names;
  std::map<std::string, std::pair<double, int64_t>> stage_ms;
  void mark_begin(const std::string &name) {
    if (!profiling) return;
    Stage st{name, nullptr, nullptr};
    FLARE_HIP_CHECK(hipEventCreate(&st.a));
    FLARE_HIP_CHECK(hipEventCreate(&st.b));
    FLARE_HIP_CHECK(hipEventRecord(st.a, stream));
    pending.push_back(st);
  }
  void mark_end() {
    if (!profiling || pending.empty()) return;
    FLARE_HIP_CHECK(hipEventRecord(pending.back().b, stream));
  }
  void collect() {
    if (pending.empty()) return;
    FLARE_HIP_CHECK(hipStreamSynchronize(stream));
    for (auto &st : pending) {
      float ms = 0;
      FLARE_HIP_CHECK(hipEventElapsedTime(&ms, st.a, st.b));
      auto it = stage_ms.find(st.name);
      if (it == stage_ms.end()) {
        stage_names.push_back(st.name);
        stage_ms[st.name] = {0.0, 0};
      }
      stage_ms[st.name].first += ms;
      stage_ms[st.name].second += 1;
      (void)hipEventDestroy(st.a);
      (void)hipEventDestroy(st.b);
    }
    pending.clear();
  }
  DevBuf ynorm_d, ynorm_f;
  std::map<std::string, std::unique_ptr<DevBuf>> ws;
  std::map<int64_t, std::unique_ptr<DevBuf>> b2idx;
  // Generation of the last mapped evaluation that wrote this context's
  // "pair.*" workspace (0: none); flare_pair_beta_grad reads those rows.
  uint64_t pair_ws_gen = 0;

  // While an arena is active (the no-allocation entry point), workspace
  // requests are carved out of the device's static scratch array instead of
  // the growable buffers: a key keeps its slot for the rest of the call.
  struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    std::map<std::string, std::pair<size_t, size_t>> slots;  // offset, bytes
  };
  Arena *arena = nullptr;

  void *buf(const std::string &key, size_t bytes) {
    if (arena) {
      auto it = arena->slots.find(key);
      if (it != arena->slots.end()) {
        REQUIRE(bytes <= it->second.second, "scratch slot " + key + " reused with a larger size");
        return arena->base + it->second.first;
      }
      const size_t off = (arena->used + 255) & ~size_t(255);
      REQUIRE(off <= arena->cap && bytes <= arena->cap - off,
              "input too large for the no-allocation entry point: its scratch holds " +
                  std::to_string(arena->cap) + " bytes and this evaluation needs more (at " +
                  key + ")");
      arena->slots[key] = {off, bytes};
      arena->used = off + bytes;
      return arena->base + off;
    }
    auto &p = ws[key];
    if (!p) p.reset(new DevBuf);
    p->ensure(bytes);
    return p->ptr;
  }
  DevBuf *ws_buf(const std::string &key) {
    auto &p = ws[key];
    if (!p) p.reset(new DevBuf);
    return p.get();
  }
  template <typename T>
  T *buf_t(const std::string &key, size_t count) {
    return reinterpret_cast<T *>(buf(key, std::max<size_t>(count, 1) * sizeof(T)));
  }
  void use() { FLARE_HIP_CHECK(hipSetDevice(device)); }
  void sync() { FLARE_HIP_CHECK(hipStreamSynchronize(stream)); }
  // end of an entry point: wait for results unless running non-blocking
  void finish() {
    collect();
    if (blocking) sync();
  }
  const int *b2_index(int S, int N, int L) {
    if (arena) {  // no upload into a new buffer: built in the scratch array
      const int F = flare::b2_n_descriptors(S, N, L);
      int *code = buf_t<int>("b2idx", F);
      flare::launch_b2_index(stream, S * N, L, code);
      return code;
    }
    int64_t key = ((int64_t)S << 32) | (N << 16) | L;
    auto &p = b2idx[key];
    if (!p) {
      int nr = S * N;
      std::vector<int> code;
      for (int c1 = 0; c1 < nr; c1++)
        for (int c2 = c1; c2 < nr; c2++)
          for (int l = 0; l <= L; l++) code.push_back((c1 << 20) | (c2 << 8) | l);
      p.reset(new DevBuf);
      p->ensure(code.size() * sizeof(int));
      FLARE_HIP_CHECK(hipMemcpy(p->ptr, code.data(), code.size() * sizeof(int),
                                hipMemcpyHostToDevice));
    }
    return p->as<int>();
  }
};

// The wide-row GEMMs pad N to 352-column tiles, the batched ones to `unit`
// columns (256 for BF16/FP16 x 3, 128 for FP32). The wide tiles are taken
// when they pad no more than `slack_pct` percent beyond what the batched
// kernel pads: measured at F = 680 (704 vs 768 columns, wide faster in every
// form) and F = 1,200 (1,408 vs 1,280: mixed still faster on wide tiles,
// 391 vs 436 us; FP32 slower, 1.10 vs 0.94 ms).
static bool wide_cols_fit(int F, int unit, int slack_pct) {
  if (F <= 256) return false;
  const int64_t wide = (F + 351) / 352 * 352, narrow = (F + unit - 1) / unit * unit;
  return 100 * wide <= (100 + slack_pct) * narrow;
}

static void gemm_rm(flare_ctx *c, bool ta, bool tb, int M, int N, int K,
                    double alpha, const double *A, int lda, const double *B,
                    int ldb, double beta, double *C, int ldc) {
  if (M == 0 || N == 0) return;
  REQUIRE(K > 0, "empty inner dimension");
  BLAS_CHECK(rocblas_dgemm(c->blas, tb ? rocblas_operation_transpose
                                       : rocblas_operation_none,
                           ta ? rocblas_operation_transpose
                              : rocblas_operation_none,
                           N, M, K, &alpha, B, ldb, A, lda, &beta, C, ldc));
}

static void sgemm_rm(flare_ctx *c, bool ta, bool tb, int M, int N, int K,
                     float alpha, const float *A, int lda, const float *B,
                     int ldb, float beta, float *C, int ldc) {
  if (M == 0 || N == 0) return;
  REQUIRE(K > 0, "empty inner dimension");
  BLAS_CHECK(rocblas_sgemm(c->blas, tb ? rocblas_operation_transpose
                                       : rocblas_operation_none,
                           ta ? rocblas_operation_transpose
                              : rocblas_operation_none,
                           N, M, K, &alpha, B, ldb, A, lda, &beta, C, ldc));
}

// ---------------------------------------------------------------------------
// cell matrices, restating Eigen's dynamic-size 3x3 arithmetic
// (structure.cpp:12-24, 120-133): products summed left to right without
// FMA; inverses through partial-pivot LU with reciprocal back substitution.

namespace {

void mm3(const double *a, const double *b, double *c) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      volatile double s = a[i * 3] * b[j];
      volatile double p1 = a[i * 3 + 1] * b[3 + j];
      s = s + p1;
      volatile double p2 = a[i * 3 + 2] * b[6 + j];
      s = s + p2;
      t[i * 3 + j] = s;
    }
  std::memcpy(c, t, sizeof(t));
}

int lu3(double lu[3][3], int tr[3]) {
  int nt = 0;
  for (int k = 0; k < 3; k++) {
    int p = k;
    double best = std::fabs(lu[k][k]);
    for (int i = k + 1; i < 3; i++)
      if (std::fabs(lu[i][k]) > best) {
        best = std::fabs(lu[i][k]);
        p = i;
      }
    tr[k] = p;
    if (best != 0) {
      if (p != k) {
        for (int j = 0; j < 3; j++) std::swap(lu[k][j], lu[p][j]);
        nt++;
      }
      for (int i = k + 1; i < 3; i++) lu[i][k] = lu[i][k] / lu[k][k];
    }
    for (int i = k + 1; i < 3; i++)
      for (int j = k + 1; j < 3; j++) {
        volatile double pr = lu[i][k] * lu[k][j];
        lu[i][j] = lu[i][j] - pr;
      }
  }
  return nt;
}

void inv3(const double *a, double *out) {
  double lu[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) lu[i][j] = a[i * 3 + j];
  int tr[3];
  lu3(lu, tr);
  double x[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int k = 0; k < 3; k++)
    if (tr[k] != k)
      for (int j = 0; j < 3; j++) std::swap(x[k][j], x[tr[k]][j]);
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      double b = x[i][j];
      for (int r = i + 1; r < 3; r++) {
        volatile double pr = b * lu[r][i];
        x[r][j] = x[r][j] - pr;
      }
    }
  for (int j = 0; j < 3; j++)
    for (int i = 2; i >= 0; i--) {
      double ainv = 1.0 / lu[i][i];
      double b = x[i][j] * ainv;
      x[i][j] = b;
      for (int r = 0; r < i; r++) {
        volatile double pr = b * lu[r][i];
        x[r][j] = x[r][j] - pr;
      }
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) out[i * 3 + j] = x[i][j];
}

double volume3(const double *cell) {
  double lu[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) lu[i][j] = cell[i * 3 + j];
  int tr[3];
  int nt = lu3(lu, tr);
  volatile double prod = lu[0][0] * lu[1][1];
  prod = prod * lu[2][2];
  return std::fabs((nt % 2 ? -1.0 : 1.0) * prod);
}

// structure.cpp:135-165
double single_sweep(const double *cell) {
  const double *v1 = cell, *v2 = cell + 3, *v3 = cell + 6;
  auto dot = [](const double *a, const double *b) {
    volatile double s = a[0] * b[0];
    volatile double p = a[1] * b[1];
    s = s + p;
    p = a[2] * b[2];
    s = s + p;
    return (double)s;
  };
  double ab = dot(v1, v2), ac = dot(v1, v3), bc = dot(v2, v3);
  double a = std::sqrt(dot(v1, v1)), b = std::sqrt(dot(v2, v2)),
         c = std::sqrt(dot(v3, v3));
  double m[6] = {a * std::sqrt(1 - std::pow(ab / (a * b), 2)),
                 b * std::sqrt(1 - std::pow(ab / (a * b), 2)),
                 a * std::sqrt(1 - std::pow(ac / (a * c), 2)),
                 c * std::sqrt(1 - std::pow(ac / (a * c), 2)),
                 b * std::sqrt(1 - std::pow(bc / (b * c), 2)),
                 c * std::sqrt(1 - std::pow(bc / (b * c), 2))};
  double s = m[0];
  for (int i = 0; i < 6; i++)
    if (m[i] < s) s = m[i];
  return s;
}

flare::WrapHost make_wrap(const double *cell, double cutoff) {
  flare::WrapHost w;
  std::memcpy(w.cell, cell, sizeof(w.cell));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) w.ct[j * 3 + i] = cell[i * 3 + j];
  inv3(w.ct, w.cti);
  mm3(cell, w.ct, w.cd);
  inv3(w.cd, w.cdi);
  w.volume = volume3(cell);
  w.single_sweep = single_sweep(cell);
  w.cutoff = cutoff;
  double sw = std::ceil(cutoff / w.single_sweep);
  REQUIRE(std::isfinite(sw) && sw >= 1 && sw < 64,
          "cutoff / cell combination needs too many periodic images");
  w.sweep = (int)sw;
  auto cross = [](const double *u, const double *v) {
    double c0 = u[1] * v[2] - u[2] * v[1];
    double c1 = u[2] * v[0] - u[0] * v[2];
    double c2 = u[0] * v[1] - u[1] * v[0];
    return std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  };
  const double *a0 = cell, *a1 = cell + 3, *a2 = cell + 6;
  double width[3] = {w.volume / cross(a1, a2), w.volume / cross(a2, a0),
                     w.volume / cross(a0, a1)};
  for (int k = 0; k < 3; k++) {
    double b = std::floor(width[k] / (cutoff * (1.0 + 1e-9)));
    if (!(b >= 1)) b = 1;
    if (b > 256) b = 256;
    w.nb[k] = (int)b;
  }
  return w;
}

}  // namespace

// ---------------------------------------------------------------------------
// neighbor list

struct flare_nlist {
  flare_ctx *ctx = nullptr;
  flare::NListBuffers b;
  flare::WrapHost wh;
  int n = 0;
  int species_count[flare::kMaxSpecies] = {0};
  int n_species_present = 0;
  bool built = false;
  // a build is enqueued without waiting; its host-side summary (pinned
  // coherent stats_host, stored by the device with sequence number
  // stats_seq) is read on the list's next use
  bool pending = false;
  int *stats_host = nullptr, *stats_dev = nullptr;
  unsigned stats_seq = 0;
  hipStream_t build_stream = nullptr;
  const int *d_species = nullptr;

  ~flare_nlist() {
    if (stats_host) (void)hipHostFree(stats_host);
  }
  void wait_stats() const {
    flare::nlist_wait_stats(build_stream, stats_host, stats_seq);
  }
  void resolve() {
    if (!pending) return;
    pending = false;
    built = false;
    ctx->use();
    wait_stats();
    n_species_present = flare::nlist_resolve(build_stream, b, n, d_species, wh, stats_host,
                                             stats_dev, &stats_seq, species_count);
    // the list's last kernels (the compaction) run after the summary is
    // published: work on another stream must not start before them
    if (build_stream != ctx->stream) FLARE_HIP_CHECK(hipStreamSynchronize(build_stream));
    built = true;
  }

  flare::NListDev dev() const {
    flare::NListDev d;
    d.n = n;
    d.P = b.P;
    d.cum = b.cum.as<int>();
    d.idx = b.idx.as<int>();
    d.nspec = b.nspec.as<int>();
    d.rel = b.rel.as<double>();
    d.rev = b.rev.as<int>();
    d.center = b.center.as<int>();
    return d;
  }
  void row_offsets(int S, int *off) const {
    off[0] = 0;
    for (int s = 0; s < S; s++) off[s + 1] = off[s] + species_count[s];
  }
};

// ---------------------------------------------------------------------------
// B2 parameters

static flare::B2Dev make_b2(const flare_b2_params *p, int pair_pred) {
  REQUIRE(p, "null B2 parameters");
  REQUIRE(p->n_species >= 1 && p->n_species <= flare::kMaxSpecies,
          "n_species must be in [1, 8]");
  REQUIRE(p->n_max >= 1 && p->n_max <= flare::kMaxN, "n_max must be in [1, 16]");
  REQUIRE(p->l_max >= 0 && p->l_max <= flare::kMaxL, "l_max must be in [0, 10]");
  REQUIRE(p->radial_basis == FLARE_RADIAL_CHEBYSHEV,
          "only the chebyshev radial basis is supported");
  REQUIRE(p->cutoff_function == FLARE_CUTOFF_QUADRATIC ||
              p->cutoff_function == FLARE_CUTOFF_COSINE,
          "cutoff function must be quadratic or cosine");
  REQUIRE(p->cutoffs, "null cutoff matrix");
  flare::B2Dev d;
  d.S = p->n_species;
  d.N = p->n_max;
  d.L = p->l_max;
  d.H = (d.L + 1) * (d.L + 1);
  d.NH = d.N * d.H;
  d.W = d.S * d.NH;
  d.nr = d.S * d.N;
  d.F = flare::b2_n_descriptors(d.S, d.N, d.L);
  d.cutoff_kind = p->cutoff_function;
  d.pair_pred = pair_pred;
  d.r1 = p->radial_r1;
  for (int q = 0; q < flare::kMaxSpecies * flare::kMaxSpecies; q++) d.cut[q] = 0;
  for (int q = 0; q < d.S * d.S; q++) {
    REQUIRE(p->cutoffs[q] > d.r1, "pair cutoffs must exceed radial_hyps[0]");
    d.cut[q] = p->cutoffs[q];
  }
  return d;
}

// Completes a pending build (host-side summary) before the list is used.
static void use_nlist(const flare_nlist *nl) {
  REQUIRE(nl && nl->built, "neighbor list not built");
  const_cast<flare_nlist *>(nl)->resolve();
}

static void check_species(const flare_nlist *nl, int S) {
  use_nlist(nl);
  for (int s = S; s < flare::kMaxSpecies; s++)
    REQUIRE(nl->species_count[s] == 0,
            "structure holds species codes beyond n_species");
}

// Shared forward: A (atom order), B (species-grouped rows), |B|^2.
struct Forward {
  void *A = nullptr, *B = nullptr;
  double *norm2 = nullptr;
  void *xh = nullptr, *xl = nullptr;
  int kp = 0;
};

// out_mode 0 writes B rows, 1 writes B/|B|; with bf16 true the latter is
// also split into BF16 hi/lo rows (K padded) for the matrix-core GEMM;
// out_mode 2 writes only that split.
static Forward run_forward(flare_ctx *c, const flare::B2Dev &p,
                           const flare_nlist *nl, const int *d_species,
                           int prec, const std::string &tag,
                           void *B_out = nullptr, int out_mode = 0,
                           bool bf16 = false, void *f16_hi = nullptr,
                           void *f16_lo = nullptr, float *f16_is = nullptr) {
  Forward f;
  const size_t ts = prec == flare::kPrecF64 ? sizeof(double) : sizeof(float);
  const int n = nl->n;
  f.A = c->buf(tag + ".A", std::max<size_t>(1, (size_t)n * p.W) * ts);
  if (out_mode != 2)
    f.B = B_out ? B_out : c->buf(tag + ".B", std::max<size_t>(1, (size_t)n * p.F) * ts);
  f.norm2 = c->buf_t<double>(tag + ".norm2", n);
  f.kp = flare::bf16_padded_k(p.F);
  if (bf16) {
    f.xh = c->buf(tag + ".xh", std::max<size_t>(1, (size_t)n * f.kp) * 2);
    f.xl = c->buf(tag + ".xl", std::max<size_t>(1, (size_t)n * f.kp) * 2);
  } else if (f16_is) {
    // the caller's FP16 row-scaled split (rows of kp)
    f.xh = f16_hi;
    f.xl = f16_lo;
  }
  c->mark_begin("forward");
  flare::launch_b2_forward(c->stream, p, nl->dev(), d_species,
                           nl->b.row_of.as<int>(), c->ynorm_d.as<double>(),
                           c->ynorm_f.as<float>(), c->b2_index(p.S, p.N, p.L),
                           prec, out_mode, f.kp, f.xh, f.xl, f.A, f.B, f.norm2, f16_is);
  c->mark_end();
  return f;
}

// Per-pair forces f_ik (P x 3) of every centre of nl, shared by the mapped
// and GP paths.
static flare::PairForces run_pair_forces(flare_ctx *c, const flare::B2Dev &p,
                                         const flare_nlist *nl, const int *d_species,
                                         int prec, int mode, const Forward &fw,
                                         const flare::WInput &wi, const std::string &tag) {
  const int n = nl->n;
  const int64_t P = nl->b.P;
  flare::PairForces pf;
  pf.f32 = prec != flare::kPrecF64;
  pf.p = c->buf(tag + (pf.f32 ? ".pf32" : ".pf"),
                (size_t)std::max<int64_t>(P, 1) * 3 * (pf.f32 ? sizeof(float) : sizeof(double)));
  if (mode == 0) {
    // plain form: dA/dr materialised for a range of atoms at a time, at most
    // c->plain_chunk_bytes of derivatives (one range when everything fits;
    // an atom with more pairs than the budget still gets a range of its own)
    const int64_t per_pair = 3 * (int64_t)p.NH * (int64_t)sizeof(double);
    int64_t budget = c->plain_chunk_bytes;
    if (c->arena) {  // what the scratch has left once the gather's partials are set aside
      const int64_t left = (int64_t)c->arena->cap - (int64_t)c->arena->used -
                           (int64_t)flare::gather_parts(n) * 10 * (int64_t)sizeof(double) - 1024;
      budget = std::min<int64_t>(budget, std::max<int64_t>(left, 0));
    }
    const int64_t max_pairs = std::max<int64_t>(1, budget / per_pair);
    std::vector<int> bounds{0, n};
    std::vector<int> cum_host;
    int64_t chunk_pairs = P;
    if (P > max_pairs) {
      cum_host.resize(n + 1);
      FLARE_HIP_CHECK(hipMemcpyAsync(cum_host.data(), nl->b.cum.ptr,
                                     (n + 1) * sizeof(int), hipMemcpyDeviceToHost,
                                     c->stream));
      FLARE_HIP_CHECK(hipStreamSynchronize(c->stream));
      bounds.assign(1, 0);
      chunk_pairs = 0;
      for (int a0 = 0; a0 < n;) {
        int a1 = a0 + 1;
        while (a1 < n && (int64_t)cum_host[a1 + 1] - cum_host[a0] <= max_pairs) a1++;
        chunk_pairs = std::max<int64_t>(chunk_pairs, cum_host[a1] - cum_host[a0]);
        bounds.push_back(a1);
        a0 = a1;
      }
    }
    double *d = c->buf_t<double>(tag + ".dA", (size_t)std::max<int64_t>(chunk_pairs, 1) * 3 * p.NH);
    for (size_t q = 0; q + 1 < bounds.size(); q++) {
      const int a0 = bounds[q], a1 = bounds[q + 1];
      const int64_t p0 = cum_host.empty() ? 0 : cum_host[a0];
      const int64_t p1 = cum_host.empty() ? P : cum_host[a1];
      c->mark_begin("pair_dA");
      flare::launch_pair_dA(c->stream, p, nl->dev(), d_species,
                            c->ynorm_d.as<double>(), p0, p1, d);
      c->mark_end();
      c->mark_begin("forces_plain");
      flare::launch_forces(c->stream, prec, mode, p, nl->dev(), d_species,
                           nl->b.row_of.as<int>(), c->ynorm_d.as<double>(),
                           c->ynorm_f.as<float>(), fw.A, wi, d, p0, a0, a1, pf);
      c->mark_end();
    }
  } else {
    c->mark_begin("forces_adjoint");
    flare::launch_forces(c->stream, prec, mode, p, nl->dev(), d_species,
                         nl->b.row_of.as<int>(), c->ynorm_d.as<double>(),
                         c->ynorm_f.as<float>(), fw.A, wi, nullptr, 0, 0, n, pf);
    c->mark_end();
  }
  return pf;
}

// Forces (gather over own and reverse pairs), per-centre virial rows, the
// energy and virial sums.  On a caller's list (no reverse pairs) `forces`
// receives each centre's own-pair sums.
static void run_gather_reduce(flare_ctx *c, const flare_nlist *nl, const flare::PairForces &pf,
                              const flare::WInput &wi, double *forces, double *energy,
                              double *virial6, double *stress6,
                              const std::string &tag) {
  const int n = nl->n;
  c->mark_begin("gather_reduce");
  const int nparts = flare::gather_parts(n);
  double *part = c->buf_t<double>(tag + ".part", (size_t)std::max(nparts, 1) * 10);
  flare::launch_gather(c->stream, nl->dev(), pf, forces, nl->b.row_of.as<int>(),
                       wi.E_row, part, 1.0 / nl->wh.volume, energy, virial6, stress6);
  c->mark_end();
}

// Forces / virial assembly shared by the mapped and GP paths.
static void run_forces(flare_ctx *c, const flare::B2Dev &p,
                       const flare_nlist *nl, const int *d_species, int prec,
                       int mode, const Forward &fw, const flare::WInput &wi,
                       double *forces, double *energy, double *virial6,
                       double *stress6, const std::string &tag) {
  const flare::PairForces pf = run_pair_forces(c, p, nl, d_species, prec, mode, fw, wi, tag);
  run_gather_reduce(c, nl, pf, wi, forces, energy, virial6, stress6, tag);
}

// ---------------------------------------------------------------------------
// coefficient files (pair_flare.cpp:263-416; sparse_gp.cpp:1142-1206)

struct flare_pair {
  flare_ctx *ctx = nullptr;
  int power = 2, S = 1, N = 1, L = 0, F = 0, cutoff_kind = 0, legacy = 0;
  int64_t beta_size = 0;
  double cutoff = 0;
  std::string radial = "chebyshev", cutoff_name = "quadratic";
  std::vector<double> cutoffs;
  DevBuf beta, beta_f32, beta_h, beta_l;  // FP64, FP32, BF16 hi/lo (K padded)
  // The last evaluation: its descriptor rows (B, or B/|B| in FP32, or the
  // BF16 hi + lo split) stay in the evaluating context's "pair.*" workspace,
  // or in the device's static scratch, until a later evaluation overwrites
  // them; last_gen numbers the evaluation so flare_pair_beta_grad can tell.
  bool have_last = false, last_mixed = false, last_scratch = false;
  int last_prec = 0, last_n = 0, last_device = -1;
  int last_off[flare::kMaxSpecies + 1] = {0};
  const flare_ctx *last_ctx = nullptr;
  uint64_t last_gen = 0;
  const void *last_B = nullptr, *last_xh = nullptr, *last_xl = nullptr;
  const double *last_norm2 = nullptr;
  std::unique_ptr<flare_nlist> ext;  // list built from a caller's neighbor list
};

namespace {

bool parse_int_line(const std::string &s, int &v) {
  const char *p = s.c_str();
  while (*p == ' ' || *p == '\t') p++;
  char *end = nullptr;
  errno = 0;
  long x = std::strtol(p, &end, 10);
  if (end == p || errno) return false;
  while (*end == ' ' || *end == '\t' || *end == '\r' || *end == '\n') end++;
  if (*end != 0) return false;
  v = (int)x;
  return true;
}

std::string first_word(const std::string &s) {
  std::istringstream is(s);
  std::string w;
  is >> w;
  return w;
}

void setup_pair(flare_pair *pr, const std::vector<double> &packed) {
  REQUIRE(pr->power == 1 || pr->power == 2, "Power should be 1 or 2.");
  REQUIRE(pr->S >= 1 && pr->S <= flare::kMaxSpecies, "n_species must be in [1, 8]");
  REQUIRE(pr->N >= 1 && pr->N <= flare::kMaxN, "n_max must be in [1, 16]");
  REQUIRE(pr->L >= 0 && pr->L <= flare::kMaxL, "l_max must be in [0, 10]");
  pr->F = flare::b2_n_descriptors(pr->S, pr->N, pr->L);
  int64_t check = pr->power == 1 ? (int64_t)pr->F : (int64_t)pr->F * (pr->F + 1) / 2;
  if (check != pr->beta_size)
    throw Error{FLARE_EFORMAT, "Beta size doesn't match the number of descriptors."};
  REQUIRE((int64_t)packed.size() == pr->S * pr->beta_size, "beta value count mismatch");
  pr->cutoff = -1;
  for (double v : pr->cutoffs) pr->cutoff = std::max(pr->cutoff, v);
  flare_ctx *c = pr->ctx;
  size_t per = pr->power == 1 ? (size_t)pr->F : (size_t)pr->F * pr->F;
  pr->beta.ensure(pr->S * per * sizeof(double));
  pr->beta_f32.ensure(pr->S * per * sizeof(float));
  double *tmp = c->buf_t<double>("pair.packed", packed.size());
  FLARE_HIP_CHECK(hipMemcpyAsync(tmp, packed.data(), packed.size() * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
  for (int s = 0; s < pr->S; s++)
    flare::launch_unpack_beta(c->stream, pr->F, pr->power, tmp + s * pr->beta_size,
                              pr->beta.as<double>() + s * per);
  flare::launch_to_f32(c->stream, (int64_t)(pr->S * per), pr->beta.as<double>(),
                       pr->beta_f32.as<float>());
  if (pr->power == 2) {
    const int kp = flare::bf16_padded_k(pr->F);
    pr->beta_h.ensure((size_t)pr->S * pr->F * kp * 2);
    pr->beta_l.ensure((size_t)pr->S * pr->F * kp * 2);
    flare::split_bf16(c->stream, pr->S * pr->F, pr->F, kp, pr->beta.as<double>(),
                      pr->beta_h.ptr, pr->beta_l.ptr);
  }
  c->sync();
}

}  // namespace

// ---------------------------------------------------------------------------
// sparse GP

struct flare_sgp {
  flare_ctx *ctx = nullptr;
  int S = 1, F = 0, M = 0;
  double sigma = 1, sig2 = 1, power = 2, jitter = 1e-8;
  std::vector<int> m_by_type, m_off;
  DevBuf P, Phat, pnorm, pnorm2, alpha, Kuu, Lcm, Linv_cm;
  bool have_alpha = false;
  DevBuf R_inv;  // M x M row-major, from training (sparse_gp.cpp:604-607)
  bool have_R_inv = false;
  // low-precision copies of the normalised sparse rows, built on first use:
  // FP32 rows; FP16 hi/lo rows of kp with per-row inverse scales; per
  // species the FP16 hi/lo transpose (F rows of mp[s], the species' count
  // padded to 32, at tpoff[s]) with per-feature inverse scales (F per species)
  bool have_lowp = false;
  int kp = 0;
  std::vector<int> mp;
  std::vector<int64_t> tpoff;
  DevBuf Phat_f, Phat_h, Phat_l, Phat_is, PhatT_h, PhatT_l, PhatT_is;
  // matrix-core operands, built on first use: per species P_hat^T (F rows
  // of ldk[s], at toff[s]; FP64 and FP32) and L^-1 rows (M_s rows of
  // ldk[s], at loff[s]), ldk[s] = M_s rounded up to 16
  bool have_pad = false;
  std::vector<int> ldk;
  std::vector<int64_t> toff, loff;
  DevBuf PhatT_d, PhatT_f, Linv_rm;
};

// ---------------------------------------------------------------------------
// compute flare/std/atom model (compute_flare_std_atom.cpp:297-650)

struct flare_std {
  flare_ctx *ctx = nullptr;
  int use_map = 1, power = 2, S = 1, N = 1, L = 0, F = 0, cutoff_kind = 0;
  int n_hyps = 0, n_kernels = 1, M = 0;
  double hyps[4] = {0, 0, 0, 0};
  double cutoff = 0;
  std::string radial = "chebyshev", cutoff_name = "quadratic";
  std::vector<double> cutoffs;
  std::vector<int> m_by_type, m_off;
  DevBuf beta;  // variance map: S x F x F, row-major as stored
  DevBuf Phat;  // L^-1 form: normalised sparse rows, M x F
  DevBuf Linv;  // L^-1 form: M x M row-major, lower
};

namespace {

// Line-oriented reader with the grab semantics of the reference readers
// (compute_flare_std_atom.cpp:652-664): header lines one at a time, value
// blocks line by line until the expected count is reached.
struct LineFile {
  FILE *f = nullptr;
  std::string path;
  explicit LineFile(const char *p) : path(p ? p : "") {
    f = p ? std::fopen(p, "r") : nullptr;
    if (!f) throw Error{FLARE_EIO, "Cannot open variance file " + path};
  }
  ~LineFile() {
    if (f) std::fclose(f);
  }
  LineFile(const LineFile &) = delete;
  LineFile &operator=(const LineFile &) = delete;
  bool next(std::string &line) {
    line.clear();
    char buf[4096];
    bool any = false;
    while (std::fgets(buf, sizeof buf, f)) {
      any = true;
      line += buf;
      if (line.back() == '\n') break;
    }
    return any;
  }
  std::string need(const char *what) {
    std::string l;
    if (!next(l)) throw Error{FLARE_EFORMAT, path + ": file ends before " + what};
    return l;
  }
  // a count line; blank lines before it are skipped
  std::string need_nonblank(const char *what) {
    for (;;) {
      std::string l = need(what);
      if (l.find_first_not_of(" \t\r\n\f") != std::string::npos) return l;
    }
  }
  void grab(int64_t n, double *out, const char *what) {
    int64_t i = 0;
    std::string line;
    while (i < n) {
      if (!next(line)) throw Error{FLARE_EFORMAT, path + ": file ends inside " + what};
      const char *q = line.c_str();
      for (;;) {
        while (*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r' || *q == '\f') q++;
        if (!*q) break;
        char *end = nullptr;
        double v = std::strtod(q, &end);
        if (end == q) throw Error{FLARE_EFORMAT, path + ": bad number in " + what};
        if (i >= n)
          throw Error{FLARE_EFORMAT, path + ": more values than expected in " + what};
        if (out) out[i] = v;
        i++;
        q = end;
      }
    }
  }
};

std::vector<long long> ints_of(const std::string &line, size_t want, const char *what) {
  std::istringstream is(line);
  std::vector<long long> v;
  long long x;
  while (v.size() < want && (is >> x)) v.push_back(x);
  if (v.size() < want) throw Error{FLARE_EFORMAT, std::string("malformed ") + what + " line"};
  return v;
}

// hyperparameter line: sscanf "%lg %lg %lg %lg" (compute_flare_std_atom.cpp:
// 367-374): sigma first, up to four values
void parse_hyps(const std::string &line, flare_std &sd) {
  std::istringstream is(line);
  double x;
  int k = 0;
  while (k < 4 && (is >> x)) sd.hyps[k++] = x;
  if (k == 0) throw Error{FLARE_EFORMAT, "malformed hyperparameter line"};
}

void set_basis(flare_std &sd) {
  if (sd.radial != "chebyshev")
    throw Error{FLARE_EINVAL, "unsupported radial basis '" + sd.radial + "'"};
  if (sd.cutoff_name == "quadratic")
    sd.cutoff_kind = FLARE_CUTOFF_QUADRATIC;
  else if (sd.cutoff_name == "cosine")
    sd.cutoff_kind = FLARE_CUTOFF_COSINE;
  else
    throw Error{FLARE_EINVAL, "unsupported cutoff function '" + sd.cutoff_name + "'"};
  REQUIRE(sd.S >= 1 && sd.S <= flare::kMaxSpecies, "n_species must be in [1, 8]");
  REQUIRE(sd.N >= 1 && sd.N <= flare::kMaxN, "n_max must be in [1, 16]");
  REQUIRE(sd.L >= 0 && sd.L <= flare::kMaxL, "l_max must be in [0, 10]");
  sd.F = flare::b2_n_descriptors(sd.S, sd.N, sd.L);
}

void read_cutoffs(LineFile &lf, flare_std &sd) {
  sd.cutoffs.assign((size_t)sd.S * sd.S, 0.0);
  lf.grab((int64_t)sd.S * sd.S, sd.cutoffs.data(), "cutoffs");
  sd.cutoff = -1;
  for (double v : sd.cutoffs) sd.cutoff = std::max(sd.cutoff, v);
}

// ComputeFlareStdAtom::read_file (compute_flare_std_atom.cpp:349-451)
void read_varmap(const char *path, flare_std &sd, std::vector<double> *beta) {
  LineFile lf(path);
  lf.need("header");
  sd.use_map = 1;
  sd.power = 2;
  sd.n_hyps = (int)ints_of(lf.need("hyperparameter count"), 1, "hyperparameter count")[0];
  parse_hyps(lf.need("hyperparameters"), sd);
  sd.radial = first_word(lf.need("radial basis"));
  auto v = ints_of(lf.need("descriptor settings"), 4, "'n_species n_max l_max beta_size'");
  sd.S = (int)v[0];
  sd.N = (int)v[1];
  sd.L = (int)v[2];
  int64_t beta_size = v[3];
  sd.cutoff_name = first_word(lf.need("cutoff function"));
  set_basis(sd);
  read_cutoffs(lf, sd);
  if ((int64_t)sd.F * sd.F != beta_size)
    throw Error{FLARE_EFORMAT, "Beta size doesn't match the number of descriptors."};
  if (beta) {
    beta->assign((size_t)sd.S * beta_size, 0.0);
    lf.grab(sd.S * beta_size, beta->data(), "variance map");
  }
}

// ComputeFlareStdAtom::read_L_inverse (compute_flare_std_atom.cpp:453-563)
void read_L_inverse(const char *path, flare_std &sd, std::vector<double> *Linv) {
  LineFile lf(path);
  lf.need("header");
  sd.use_map = 0;
  sd.power = (int)ints_of(lf.need("power"), 1, "power")[0];
  sd.n_hyps = (int)ints_of(lf.need("hyperparameter count"), 1, "hyperparameter count")[0];
  parse_hyps(lf.need("hyperparameters"), sd);
  sd.radial = first_word(lf.need("radial basis"));
  auto v = ints_of(lf.need("descriptor settings"), 4, "'n_species n_max l_max n_kernels'");
  sd.S = (int)v[0];
  sd.N = (int)v[1];
  sd.L = (int)v[2];
  sd.n_kernels = (int)v[3];
  sd.cutoff_name = first_word(lf.need("cutoff function"));
  set_basis(sd);
  if (sd.n_kernels != 1) throw Error{FLARE_EINVAL, "one kernel is supported"};
  read_cutoffs(lf, sd);
  sd.M = (int)ints_of(lf.need_nonblank("sparse count"), 1, "sparse count")[0];
  REQUIRE(sd.M >= 0, "negative sparse count");
  const int64_t M = sd.M;
  std::vector<double> packed((size_t)(M * (M + 1) / 2));
  lf.grab(M * (M + 1) / 2, packed.data(), "L inverse");
  if (Linv) {
    Linv->assign((size_t)(M * M), 0.0);
    int64_t c = 0;
    for (int64_t j = 0; j < M; j++)
      for (int64_t k = 0; k <= j; k++) (*Linv)[j * M + k] = packed[c++];
  }
}

// ComputeFlareStdAtom::read_sparse_descriptors (compute_flare_std_atom.cpp:
// 565-650)
void read_sparse(const char *path, flare_std &sd, std::vector<double> *rows) {
  LineFile lf(path);
  lf.need("header");
  int n_kern = (int)ints_of(lf.need("kernel count"), 1, "kernel count")[0];
  if (n_kern != sd.n_kernels)
    throw Error{FLARE_EFORMAT, "n_kernals in sparse_descriptors and L_inv not match"};
  auto v = ints_of(lf.need_nonblank("kernel line"), 3, "'kernel n_clusters n_types'");
  if ((int)v[1] != sd.M)
    throw Error{FLARE_EFORMAT, "n_clusters in sparse_descriptors and L_inv not match"};
  const int n_types = (int)v[2];
  if (n_types != sd.S)
    throw Error{FLARE_EFORMAT, "n_types in sparse_descriptors differs from n_species"};
  sd.m_by_type.assign(sd.S, 0);
  sd.m_off.assign(sd.S + 1, 0);
  if (rows) rows->assign((size_t)sd.M * sd.F, 0.0);
  for (int s = 0; s < n_types; s++) {
    int c = (int)ints_of(lf.need_nonblank("cluster count"), 1, "cluster count")[0];
    REQUIRE(c >= 0 && sd.m_off[s] + c <= sd.M, "cluster counts exceed n_clusters");
    sd.m_by_type[s] = c;
    sd.m_off[s + 1] = sd.m_off[s] + c;
    lf.grab((int64_t)c * sd.F, rows ? rows->data() + (size_t)sd.m_off[s] * sd.F : nullptr,
            "sparse descriptors");
  }
  if (sd.m_off[sd.S] != sd.M)
    throw Error{FLARE_EFORMAT, "cluster counts do not add up to n_clusters"};
}

void read_std_files(const char *path1, const char *path2, flare_std &sd,
                    std::vector<double> *v1, std::vector<double> *v2) {
  if (!path2) {
    read_varmap(path1, sd, v1);
  } else {
    read_L_inverse(path1, sd, v1);
    read_sparse(path2, sd, v2);
  }
}

void fill_std_info(const flare_std &sd, flare_std_info *o) {
  std::memset(o, 0, sizeof(*o));
  o->use_map = sd.use_map;
  o->power = sd.power;
  o->n_hyps = sd.n_hyps;
  for (int k = 0; k < 4; k++) o->hyps[k] = sd.hyps[k];
  o->n_species = sd.S;
  o->n_max = sd.N;
  o->l_max = sd.L;
  o->n_descriptors = sd.F;
  o->beta_size = sd.use_map ? (int64_t)sd.F * sd.F : 0;
  o->n_kernels = sd.n_kernels;
  o->n_clusters = sd.use_map ? 0 : sd.M;
  for (int s = 0; s < sd.S && !sd.use_map; s++) o->n_clusters_by_type[s] = sd.m_by_type[s];
  o->cutoff_function = sd.cutoff_kind;
  o->cutoff = sd.cutoff;
  std::snprintf(o->radial_basis, sizeof(o->radial_basis), "%s", sd.radial.c_str());
  std::snprintf(o->cutoff_name, sizeof(o->cutoff_name), "%s", sd.cutoff_name.c_str());
}

void upload_std(flare_std &sd, const double *v1, const double *v2) {
  flare_ctx *c = sd.ctx;
  if (sd.use_map) {
    const size_t nb = (size_t)sd.S * sd.F * sd.F * sizeof(double);
    sd.beta.ensure(nb);
    FLARE_HIP_CHECK(hipMemcpyAsync(sd.beta.ptr, v1, nb, hipMemcpyHostToDevice, c->stream));
  } else {
    const size_t M = sd.M;
    sd.Phat.ensure(std::max<size_t>(M, 1) * sd.F * sizeof(double));
    sd.Linv.ensure(std::max<size_t>(M * M, 1) * sizeof(double));
    if (M) {
      FLARE_HIP_CHECK(hipMemcpyAsync(sd.Phat.ptr, v2, M * sd.F * sizeof(double),
                                     hipMemcpyHostToDevice, c->stream));
      FLARE_HIP_CHECK(hipMemcpyAsync(sd.Linv.ptr, v1, M * M * sizeof(double),
                                     hipMemcpyHostToDevice, c->stream));
    }
  }
  c->sync();
}

// one line of values: "%.16e " each, five to a line with the counter c
// running on; the layout of write_L_inverse / write_sparse_descriptors
// (sparse_gp.cpp:1329-1336, 1378-1388)
void put_value(FILE *f, double v, int &count) {
  std::fprintf(f, "%.16e ", v);
  if (count % 5 == 0) std::fprintf(f, "\n");
  count++;
}

FILE *open_out(const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) throw Error{FLARE_EIO, std::string("cannot write ") + path};
  return f;
}

void close_out(FILE *f, const char *path) {
  if (std::fclose(f) != 0) throw Error{FLARE_EIO, std::string("cannot write ") + path};
}

std::string date_now() {
  time_t now = std::time(nullptr);
  std::string t(std::ctime(&now));
  if (!t.empty() && t.back() == '\n') t.pop_back();
  return t;
}

// hyperparameter count and values (sparse_gp.cpp:1230-1236, 1306-1312)
void put_hyps(FILE *f, int n_hyps, const double *hyps) {
  std::fprintf(f, "%d\n", n_hyps);
  for (int i = 0; i < n_hyps; i++) std::fprintf(f, "%.16e ", hyps[i]);
  std::fprintf(f, "\n");
}

// B2::write_to_file (b2.cpp:52-74)
void put_b2_settings(FILE *f, const char *radial, int S, int N, int L,
                     long long coeff_size, const char *cutoff_fn,
                     const double *cutoffs) {
  std::fprintf(f, "%s\n", radial ? radial : "chebyshev");
  std::fprintf(f, "%d %d %d %lld\n", S, N, L, coeff_size);
  std::fprintf(f, "%s\n", cutoff_fn ? cutoff_fn : "quadratic");
  for (int q = 0; q < S * S; q++) std::fprintf(f, "%.2f ", cutoffs[q]);
  std::fprintf(f, "\n");
}

}  // namespace

// ===========================================================================
extern "C" {

const char *flare_version(void) { return "flare_hip 0.1.0 (gfx950)"; }

const char *flare_last_error(void) { return flare::g_err.c_str(); }

int flare_ctx_create(int device, flare_ctx **out) {
  return guarded([&] {
    REQUIRE(out, "null output handle");
    *out = nullptr;
    std::unique_ptr<flare_ctx> c(new flare_ctx);
    c->device = device;
    c->use();
    FLARE_HIP_CHECK(
        hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
    scratch_setup(device);
    BLAS_CHECK(rocblas_create_handle(&c->blas));
    BLAS_CHECK(rocblas_set_stream(c->blas, c->stream));
    double tab[flare::kNormTableSize];
    float tabf[flare::kNormTableSize];
    flare::ylm_norm_table(tab);
    for (int q = 0; q < flare::kNormTableSize; q++) tabf[q] = (float)tab[q];
    c->ynorm_d.ensure(sizeof(tab));
    c->ynorm_f.ensure(sizeof(tabf));
    FLARE_HIP_CHECK(hipMemcpy(c->ynorm_d.ptr, tab, sizeof(tab), hipMemcpyHostToDevice));
    FLARE_HIP_CHECK(hipMemcpy(c->ynorm_f.ptr, tabf, sizeof(tabf), hipMemcpyHostToDevice));
    *out = c.release();
  });
}

int flare_ctx_destroy(flare_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    ctx->use();
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->blas) rocblas_destroy_handle(ctx->blas);
    delete ctx;
  });
}

int flare_ctx_set_stream(flare_ctx *ctx, void *stream) {
  return guarded([&] {
    REQUIRE(ctx, "null context");
    ctx->stream = (hipStream_t)stream;
    BLAS_CHECK(rocblas_set_stream(ctx->blas, ctx->stream));
  });
}

int flare_ctx_set_blocking(flare_ctx *ctx, int blocking) {
  return guarded([&] {
    REQUIRE(ctx, "null context");
    ctx->blocking = blocking ? 1 : 0;
  });
}

int flare_ctx_set_plain_chunk(flare_ctx *ctx, int64_t bytes) {
  return guarded([&] {
    REQUIRE(ctx, "null context");
    REQUIRE(bytes > 0, "chunk size must be positive");
    ctx->plain_chunk_bytes = bytes;
  });
}

int flare_ctx_profile(flare_ctx *ctx, int enable) {
  return guarded([&] {
    REQUIRE(ctx, "null context");
    ctx->use();
    ctx->collect();
    ctx->profiling = enable ? 1 : 0;
    ctx->stage_ms.clear();
    ctx->stage_names.clear();
  });
}

int flare_ctx_profile_get(flare_ctx *ctx, int index, char *name, int name_len,
                          double *total_ms, int64_t *count) {
  return guarded([&] {
    REQUIRE(ctx && name && total_ms && count && name_len > 0, "null argument");
    ctx->use();
    ctx->collect();
    REQUIRE(index >= 0 && index < (int)ctx->stage_names.size(), "no such stage");
    const std::string &k = ctx->stage_names[index];
    std::snprintf(name, name_len, "%s", k.c_str());
    *total_ms = ctx->stage_ms[k].first;
    *count = ctx->stage_ms[k].second;
  });
}

int flare_ctx_synchronize(flare_ctx *ctx) {
  return guarded([&] {
    REQUIRE(ctx, "null context");
    ctx->use();
    ctx->sync();
  });
}

// ---- neighbor list ---------------------------------------------------------
int flare_nlist_create(flare_ctx *ctx, flare_nlist **out) {
  return guarded([&] {
    REQUIRE(ctx && out, "null argument");
    flare_nlist *nl = new flare_nlist;
    nl->ctx = ctx;
    *out = nl;
  });
}

int flare_nlist_destroy(flare_nlist *nl) {
  return guarded([&] {
    if (!nl) return;
    nl->ctx->use();
    (void)hipStreamSynchronize(nl->ctx->stream);
    // a build enqueued on another stream may still store its summary into
    // the pinned block the destructor frees
    if (nl->build_stream != nl->ctx->stream) (void)hipStreamSynchronize(nl->build_stream);
    delete nl;
  });
}

int flare_nlist_build(flare_ctx *ctx, flare_nlist *nl, int n_atoms,
                      const double *cell, const double *d_positions,
                      const int *d_species, double cutoff) {
  return guarded([&] {
    REQUIRE(ctx && nl && cell, "null argument");
    REQUIRE(n_atoms >= 0, "negative atom count");
    REQUIRE(n_atoms == 0 || (d_positions && d_species), "null positions/species");
    REQUIRE(cutoff > 0 && std::isfinite(cutoff), "cutoff must be positive");
    for (int q = 0; q < 9; q++) REQUIRE(std::isfinite(cell[q]), "non-finite cell");
    ctx->use();
    nl->resolve();  // an earlier build's fallback must not run on new data
    nl->built = false;
    nl->wh = make_wrap(cell, cutoff);
    REQUIRE(nl->wh.volume > 0, "singular cell");
    nl->n = n_atoms;
    nl->ctx = ctx;
    nl->d_species = d_species;
    if (!nl->stats_host) {
      FLARE_HIP_CHECK(hipHostMalloc((void **)&nl->stats_host,
                                    flare::kNlistHostInts * sizeof(int),
                                    hipHostMallocMapped | hipHostMallocCoherent));
      std::memset(nl->stats_host, 0, flare::kNlistHostInts * sizeof(int));
      FLARE_HIP_CHECK(hipHostGetDevicePointer((void **)&nl->stats_dev, nl->stats_host, 0));
    }
    nl->stats_seq++;
    if (nl->stats_seq == 0) nl->stats_seq = 1;  // 0 is the block's initial value
    nl->build_stream = ctx->stream;
    ctx->mark_begin("nlist");
    flare::nlist_build_async(ctx->stream, nl->b, n_atoms, d_positions, d_species,
                             nl->wh, nl->stats_dev, nl->stats_seq);
    ctx->mark_end();
    nl->pending = true;
    nl->built = true;
    if (ctx->blocking) {
      nl->resolve();
      ctx->sync();
    }
  });
}

int flare_nlist_get_view(const flare_nlist *nl, flare_nlist_view *v) {
  return guarded([&] {
    REQUIRE(nl && v, "null argument");
    use_nlist(nl);
    v->n_atoms = nl->n;
    v->n_pairs = nl->b.P;
    v->sweep = nl->wh.sweep;
    v->volume = nl->wh.volume;
    v->single_sweep_cutoff = nl->wh.single_sweep;
    v->cutoff = nl->wh.cutoff;
    for (int s = 0; s < flare::kMaxSpecies; s++) v->species_count[s] = nl->species_count[s];
    v->d_neighbor_count = nl->b.count.as<int>();
    v->d_cumulative_neighbor_count = nl->b.cum.as<int>();
    v->d_structure_indices = nl->b.idx.as<int>();
    v->d_neighbor_species = nl->b.nspec.as<int>();
    v->d_relative_positions = nl->b.rel.as<double>();
    v->d_wrapped_positions = nl->b.wrapped.as<double>();
    v->d_reverse_pair = nl->b.rev.as<int>();
    v->d_atom_indices = nl->b.atom_of_row.as<int>();
  });
}

// ---- B2 -------------------------------------------------------------------
int flare_b2_n_descriptors(const flare_b2_params *p) {
  if (!p) return -1;
  return flare::b2_n_descriptors(p->n_species, p->n_max, p->l_max);
}

int flare_b2_compute(flare_ctx *ctx, const flare_b2_params *params,
                     const flare_nlist *nl, const int *d_species,
                     double *d_descriptors, double *d_norms,
                     double *d_single_bond) {
  return guarded([&] {
    REQUIRE(ctx, "null context");
    flare::B2Dev p = make_b2(params, 0);
    check_species(nl, p.S);
    ctx->use();
    Forward fw = run_forward(ctx, p, nl, d_species, flare::kPrecF64, "b2",
                             d_descriptors);
    if (d_norms) flare::launch_sqrt(ctx->stream, nl->n, fw.norm2, d_norms);
    if (d_single_bond && nl->n)
      FLARE_HIP_CHECK(hipMemcpyAsync(d_single_bond, fw.A,
                                     (size_t)nl->n * p.W * sizeof(double),
                                     hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sync();
  });
}

// host bookkeeping of the species-grouped B2 pair list
struct B2Pairs {
  std::vector<int> count, atom_of_row;
  std::vector<int64_t> q0;  // per row
  int64_t total = 0;
  int64_t by_type[flare::kMaxSpecies] = {0};
};

static B2Pairs b2_pairs(flare_ctx *c, const flare::B2Dev &p,
                        const flare_nlist *nl, const int *d_species) {
  B2Pairs r;
  const int n = nl->n;
  int *cnt = c->buf_t<int>("b2.count", n);
  flare::launch_b2_filter_count(c->stream, p, nl->dev(), d_species, cnt);
  r.count.resize(n);
  r.atom_of_row.resize(n);
  if (n) {
    FLARE_HIP_CHECK(hipMemcpyAsync(r.count.data(), cnt, n * sizeof(int),
                                   hipMemcpyDeviceToHost, c->stream));
    FLARE_HIP_CHECK(hipMemcpyAsync(r.atom_of_row.data(), nl->b.atom_of_row.ptr,
                                   n * sizeof(int), hipMemcpyDeviceToHost,
                                   c->stream));
  }
  c->sync();
  r.q0.resize(n);
  int off[flare::kMaxSpecies + 1];
  nl->row_offsets(p.S, off);
  int64_t q = 0;
  for (int s = 0; s < p.S; s++) {
    for (int row = off[s]; row < off[s + 1]; row++) {
      r.q0[row] = q;
      q += r.count[r.atom_of_row[row]];
      r.by_type[s] += r.count[r.atom_of_row[row]];
    }
  }
  r.total = q;
  return r;
}

int flare_b2_count_pairs(flare_ctx *ctx, const flare_b2_params *params,
                         const flare_nlist *nl, const int *d_species,
                         int64_t *n_pairs, int64_t *n_pairs_by_type) {
  return guarded([&] {
    REQUIRE(ctx && n_pairs, "null argument");
    flare::B2Dev p = make_b2(params, 0);
    check_species(nl, p.S);
    ctx->use();
    B2Pairs r = b2_pairs(ctx, p, nl, d_species);
    *n_pairs = r.total;
    if (n_pairs_by_type)
      for (int s = 0; s < p.S; s++) n_pairs_by_type[s] = r.by_type[s];
  });
}

int flare_b2_compute_dervs(flare_ctx *ctx, const flare_b2_params *params,
                           const flare_nlist *nl, const int *d_species,
                           double *d_force_dervs, double *d_force_dots,
                           double *d_neighbor_coordinates,
                           int *d_neighbor_counts,
                           int *d_cumulative_neighbor_counts,
                           int *d_neighbor_indices) {
  return guarded([&] {
    REQUIRE(ctx, "null context");
    flare::B2Dev p = make_b2(params, 0);
    check_species(nl, p.S);
    ctx->use();
    const int n = nl->n;
    Forward fw = run_forward(ctx, p, nl, d_species, flare::kPrecF64, "b2d");
    B2Pairs r = b2_pairs(ctx, p, nl, d_species);
    std::vector<int> ncount(n), ncum(n);
    int off[flare::kMaxSpecies + 1];
    nl->row_offsets(p.S, off);
    for (int s = 0; s < p.S; s++) {
      int64_t start = off[s] < off[s + 1] ? r.q0[off[s]] : 0;
      for (int row = off[s]; row < off[s + 1]; row++) {
        ncount[row] = r.count[r.atom_of_row[row]];
        ncum[row] = (int)(r.q0[row] - start);
      }
    }
    if (n && d_neighbor_counts)
      FLARE_HIP_CHECK(hipMemcpyAsync(d_neighbor_counts, ncount.data(), n * sizeof(int),
                                     hipMemcpyHostToDevice, ctx->stream));
    if (n && d_cumulative_neighbor_counts)
      FLARE_HIP_CHECK(hipMemcpyAsync(d_cumulative_neighbor_counts, ncum.data(),
                                     n * sizeof(int), hipMemcpyHostToDevice,
                                     ctx->stream));
    const int64_t Q = r.total;
    if (Q > 0) {
      int64_t *q0 = ctx->buf_t<int64_t>("b2d.q0", n);
      int64_t *fpair = ctx->buf_t<int64_t>("b2d.fpair", Q);
      FLARE_HIP_CHECK(hipMemcpyAsync(q0, r.q0.data(), n * sizeof(int64_t),
                                     hipMemcpyHostToDevice, ctx->stream));
      flare::launch_b2_fill_fpair(ctx->stream, p, nl->dev(), d_species,
                                  nl->b.row_of.as<int>(), q0, fpair);
      double *dB = d_force_dervs ? d_force_dervs
                                 : ctx->buf_t<double>("b2d.dB", (size_t)Q * 3 * p.F);
      double *fd = d_force_dots ? d_force_dots : ctx->buf_t<double>("b2d.fd", Q * 3);
      double *co = d_neighbor_coordinates ? d_neighbor_coordinates
                                          : ctx->buf_t<double>("b2d.co", Q * 3);
      int *ni = d_neighbor_indices ? d_neighbor_indices : ctx->buf_t<int>("b2d.ni", Q);
      flare::launch_b2_dervs(ctx->stream, p, nl->dev(), d_species,
                             nl->b.row_of.as<int>(), ctx->ynorm_d.as<double>(),
                             ctx->b2_index(p.S, p.N, p.L), Q, fpair,
                             (const double *)fw.A, (const double *)fw.B, dB, fd,
                             co, ni);
    }
    ctx->sync();
  });
}

// ---- mapped pair style -----------------------------------------------------
// PairFLARE::read_file (pair_flare.cpp:263-402) on the host: header lines,
// then whitespace-separated values (grab, pair_flare.cpp:404-416).
static void parse_coefficients(const char *path, flare_pair &pr,
                               std::vector<double> &packed) {
  std::ifstream f(path);
  if (!f) throw Error{FLARE_EIO, std::string("Cannot open potential file ") + path};
  std::string line;
  auto next = [&](const char *what) {
    if (!std::getline(f, line))
      throw Error{FLARE_EFORMAT, std::string("unexpected end of file before ") + what};
  };
  next("header");  // date and contributor
  next("power");
  int power = 0;
  if (parse_int_line(line, power)) {
    pr.power = power;
    next("radial basis");
    pr.radial = first_word(line);
  } else {  // legacy layout: no power line, power 2
    pr.legacy = 1;
    pr.power = 2;
    pr.radial = first_word(line);
  }
  next("descriptor settings");
  {
    std::istringstream is(line);
    long long bs = 0;
    if (!(is >> pr.S >> pr.N >> pr.L >> bs))
      throw Error{FLARE_EFORMAT, "malformed 'n_species n_max l_max beta_size' line"};
    pr.beta_size = bs;
  }
  next("cutoff function");
  pr.cutoff_name = first_word(line);
  if (pr.radial != "chebyshev")
    throw Error{FLARE_EINVAL, "unsupported radial basis '" + pr.radial + "'"};
  if (pr.cutoff_name == "quadratic")
    pr.cutoff_kind = FLARE_CUTOFF_QUADRATIC;
  else if (pr.cutoff_name == "cosine")
    pr.cutoff_kind = FLARE_CUTOFF_COSINE;
  else
    throw Error{FLARE_EINVAL, "unsupported cutoff function '" + pr.cutoff_name + "'"};
  if (pr.power != 1 && pr.power != 2) throw Error{FLARE_EINVAL, "Power should be 1 or 2."};
  REQUIRE(pr.S >= 1 && pr.S <= flare::kMaxSpecies, "n_species must be in [1, 8]");
  REQUIRE(pr.N >= 1 && pr.N <= flare::kMaxN, "n_max must be in [1, 16]");
  REQUIRE(pr.L >= 0 && pr.L <= flare::kMaxL, "l_max must be in [0, 10]");
  pr.F = flare::b2_n_descriptors(pr.S, pr.N, pr.L);
  int64_t check = pr.power == 1 ? (int64_t)pr.F : (int64_t)pr.F * (pr.F + 1) / 2;
  if (check != pr.beta_size)
    throw Error{FLARE_EFORMAT, "Beta size doesn't match the number of descriptors."};
  const int n_cut = pr.legacy ? 1 : pr.S * pr.S;
  std::vector<double> vals;
  const int64_t need = n_cut + pr.S * pr.beta_size;
  vals.reserve(need);
  std::string tok;
  while ((int64_t)vals.size() < need && (f >> tok)) {
    char *end = nullptr;
    double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str() || *end != 0)
      throw Error{FLARE_EFORMAT, "non-numeric token '" + tok + "'"};
    vals.push_back(v);
  }
  if ((int64_t)vals.size() < need)
    throw Error{FLARE_EFORMAT, "file ends after " + std::to_string(vals.size()) +
                                   " of " + std::to_string(need) + " values"};
  pr.cutoffs.assign(pr.S * pr.S, vals[0]);
  if (!pr.legacy)
    for (int q = 0; q < pr.S * pr.S; q++) pr.cutoffs[q] = vals[q];
  pr.cutoff = -1;
  for (double v : pr.cutoffs) pr.cutoff = std::max(pr.cutoff, v);
  packed.assign(vals.begin() + n_cut, vals.end());
}

static void fill_info(const flare_pair &pr, flare_pair_info *o) {
  std::memset(o, 0, sizeof(*o));
  o->power = pr.power;
  o->n_species = pr.S;
  o->n_max = pr.N;
  o->l_max = pr.L;
  o->n_descriptors = pr.F;
  o->beta_size = pr.beta_size;
  o->cutoff_function = pr.cutoff_kind;
  o->legacy = pr.legacy;
  o->cutoff = pr.cutoff;
  std::snprintf(o->radial_basis, sizeof(o->radial_basis), "%s", pr.radial.c_str());
  std::snprintf(o->cutoff_name, sizeof(o->cutoff_name), "%s", pr.cutoff_name.c_str());
}

int flare_coeff_read(const char *path, flare_pair_info *info, double *cutoffs,
                     double *beta_packed) {
  return guarded([&] {
    REQUIRE(path && info, "null argument");
    flare_pair pr;
    std::vector<double> packed;
    parse_coefficients(path, pr, packed);
    fill_info(pr, info);
    if (cutoffs) std::copy(pr.cutoffs.begin(), pr.cutoffs.end(), cutoffs);
    if (beta_packed) std::copy(packed.begin(), packed.end(), beta_packed);
  });
}

int flare_pair_load(flare_ctx *ctx, const char *path, flare_pair **out) {
  return guarded([&] {
    REQUIRE(ctx && path && out, "null argument");
    *out = nullptr;
    ctx->use();
    std::unique_ptr<flare_pair> pr(new flare_pair);
    pr->ctx = ctx;
    std::vector<double> packed;
    parse_coefficients(path, *pr, packed);
    setup_pair(pr.get(), packed);
    *out = pr.release();
  });
}

int flare_pair_create(flare_ctx *ctx, int power, int n_species, int n_max,
                      int l_max, int cutoff_function, const double *cutoffs,
                      const double *beta_packed, flare_pair **out) {
  return guarded([&] {
    REQUIRE(ctx && cutoffs && beta_packed && out, "null argument");
    *out = nullptr;
    ctx->use();
    std::unique_ptr<flare_pair> pr(new flare_pair);
    pr->ctx = ctx;
    pr->power = power;
    pr->S = n_species;
    pr->N = n_max;
    pr->L = l_max;
    REQUIRE(cutoff_function == FLARE_CUTOFF_QUADRATIC ||
                cutoff_function == FLARE_CUTOFF_COSINE,
            "cutoff function must be quadratic or cosine");
    pr->cutoff_kind = cutoff_function;
    pr->cutoff_name = cutoff_function == FLARE_CUTOFF_QUADRATIC ? "quadratic" : "cosine";
    REQUIRE(n_species >= 1 && n_species <= flare::kMaxSpecies, "n_species must be in [1, 8]");
    int F = flare::b2_n_descriptors(n_species, n_max, l_max);
    pr->beta_size = power == 1 ? (int64_t)F : (int64_t)F * (F + 1) / 2;
    pr->cutoffs.assign(cutoffs, cutoffs + n_species * n_species);
    std::vector<double> packed(beta_packed, beta_packed + n_species * pr->beta_size);
    setup_pair(pr.get(), packed);
    *out = pr.release();
  });
}

int flare_pair_destroy(flare_pair *pr) {
  return guarded([&] {
    if (!pr) return;
    pr->ctx->use();
    (void)hipStreamSynchronize(pr->ctx->stream);
    delete pr;
  });
}

int flare_pair_get_info(const flare_pair *pr, flare_pair_info *o) {
  return guarded([&] {
    REQUIRE(pr && o, "null argument");
    fill_info(*pr, o);
  });
}

int flare_pair_get_cutoffs(const flare_pair *pr, double *cutoffs) {
  return guarded([&] {
    REQUIRE(pr && cutoffs, "null argument");
    std::copy(pr->cutoffs.begin(), pr->cutoffs.end(), cutoffs);
  });
}

int flare_pair_get_beta(flare_ctx *ctx, const flare_pair *pr, double *beta) {
  return guarded([&] {
    REQUIRE(ctx && pr && beta, "null argument");
    ctx->use();
    size_t per = pr->power == 1 ? (size_t)pr->F : (size_t)pr->F * pr->F;
    FLARE_HIP_CHECK(hipMemcpy(beta, pr->beta.ptr, pr->S * per * sizeof(double),
                              hipMemcpyDeviceToHost));
  });
}

}  // extern "C"

// Forward, GEMM and w of a mapped force evaluation over every centre of nl
// (pair_flare.cpp:95-137; lammps_descriptor.cpp:209-304).
struct PairEval {
  flare::B2Dev p;
  Forward fw;
  flare::WInput wi;
  int prec = 0, mode = 0;
  bool mixed = false;
  int off[flare::kMaxSpecies + 1];
};

static PairEval pair_eval(flare_ctx *ctx, flare_pair *pr, const flare_nlist *nl,
                          const int *d_species, int form, const double *d_beta,
                          int pair_pred) {
  REQUIRE(form >= FLARE_FORM_PLAIN_FP64 && form <= FLARE_FORM_ADJOINT_MIXED,
          "unknown form");
  // the "pair.*" rows (or the scratch) are about to be overwritten: no
  // earlier evaluation can be differentiated from them any more
  if (ctx->arena) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    g_scratch_gen[ctx->device] = ++g_eval_gen;
  } else {
    ctx->pair_ws_gen = ++g_eval_gen;
  }
  PairEval ev;
  flare_b2_params bp;
  bp.n_species = pr->S;
  bp.n_max = pr->N;
  bp.l_max = pr->L;
  bp.radial_basis = FLARE_RADIAL_CHEBYSHEV;
  bp.cutoff_function = pr->cutoff_kind;
  bp.radial_r1 = 0.0;  // radial_hyps = {0, cutoff} (pair_flare.cpp:346-349)
  bp.cutoffs = pr->cutoffs.data();
  flare::B2Dev p = make_b2(&bp, pair_pred);
  ev.p = p;
  check_species(nl, p.S);
  const int n = nl->n;
  const int F = p.F;
  const int prec = (form == FLARE_FORM_PLAIN_FP64 || form == FLARE_FORM_ADJOINT_FP64)
                       ? flare::kPrecF64
                       : flare::kPrecF32;
  ev.prec = prec;
  ev.mode = form == FLARE_FORM_PLAIN_FP64 ? 0 : 1;
  const size_t per = pr->power == 1 ? (size_t)F : (size_t)F * F;
  const size_t ts = prec == flare::kPrecF64 ? sizeof(double) : sizeof(float);
  const bool mixed = form == FLARE_FORM_ADJOINT_MIXED && pr->power == 2;
  ev.mixed = mixed;
  // GEMM rows: B itself in FP64 (as pair_flare), B/|B| in the FP32 forms
  const int normalized = prec == flare::kPrecF32 ? 1 : 0;
  Forward fw = run_forward(ctx, p, nl, d_species, prec, "pair", nullptr,
                           mixed ? 2 : normalized, mixed);
  int *off = ev.off;
  nl->row_offsets(p.S, off);
  // beta in the working precision
  const void *beta = pr->beta.ptr;
  const void *bh = pr->beta_h.ptr, *bl = pr->beta_l.ptr;
  const int kp = fw.kp;
  if (prec == flare::kPrecF64) {
    if (d_beta) beta = d_beta;
  } else {
    beta = pr->beta_f32.ptr;
    if (d_beta) {
      float *bf = ctx->buf_t<float>("pair.beta_override", p.S * per);
      flare::launch_to_f32(ctx->stream, (int64_t)(p.S * per), d_beta, bf);
      beta = bf;
      if (mixed) {
        void *h = ctx->buf("pair.beta_oh", (size_t)p.S * F * kp * 2);
        void *l = ctx->buf("pair.beta_ol", (size_t)p.S * F * kp * 2);
        flare::split_bf16(ctx->stream, p.S * F, F, kp, d_beta, h, l);
        bh = h;
        bl = l;
      }
    }
  }
  const void *X = fw.B;
  void *Z = nullptr;
  ctx->mark_begin("gemm");
  if (pr->power == 2) {
    Z = ctx->buf("pair.Z", std::max<size_t>(1, (size_t)n * F) * ts);
    if (mixed) {  // all species in one launch
      flare::GemmProduct g[flare::kMaxSpecies];
      for (int s = 0; s < p.S; s++)
        g[s] = flare::GemmProduct{off[s + 1] - off[s], F, kp, F,
                                  (int64_t)off[s] * kp, (int64_t)s * F * kp,
                                  (int64_t)off[s] * F};
      // wide rows: 256 x 352 tiles (one round of workgroups at F = 680)
      if (wide_cols_fit(F, 256, 10))
        flare::gemm_bf16x3_wide(ctx->stream, p.S, g, fw.xh, fw.xl, bh, bl, (float *)Z);
      else
        flare::gemm_bf16x3_batched(ctx->stream, p.S, g, fw.xh, fw.xl, bh, bl,
                                   (float *)Z);
    }
    // FP64 and FP32: the hand-written matrix-core GEMMs where they were
    // measured faster than rocBLAS (F = 680 at 32,000 atoms: FP64 0.575 vs
    // 0.636 ms, FP32 0.25 on 256 x 352 tiles, 0.296 on 128 x 128, 0.413 in
    // rocBLAS); FP32 wide rows on the 256 x 352 tiles past 1,024 columns
    // (F = 2,100 at 262,144 atoms: as fast as rocBLAS) and below where they
    // pad no more than the 128 x 128 tiles, FP64 rows past 1,024 on rocBLAS
    // (35.3 vs 39.7 ms there)
    const bool f32_wide = !mixed && prec == flare::kPrecF32 && F % 4 == 0 &&
                          (F > 1024 || wide_cols_fit(F, 128, 0));
    const bool mcore = !mixed && (F <= 1024 || f32_wide) &&
                       (prec == flare::kPrecF64 ? F % 2 == 0 : F % 4 == 0);
    if (mcore) {  // all species in one launch
      flare::GemmProduct g[flare::kMaxSpecies];
      for (int s = 0; s < p.S; s++)
        g[s] = flare::GemmProduct{off[s + 1] - off[s], F, F, F, (int64_t)off[s] * F,
                                  (int64_t)s * (int64_t)per, (int64_t)off[s] * F};
      if (prec == flare::kPrecF64)
        flare::dgemm_batched(ctx->stream, p.S, g, (const double *)X, (const double *)beta,
                             (double *)Z);
      else if (f32_wide)
        flare::sgemm_wide(ctx->stream, p.S, g, (const float *)X, (const float *)beta,
                          (float *)Z);
      else
        flare::sgemm_batched(ctx->stream, p.S, g, (const float *)X, (const float *)beta,
                             (float *)Z);
    }
    for (int s = 0; s < p.S && !mixed && !mcore; s++) {
      int rows = off[s + 1] - off[s];
      if (!rows) continue;
      if (prec == flare::kPrecF64) {
        gemm_rm(ctx, false, false, rows, F, F, 1.0,
                (const double *)X + (size_t)off[s] * F, F,
                (const double *)beta + s * per, F, 0.0,
                (double *)Z + (size_t)off[s] * F, F);
      } else {
        sgemm_rm(ctx, false, false, rows, F, F, 1.0f,
                 (const float *)X + (size_t)off[s] * F, F,
                 (const float *)beta + s * per, F, 0.0f,
                 (float *)Z + (size_t)off[s] * F, F);
      }
    }
  }
  ctx->mark_end();
  double *E_row = ctx->buf_t<double>("pair.E", n);
  flare::WInput wi{mixed ? 3 : pr->power == 2 ? 1 : 2, normalized, X, Z,
                   pr->power == 1 ? beta : nullptr, nullptr, fw.norm2, E_row};
  if (mixed) {
    wi.xh = fw.xh;
    wi.xl = fw.xl;
    wi.kp = kp;
  }
  ev.fw = fw;
  ev.wi = wi;
  return ev;
}

static void pair_remember(flare_ctx *ctx, flare_pair *pr, const PairEval &ev, int n) {
  const uint64_t gen = ++g_eval_gen;
  pr->have_last = true;
  pr->last_mixed = ev.mixed;
  pr->last_prec = ev.prec;
  pr->last_n = n;
  for (int s = 0; s <= ev.p.S; s++) pr->last_off[s] = ev.off[s];
  pr->last_ctx = ctx;
  pr->last_device = ctx->device;
  pr->last_gen = gen;
  pr->last_scratch = ctx->arena != nullptr;
  pr->last_B = ev.fw.B;
  pr->last_xh = ev.fw.xh;
  pr->last_xl = ev.fw.xl;
  pr->last_norm2 = ev.fw.norm2;
  if (pr->last_scratch) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    g_scratch_gen[ctx->device] = gen;
  } else {
    ctx->pair_ws_gen = gen;
  }
}

extern "C" {

int flare_pair_compute(flare_ctx *ctx, flare_pair *pr, const flare_nlist *nl,
                       const int *d_species, int form, const double *d_beta,
                       double *d_energy, double *d_forces, double *d_virial,
                       double *d_eatom) {
  return guarded([&] {
    REQUIRE(ctx && pr, "null argument");
    ctx->use();
    PairEval ev = pair_eval(ctx, pr, nl, d_species, form, d_beta, 1);
    const int n = nl->n;
    run_forces(ctx, ev.p, nl, d_species, ev.prec, ev.mode, ev.fw, ev.wi, d_forces,
               d_energy, d_virial, nullptr, "pair");
    if (d_eatom)
      flare::launch_rows_to_atoms(ctx->stream, n, nl->b.atom_of_row.as<int>(),
                                  ev.wi.E_row, d_eatom);
    pair_remember(ctx, pr, ev, n);
    ctx->finish();
  });
}

int flare_pair_compute_noalloc(flare_ctx *ctx, flare_pair *pr, const flare_nlist *nl,
                               const int *d_species, int form, const double *d_beta,
                               double *d_energy, double *d_forces, double *d_virial,
                               double *d_eatom) {
  return guarded([&] {
    REQUIRE(ctx && pr, "null argument");
    REQUIRE(nl && nl->built, "neighbor list not built");
    ctx->use();
    char *base = scratch_of(ctx->device);
    REQUIRE(base, "no static scratch on this device");
    // a pending build is completed here only when that needs no new buffers
    // (rows over the one-pass capacity are rebuilt into larger ones)
    if (nl->pending) {
      nl->wait_stats();
      REQUIRE(nl->stats_host[flare::kMaxSpecies + 1] <= flare::kNbrCap,
              "the neighbor list has atoms with more than " +
                  std::to_string(flare::kNbrCap) +
                  " neighbors, whose rebuild allocates: complete it first "
                  "(flare_nlist_get_view or flare_pair_compute)");
    }
    // rocBLAS manages device workspace of its own: only the forms whose
    // contraction is a hand-written kernel are served here (FP64 up to
    // F = 1,024, FP32 with F a multiple of 4, mixed at any F)
    if (pr->power == 2 && form != FLARE_FORM_ADJOINT_MIXED) {
      const bool f64 = form == FLARE_FORM_PLAIN_FP64 || form == FLARE_FORM_ADJOINT_FP64;
      REQUIRE(f64 ? pr->F <= 1024 && pr->F % 2 == 0 : pr->F % 4 == 0,
              std::string(f64 ? "the no-allocation entry point needs F <= 1024 and F a "
                                "multiple of 2 in this form (F = "
                              : "the no-allocation entry point needs F a multiple of 4 "
                                "in this form (F = ") +
                  std::to_string(pr->F) + ")");
    }
    flare_ctx::Arena arena;
    arena.base = base;
    arena.cap = FLARE_NOALLOC_SCRATCH_BYTES;
    // stage timing would create HIP events: off for the call
    struct Scope {
      flare_ctx *c;
      int profiling;
      ~Scope() {
        c->arena = nullptr;
        c->profiling = profiling;
      }
    } scope{ctx, ctx->profiling};
    ctx->collect();
    ctx->profiling = 0;
    ctx->arena = &arena;
    PairEval ev = pair_eval(ctx, pr, nl, d_species, form, d_beta, 1);
    const int n = nl->n;
    run_forces(ctx, ev.p, nl, d_species, ev.prec, ev.mode, ev.fw, ev.wi, d_forces,
               d_energy, d_virial, nullptr, "pair");
    if (d_eatom)
      flare::launch_rows_to_atoms(ctx->stream, n, nl->b.atom_of_row.as<int>(),
                                  ev.wi.E_row, d_eatom);
    pair_remember(ctx, pr, ev, n);
    ctx->finish();
  });
}

int flare_pair_compute_neigh(flare_ctx *ctx, flare_pair *pr, int n_all, int inum,
                             const int *d_ilist, const int *d_numneigh,
                             const int *d_neighbors, int64_t stride_i,
                             int64_t stride_j, const double *d_x,
                             const int *d_type, int form, double *d_energy,
                             double *d_forces, double *d_virial, double *d_eatom) {
  return guarded([&] {
    REQUIRE(ctx && pr, "null argument");
    REQUIRE(n_all >= 0 && inum >= 0 && inum <= n_all, "need 0 <= inum <= n_all");
    REQUIRE(inum == 0 || (d_ilist && d_numneigh && d_neighbors && d_x && d_type),
            "null neighbor list or atom arrays");
    REQUIRE(d_energy && d_forces && d_virial, "null output");
    ctx->use();
    // the caller's list as a CSR list over the inum centres
    if (!pr->ext) {
      pr->ext.reset(new flare_nlist);
      pr->ext->ctx = ctx;
    }
    flare_nlist *nl = pr->ext.get();
    flare::ExtList e{n_all, inum, d_ilist, d_numneigh, d_neighbors, stride_i,
                     stride_j, d_x, d_type};
    int *cspec = ctx->buf_t<int>("ext.cspec", inum);
    ctx->mark_begin("nlist");
    flare::nlist_from_external(ctx->stream, nl->b, e, pr->cutoffs.data(), pr->S, cspec,
                               nl->species_count);
    ctx->mark_end();
    nl->n = inum;
    nl->wh.volume = 1.0;  // no stress on this path
    nl->pending = false;
    nl->built = true;
    PairEval ev = pair_eval(ctx, pr, nl, cspec, form, nullptr, 2);
    const flare::PairForces pf = run_pair_forces(ctx, ev.p, nl, cspec, ev.prec, ev.mode,
                                                 ev.fw, ev.wi, "pair");
    double *own = ctx->buf_t<double>("ext.own", (size_t)inum * 3);
    run_gather_reduce(ctx, nl, pf, ev.wi, own, d_energy, d_virial, nullptr, "pair");
    // F_j -= f_ij over the pairs whose neighbor is j (ghosts included), then
    // F_i += own sums; LAMMPS sums ghost forces back (newton on)
    ctx->mark_begin("assemble");
    int *in_cum = ctx->buf_t<int>("ext.in_cum", (size_t)n_all + 1);
    int *in_pairs = ctx->buf_t<int>("ext.in_pairs", (size_t)std::max<int64_t>(nl->b.P, 1));
    flare::ext_incoming(ctx->stream, nl->b, n_all, nl->b.P, in_cum, in_pairs);
    flare::ext_assemble(ctx->stream, n_all, inum, d_ilist, in_cum, in_pairs, pf, own,
                        nl->b.row_of.as<int>(), ev.wi.E_row, d_forces, d_eatom);
    ctx->mark_end();
    pair_remember(ctx, pr, ev, inum);
    ctx->finish();
  });
}

int flare_pair_beta_grad(flare_ctx *ctx, flare_pair *pr, double *d_grad) {
  return guarded([&] {
    REQUIRE(ctx && pr && d_grad, "null argument");
    REQUIRE(pr->have_last, "no force evaluation to differentiate");
    // the rows of this pair's last evaluation must still be where it left
    // them: no later evaluation (of any pair) may have reused that memory
    bool current;
    if (pr->last_scratch) {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      current = ctx->device == pr->last_device && g_scratch_gen[pr->last_device] == pr->last_gen;
    } else {
      current = ctx == pr->last_ctx && ctx->pair_ws_gen == pr->last_gen;
    }
    REQUIRE(current, "dE/dbeta needs the descriptor rows of this pair's last force "
                     "evaluation, and a later evaluation (of this or another pair, on this "
                     "context or in the device's no-allocation scratch) has overwritten "
                     "them; evaluate this pair again first");
    ctx->use();
    const int n = pr->last_n, F = pr->F;
    const int *off = pr->last_off;
    const double *norm2 = pr->last_norm2;
    double *dhat = ctx->buf_t<double>("pair.dhat", (size_t)n * F);
    if (pr->last_prec == flare::kPrecF64)
      flare::launch_normalize_rows_f64(ctx->stream, n, F, (const double *)pr->last_B,
                                       norm2, dhat);
    else if (pr->last_mixed)  // B/|B| as the GEMM's BF16 hi + lo rows
      flare::launch_bf16_pair_to_f64(ctx->stream, n, F, flare::bf16_padded_k(F),
                                     pr->last_xh, pr->last_xl, dhat);
    else  // the FP32 forms keep the normalised rows B/|B|
      flare::launch_to_f64(ctx->stream, (int64_t)n * F, (const float *)pr->last_B, dhat);
    // empty environments (|B|^2 < 1e-8) carry no energy
    flare::launch_mask_rows(ctx->stream, n, F, norm2, 1e-8, dhat);
    for (int s = 0; s < pr->S; s++) {
      int rows = off[s + 1] - off[s];
      if (pr->power == 2) {
        double *g = d_grad + (size_t)s * F * F;
        if (!rows) {
          FLARE_HIP_CHECK(hipMemsetAsync(g, 0, (size_t)F * F * sizeof(double), ctx->stream));
          continue;
        }
        gemm_rm(ctx, true, false, F, F, rows, 1.0, dhat + (size_t)off[s] * F, F,
                dhat + (size_t)off[s] * F, F, 0.0, g, F);
      } else {
        double *g = d_grad + (size_t)s * F;
        if (!rows) {
          FLARE_HIP_CHECK(hipMemsetAsync(g, 0, (size_t)F * sizeof(double), ctx->stream));
          continue;
        }
        flare::launch_colsum(ctx->stream, rows, F, dhat + (size_t)off[s] * F, g);
      }
    }
    ctx->sync();
  });
}

int flare_write_coefficients(const char *path, const char *contributor,
                             int power, int n_species, int n_max, int l_max,
                             const char *radial_basis,
                             const char *cutoff_function,
                             const double *cutoffs, int64_t beta_size,
                             const double *beta_packed) {
  return guarded([&] {
    REQUIRE(path && cutoffs && beta_packed, "null argument");
    FILE *f = std::fopen(path, "w");
    if (!f) throw Error{FLARE_EIO, std::string("cannot write ") + path};
    time_t now = std::time(nullptr);
    std::string t(std::ctime(&now));
    if (!t.empty() && t.back() == '\n') t.pop_back();
    std::fprintf(f, "DATE: %s CONTRIBUTOR: %s\n", t.c_str(),
                 contributor ? contributor : "");
    std::fprintf(f, "%d\n", power);
    std::fprintf(f, "%s\n", radial_basis ? radial_basis : "chebyshev");
    std::fprintf(f, "%d %d %d %lld\n", n_species, n_max, l_max, (long long)beta_size);
    std::fprintf(f, "%s\n", cutoff_function ? cutoff_function : "quadratic");
    for (int q = 0; q < n_species * n_species; q++) std::fprintf(f, "%.2f ", cutoffs[q]);
    std::fprintf(f, "\n");
    // five values a line, two spaces before positives, one before the rest;
    // the line counter runs on across species (sparse_gp.cpp:1173-1203)
    int count = 0;
    for (int s = 0; s < n_species; s++) {
      if (count != 0) std::fprintf(f, "\n");
      for (int64_t j = 0; j < beta_size; j++) {
        double v = beta_packed[(int64_t)s * beta_size + j];
        std::fprintf(f, v > 0 ? "  %.16e" : " %.16e", v);
        if (++count == 5) {
          count = 0;
          std::fprintf(f, "\n");
        }
      }
    }
    if (std::fclose(f) != 0) throw Error{FLARE_EIO, std::string("cannot write ") + path};
  });
}

// ---- sparse GP ---------------------------------------------------------------
static void sgp_factor(flare_sgp *gp) {
  flare_ctx *c = gp->ctx;
  const int M = gp->M, F = gp->F;
  int *info = c->buf_t<int>("sgp.info", 1);
  FLARE_HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int), c->stream));
  gp->Kuu.ensure((size_t)std::max(M, 1) * M * sizeof(double));
  gp->Lcm.ensure((size_t)std::max(M, 1) * M * sizeof(double));
  gp->Linv_cm.ensure((size_t)std::max(M, 1) * M * sizeof(double));
  double *K = gp->Kuu.as<double>(), *Lc = gp->Lcm.as<double>(),
         *Li = gp->Linv_cm.as<double>();
  FLARE_HIP_CHECK(hipMemsetAsync(K, 0, (size_t)M * M * sizeof(double), c->stream));
  FLARE_HIP_CHECK(hipMemsetAsync(Li, 0, (size_t)M * M * sizeof(double), c->stream));
  for (int s = 0; s < gp->S; s++) {
    int Ms = gp->m_by_type[s], o = gp->m_off[s];
    if (!Ms) continue;
    const double *Ps = gp->P.as<double>() + (size_t)o * F;
    double *Ks = K + (size_t)o * M + o;
    gemm_rm(c, false, true, Ms, Ms, F, 1.0, Ps, F, Ps, F, 0.0, Ks, M);
    flare::launch_kuu_block(c->stream, Ms, M, Ks, gp->pnorm.as<double>() + o,
                            gp->sig2, gp->power);
  }
  FLARE_HIP_CHECK(hipMemcpyAsync(Lc, K, (size_t)M * M * sizeof(double),
                                 hipMemcpyDeviceToDevice, c->stream));
  flare::launch_add_diag(c->stream, M, Lc, M, gp->jitter);
  for (int s = 0; s < gp->S; s++) {
    int Ms = gp->m_by_type[s], o = gp->m_off[s];
    if (!Ms) continue;
    double *A = Lc + (size_t)o * M + o;
    const int nb = 64;
    for (int j = 0; j < Ms; j += nb) {
      int jb = std::min(nb, Ms - j);
      double *A11 = A + (size_t)j * M + j;
      flare::launch_chol_panel(c->stream, jb, A11, M, info);
      int m2 = Ms - j - jb;
      if (m2 > 0) {
        double one = 1.0, mone = -1.0;
        double *A21 = A + (size_t)j * M + (j + jb);
        double *A22 = A + (size_t)(j + jb) * M + (j + jb);
        BLAS_CHECK(rocblas_dtrsm(c->blas, rocblas_side_right, rocblas_fill_lower,
                                 rocblas_operation_transpose,
                                 rocblas_diagonal_non_unit, m2, jb, &one, A11, M,
                                 A21, M));
        BLAS_CHECK(rocblas_dsyrk(c->blas, rocblas_fill_lower,
                                 rocblas_operation_none, m2, jb, &mone, A21, M,
                                 &one, A22, M));
      }
    }
  }
  flare::launch_zero_upper(c->stream, M, Lc, M);
  for (int s = 0; s < gp->S; s++) {
    int Ms = gp->m_by_type[s], o = gp->m_off[s];
    if (!Ms) continue;
    BLAS_CHECK(rocblas_dtrtri(c->blas, rocblas_fill_lower,
                              rocblas_diagonal_non_unit, Ms,
                              Lc + (size_t)o * M + o, M, Li + (size_t)o * M + o, M));
  }
  int hinfo = 0;
  FLARE_HIP_CHECK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost,
                                 c->stream));
  c->sync();
  REQUIRE(hinfo == 0, "K_uu + jitter is not positive definite");
}

int flare_sgp_create(flare_ctx *ctx, int n_species, int n_descriptors,
                     const int *n_sparse_by_type, const double *d_sparse,
                     double sigma, double power, double jitter,
                     flare_sgp **out) {
  return guarded([&] {
    REQUIRE(ctx && n_sparse_by_type && out, "null argument");
    REQUIRE(n_species >= 1 && n_species <= flare::kMaxSpecies, "n_species must be in [1, 8]");
    REQUIRE(n_descriptors > 0, "n_descriptors must be positive");
    REQUIRE(power > 0, "power must be positive");
    *out = nullptr;
    ctx->use();
    std::unique_ptr<flare_sgp> gp(new flare_sgp);
    gp->ctx = ctx;
    gp->S = n_species;
    gp->F = n_descriptors;
    gp->sigma = sigma;
    gp->sig2 = sigma * sigma;
    gp->power = power;
    gp->jitter = jitter;
    gp->m_by_type.assign(n_sparse_by_type, n_sparse_by_type + n_species);
    gp->m_off.assign(n_species + 1, 0);
    for (int s = 0; s < n_species; s++) {
      REQUIRE(gp->m_by_type[s] >= 0, "negative sparse count");
      gp->m_off[s + 1] = gp->m_off[s] + gp->m_by_type[s];
    }
    gp->M = gp->m_off[n_species];
    const int M = gp->M, F = gp->F;
    REQUIRE(M == 0 || d_sparse, "null sparse descriptors");
    gp->P.ensure((size_t)std::max(M, 1) * F * sizeof(double));
    gp->Phat.ensure((size_t)std::max(M, 1) * F * sizeof(double));
    gp->pnorm.ensure((size_t)std::max(M, 1) * sizeof(double));
    gp->pnorm2.ensure((size_t)std::max(M, 1) * sizeof(double));
    gp->alpha.ensure((size_t)std::max(M, 1) * sizeof(double));
    FLARE_HIP_CHECK(hipMemsetAsync(gp->alpha.ptr, 0, std::max(M, 1) * sizeof(double),
                                   ctx->stream));
    if (M) {
      FLARE_HIP_CHECK(hipMemcpyAsync(gp->P.ptr, d_sparse, (size_t)M * F * sizeof(double),
                                     hipMemcpyDeviceToDevice, ctx->stream));
      flare::launch_row_norms(ctx->stream, M, F, gp->P.as<double>(),
                              gp->pnorm.as<double>(), gp->Phat.as<double>());
      flare::launch_square(ctx->stream, M, gp->pnorm.as<double>(), gp->pnorm2.as<double>());
    }
    sgp_factor(gp.get());
    *out = gp.release();
  });
}

int flare_sgp_destroy(flare_sgp *gp) {
  return guarded([&] {
    if (!gp) return;
    gp->ctx->use();
    (void)hipStreamSynchronize(gp->ctx->stream);
    delete gp;
  });
}

int flare_sgp_set_alpha(flare_ctx *ctx, flare_sgp *gp, const double *d_alpha) {
  return guarded([&] {
    REQUIRE(ctx && gp, "null argument");
    REQUIRE(gp->M == 0 || d_alpha, "null alpha");
    ctx->use();
    if (gp->M)
      FLARE_HIP_CHECK(hipMemcpyAsync(gp->alpha.ptr, d_alpha, gp->M * sizeof(double),
                                     hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sync();
    gp->have_alpha = true;
  });
}

int flare_sgp_get_L_inverse(flare_ctx *ctx, const flare_sgp *gp, double *d_out) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_out, "null argument");
    ctx->use();
    flare::launch_transpose(ctx->stream, gp->M, gp->Linv_cm.as<double>(), d_out);
    ctx->sync();
  });
}

int flare_sgp_get_Kuu(flare_ctx *ctx, const flare_sgp *gp, double *d_out) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_out, "null argument");
    ctx->use();
    if (gp->M)
      FLARE_HIP_CHECK(hipMemcpyAsync(d_out, gp->Kuu.ptr,
                                     (size_t)gp->M * gp->M * sizeof(double),
                                     hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sync();
  });
}

struct GPState {
  Forward fw;
  double *kval = nullptr;  // per species blocks rows_s x ldk_s
  std::vector<size_t> koff;
  bool lowp = false;
};

static void sgp_build_lowp(flare_ctx *ctx, flare_sgp *gp) {
  if (gp->have_lowp) return;
  const int M = gp->M, F = gp->F, S = gp->S;
  gp->kp = flare::bf16_padded_k(F);
  gp->mp.assign(S, 0);
  gp->tpoff.assign(S + 1, 0);
  for (int s = 0; s < S; s++) {
    gp->mp[s] = flare::bf16_padded_k(std::max(gp->m_by_type[s], 1));
    gp->tpoff[s + 1] = gp->tpoff[s] + (int64_t)F * gp->mp[s];
  }
  gp->Phat_f.ensure((size_t)std::max(M, 1) * F * sizeof(float));
  gp->Phat_h.ensure((size_t)std::max(M, 1) * gp->kp * 2);
  gp->Phat_l.ensure((size_t)std::max(M, 1) * gp->kp * 2);
  gp->Phat_is.ensure((size_t)std::max(M, 1) * sizeof(float));
  gp->PhatT_h.ensure((size_t)std::max<int64_t>(gp->tpoff[S], 1) * 2);
  gp->PhatT_l.ensure((size_t)std::max<int64_t>(gp->tpoff[S], 1) * 2);
  gp->PhatT_is.ensure((size_t)S * F * sizeof(float));
  if (M) {
    flare::launch_to_f32(ctx->stream, (int64_t)M * F, gp->Phat.as<double>(),
                         gp->Phat_f.as<float>());
    flare::split_f16_rows(ctx->stream, M, F, F, gp->kp, gp->Phat.as<double>(),
                          gp->Phat_h.ptr, gp->Phat_l.ptr, gp->Phat_is.as<float>());
  }
  for (int s = 0; s < S; s++)
    flare::split_f16_cols(ctx->stream, gp->m_by_type[s], F, gp->mp[s],
                          gp->Phat.as<double>() + (size_t)gp->m_off[s] * F,
                          (_Float16 *)gp->PhatT_h.ptr + gp->tpoff[s],
                          (_Float16 *)gp->PhatT_l.ptr + gp->tpoff[s],
                          gp->PhatT_is.as<float>() + (size_t)s * F);
  gp->have_lowp = true;
}

// Operands of the matrix-core GP products, built once per model: per
// species the transposed normalised sparse rows P_hat^T (F rows of ldk_s,
// FP64 and FP32) and the species' block of L^-1 as rows of ldk_s (row-major,
// lower triangular), ldk_s = M_s rounded up to 16 (zero padding).
static void sgp_build_pad(flare_ctx *ctx, flare_sgp *gp) {
  if (gp->have_pad) return;
  const int M = gp->M, F = gp->F, S = gp->S;
  gp->ldk.assign(S, 16);
  gp->toff.assign(S + 1, 0);
  gp->loff.assign(S + 1, 0);
  for (int s = 0; s < S; s++) {
    gp->ldk[s] = (std::max(gp->m_by_type[s], 1) + 15) / 16 * 16;
    gp->toff[s + 1] = gp->toff[s] + (int64_t)F * gp->ldk[s];
    gp->loff[s + 1] = gp->loff[s] + (int64_t)gp->m_by_type[s] * gp->ldk[s];
  }
  gp->PhatT_d.ensure((size_t)std::max<int64_t>(gp->toff[S], 1) * sizeof(double));
  gp->PhatT_f.ensure((size_t)std::max<int64_t>(gp->toff[S], 1) * sizeof(float));
  gp->Linv_rm.ensure((size_t)std::max<int64_t>(gp->loff[S], 1) * sizeof(double));
  for (int s = 0; s < S; s++) {
    const int Ms = gp->m_by_type[s], o = gp->m_off[s];
    const double *ph = gp->Phat.as<double>() + (size_t)o * F;
    // P_hat^T: element (f, u) = P_hat[o + u][f]
    flare::launch_pad_block(ctx->stream, F, Ms, ph, 1, F,
                            gp->PhatT_d.as<double>() + gp->toff[s], gp->ldk[s]);
    flare::launch_pad_block_f32(ctx->stream, F, Ms, ph, 1, F,
                                gp->PhatT_f.as<float>() + gp->toff[s], gp->ldk[s]);
    // L^-1 rows: element (u, v) = L^-1[o + u][o + v] (column-major source)
    flare::launch_pad_block(ctx->stream, Ms, Ms, gp->Linv_cm.as<double>() + (size_t)o * M + o,
                            1, M, gp->Linv_rm.as<double>() + gp->loff[s], gp->ldk[s]);
  }
  gp->have_pad = true;
}

// Whether the matrix-core products take these rows (16-byte K rows)
static bool gp_mcore(const flare_sgp *gp, bool f32) { return gp->F % (f32 ? 4 : 2) == 0; }

// FP32 / mixed forms of the GP mean: normalised FP32 rows d, t = d . p_hat
// in FP32 (matrix-core GEMM with the kernel-value epilogue fused) or on FP16
// MFMA with hi+lo splitting and power-of-two row scaling, G = c P_hat the
// same way, k, c, e, w in FP64 from t, forces from the FP32 adjoint kernel.
static GPState gp_mean_lowp(flare_ctx *ctx, flare_sgp *gp, const flare::B2Dev &p,
                            const flare_nlist *nl, const int *d_species, int form,
                            double *d_mean_efs, bool want_k) {
  const bool mixed = form == FLARE_FORM_ADJOINT_MIXED;
  const int n = nl->n, F = p.F;
  const bool mcore = !mixed && gp_mcore(gp, true);
  sgp_build_lowp(ctx, gp);
  sgp_build_pad(ctx, gp);
  GPState st;
  st.lowp = true;
  // mixed: the forward also writes the FP16 hi + lo split of the d rows
  // with their power-of-two row scales (t's A operand)
  void *dh = nullptr, *dl = nullptr;
  float *d_is = nullptr;
  if (mixed) {
    REQUIRE(gp->kp == flare::bf16_padded_k(F), "padded row length");
    dh = ctx->buf("gpl.dh", std::max<size_t>((size_t)n * gp->kp, 1) * 2);
    dl = ctx->buf("gpl.dl", std::max<size_t>((size_t)n * gp->kp, 1) * 2);
    d_is = ctx->buf_t<float>("gpl.d_is", std::max(n, 1));
  }
  st.fw = run_forward(ctx, p, nl, d_species, flare::kPrecF32, "gpl", nullptr, 1, false, dh,
                      dl, d_is);
  const float *d = (const float *)st.fw.B;
  int off[flare::kMaxSpecies + 1];
  nl->row_offsets(p.S, off);
  size_t ktot = 0, ctot = 0, rtot = 0;
  st.koff.assign(p.S + 1, 0);
  std::vector<size_t> coff(p.S + 1, 0), roff(p.S + 1, 0);
  for (int s = 0; s < p.S; s++) {
    st.koff[s] = ktot;
    coff[s] = ctot;
    roff[s] = rtot;
    ktot += (size_t)(off[s + 1] - off[s]) * gp->ldk[s];
    ctot += (size_t)(off[s + 1] - off[s]) * gp->mp[s];
    rtot += (size_t)(off[s + 1] - off[s]) * gp->m_by_type[s];
  }
  st.koff[p.S] = ktot;
  coff[p.S] = ctot;
  roff[p.S] = rtot;
  // e partial sums: one per 64 sparse columns (matrix-core path), or the
  // row's sum itself (FP16 path)
  int np = 1;
  for (int s = 0; s < p.S; s++) np = std::max(np, (gp->m_by_type[s] + 63) / 64);
  if (!mcore) np = 1;
  st.kval = ctx->buf_t<double>("gpl.kval", std::max<size_t>(ktot, 1));
  float *G = ctx->buf_t<float>("gpl.G", std::max<size_t>((size_t)n * F, 1));
  double *e_part = ctx->buf_t<double>("gpl.e", std::max<size_t>((size_t)n * np, 1));
  float *cf = nullptr, *c_is = nullptr, *Kt = nullptr;
  void *ch = nullptr, *cl = nullptr;
  if (mixed) {
    ch = ctx->buf("gpl.ch", std::max<size_t>(ctot, 1) * 2);
    cl = ctx->buf("gpl.cl", std::max<size_t>(ctot, 1) * 2);
    c_is = ctx->buf_t<float>("gpl.c_is", std::max(n, 1));
    Kt = ctx->buf_t<float>("gpl.K", std::max<size_t>(2 * rtot, 1));  // room for a K split
  } else {
    cf = ctx->buf_t<float>("gpl.c", std::max<size_t>(ktot, 1));
    if (!mcore) Kt = ctx->buf_t<float>("gpl.K", std::max<size_t>(rtot, 1));
  }
  int ksplit = 1;
  ctx->mark_begin("gp_K");
  if (mixed) {
    const int kp = gp->kp;
    flare::GemmProduct g[2 * flare::kMaxSpecies];
    int width = 0;
    for (int s = 0; s < p.S; s++) {
      g[s] = flare::GemmProduct{off[s + 1] - off[s], gp->m_by_type[s], kp,
                                gp->m_by_type[s], (int64_t)off[s] * kp,
                                (int64_t)gp->m_off[s] * kp, (int64_t)roff[s]};
      g[s].rs = d_is + off[s];
      g[s].cs = gp->Phat_is.as<float>() + gp->m_off[s];
      width = std::max(width, std::max(gp->m_by_type[s], std::max(gp->ldk[s], gp->mp[s])));
    }
    // 256 x 352 tiles when they make a single round of workgroups, with K
    // split in two halves (summed by the epilogue) when that still fits one
    // round (32,000 atoms, 250 sparse per species: 96 us on 128 x 256 tiles,
    // 83 us on 126 wide tiles), 128 x 256 tiles otherwise (55,296 atoms,
    // 667 per species: 367 vs 418 us)
    const int wt = flare::gemm_x3_wide_tiles(p.S, g);
    if (wt <= ctx->n_cu) {
      if (2 * wt <= ctx->n_cu && 2 * p.S <= flare::kGemmMaxBatch && kp % 64 == 0 &&
          width <= 16 * flare::kWave) {
        ksplit = 2;
        for (int s = 0; s < p.S; s++) {
          flare::GemmProduct &h0 = g[s], &h1 = g[p.S + s];
          h0.K = kp / 2;
          h0.ld_ab = kp;
          h1 = h0;
          h1.a_off += kp / 2;
          h1.b_off += kp / 2;
          h1.z_off += (int64_t)rtot;
        }
      }
      flare::gemm_f16x3_wide(ctx->stream, ksplit * p.S, g, dh, dl, gp->Phat_h.ptr,
                             gp->Phat_l.ptr, Kt);
    } else {
      flare::gemm_f16x3_batched(ctx->stream, p.S, g, dh, dl, gp->Phat_h.ptr,
                                gp->Phat_l.ptr, Kt);
    }
  } else if (mcore) {
    // t = d . p_hat with the kernel values, c = alpha t^(xi-1) and the
    // partial energies in the GEMM's epilogue
    flare::GemmProduct g[flare::kMaxSpecies];
    flare::DEpilogue ep[flare::kMaxSpecies];
    for (int s = 0; s < p.S; s++) {
      const int o = gp->m_off[s];
      g[s] = flare::GemmProduct{off[s + 1] - off[s], gp->m_by_type[s], F, 0,
                                (int64_t)off[s] * F, (int64_t)o * F, 0};
      ep[s].mode = flare::DEpilogue::kKernel;
      ep[s].norm2_rows = st.fw.norm2 + off[s];
      ep[s].pnorm = gp->pnorm.as<double>() + o;
      ep[s].alpha = gp->alpha.as<double>() + o;
      ep[s].sig2 = gp->sig2;
      ep[s].power = gp->power;
      ep[s].normalized = 1;
      ep[s].kval = want_k ? st.kval + st.koff[s] : nullptr;  // the variance's k rows
      ep[s].cmat = cf + st.koff[s];
      ep[s].ldk = gp->ldk[s];
      ep[s].part = e_part + (size_t)off[s] * np;
      ep[s].nparts = np;
    }
    flare::sgemm_epi_batched(ctx->stream, p.S, g, ep, d, gp->Phat_f.as<float>());
  } else {
    for (int s = 0; s < p.S; s++) {
      const int rows = off[s + 1] - off[s], Ms = gp->m_by_type[s];
      if (!rows || !Ms) continue;
      sgemm_rm(ctx, false, true, rows, Ms, F, 1.0f, d + (size_t)off[s] * F, F,
               gp->Phat_f.as<float>() + (size_t)gp->m_off[s] * F, F, 0.0f,
               Kt + roff[s], Ms);
    }
  }
  ctx->mark_end();
  ctx->mark_begin("gp_epilogue");
  flare::GpEpiLp eps[flare::kMaxSpecies];
  int neps = 0, packed_rows = 0;
  for (int s = 0; s < p.S; s++) {
    const int rows = off[s + 1] - off[s], Ms = gp->m_by_type[s], o = gp->m_off[s];
    if (!rows) continue;
    if (!Ms) {
      // no sparse environments of this species: k, c and e are zero
      FLARE_HIP_CHECK(hipMemsetAsync(e_part + (size_t)off[s] * np, 0,
                                     (size_t)rows * np * sizeof(double), ctx->stream));
      FLARE_HIP_CHECK(hipMemsetAsync(st.kval + st.koff[s], 0,
                                     (size_t)rows * gp->ldk[s] * sizeof(double), ctx->stream));
      if (mixed) {
        FLARE_HIP_CHECK(hipMemsetAsync((_Float16 *)ch + coff[s], 0,
                                       (size_t)rows * gp->mp[s] * 2, ctx->stream));
        FLARE_HIP_CHECK(hipMemsetAsync((_Float16 *)cl + coff[s], 0,
                                       (size_t)rows * gp->mp[s] * 2, ctx->stream));
        FLARE_HIP_CHECK(hipMemsetAsync(c_is + off[s], 0, rows * sizeof(float),
                                       ctx->stream));
      } else {
        FLARE_HIP_CHECK(hipMemsetAsync(cf + st.koff[s], 0,
                                       (size_t)rows * gp->ldk[s] * sizeof(float), ctx->stream));
      }
      continue;
    }
    if (mcore) continue;  // done in the GEMM
    // kval and (FP32 fallback) cf rows of ldk, or FP16 hi/lo rows of mp (mixed)
    flare::GpEpiLp &q = eps[neps++];
    q.row0 = packed_rows;
    q.rows = rows;
    q.M = Ms;
    q.Mp = gp->mp[s];
    q.ldk = gp->ldk[s];
    q.K = Kt + roff[s];
    q.K2 = ksplit == 2 ? Kt + rtot + roff[s] : nullptr;
    q.norm2_rows = st.fw.norm2 + off[s];
    q.pnorm = gp->pnorm.as<double>() + o;
    q.alpha = gp->alpha.as<double>() + o;
    q.kval = want_k ? st.kval + st.koff[s] : nullptr;
    q.cf = mixed ? nullptr : cf + st.koff[s];
    q.ch = mixed ? (void *)((_Float16 *)ch + coff[s]) : nullptr;
    q.cl = mixed ? (void *)((_Float16 *)cl + coff[s]) : nullptr;
    q.c_is = mixed ? c_is + off[s] : nullptr;
    q.e_row = e_part + off[s];
    packed_rows += rows;
  }
  // every species in one launch, rows in registers; rows wider than 1,024
  // values one launch per species
  if (neps > 0 && !flare::launch_gp_epilogue_lp_rows(ctx->stream, neps, eps, gp->sig2,
                                                     gp->power)) {
    FLARE_REQUIRE(ksplit == 1, "gp epilogue: a K split needs the single-pass epilogue");
    for (int k = 0; k < neps; k++) {
      const flare::GpEpiLp &q = eps[k];
      flare::launch_gp_epilogue_lp(ctx->stream, q.rows, q.M, q.Mp, q.K, q.norm2_rows,
                                   q.pnorm, q.alpha, gp->sig2, gp->power, q.kval, q.ldk,
                                   q.cf, q.ch, q.cl, q.c_is, q.e_row);
    }
  }
  ctx->mark_end();
  ctx->mark_begin("gp_G");
  if (mixed) {
    flare::GemmProduct g[flare::kMaxSpecies];
    for (int s = 0; s < p.S; s++) {
      g[s] = flare::GemmProduct{off[s + 1] - off[s], F, gp->mp[s], F,
                                (int64_t)coff[s], gp->tpoff[s], (int64_t)off[s] * F};
      g[s].rs = c_is + off[s];
      g[s].cs = gp->PhatT_is.as<float>() + (size_t)s * F;
    }
    // wide rows on 256 x 352 tiles (F = 680: 58 vs 85 us; F = 1,200: 391
    // vs 436 us)
    if (wide_cols_fit(F, 256, 10))
      flare::gemm_f16x3_wide(ctx->stream, p.S, g, ch, cl, gp->PhatT_h.ptr, gp->PhatT_l.ptr, G);
    else
      flare::gemm_f16x3_batched(ctx->stream, p.S, g, ch, cl, gp->PhatT_h.ptr,
                                gp->PhatT_l.ptr, G);
  } else {
    // G = c P_hat on the matrix cores: c rows of ldk against P_hat^T rows;
    // 256 x 352 tiles where they pad no more than 128 x 128 tiles (F = 680:
    // 0.128 -> 0.111 ms at 32k atoms)
    flare::GemmProduct g[flare::kMaxSpecies];
    for (int s = 0; s < p.S; s++)
      g[s] = flare::GemmProduct{off[s + 1] - off[s], F, gp->ldk[s], F,
                                (int64_t)st.koff[s], gp->toff[s], (int64_t)off[s] * F};
    if (F % 4 == 0 && wide_cols_fit(F, 128, 0))
      flare::sgemm_wide(ctx->stream, p.S, g, cf, gp->PhatT_f.as<float>(), G);
    else
      flare::sgemm_batched(ctx->stream, p.S, g, cf, gp->PhatT_f.as<float>(), G);
  }
  ctx->mark_end();
  float *w = ctx->buf_t<float>("gpl.w", std::max<size_t>((size_t)n * F, 1));
  double *E_row = ctx->buf_t<double>("gpl.E", std::max(n, 1));
  ctx->mark_begin("gp_w");
  flare::launch_gp_w_lp(ctx->stream, n, F, G, e_part, np, d, st.fw.norm2, gp->sig2,
                        gp->power, w, E_row);
  ctx->mark_end();
  flare::WInput wi{0, 1, nullptr, nullptr, nullptr, w, st.fw.norm2, E_row};
  run_forces(ctx, p, nl, d_species, flare::kPrecF32, 1, st.fw, wi, d_mean_efs + 1,
             d_mean_efs, nullptr, d_mean_efs + 1 + 3 * (size_t)n, "gpl");
  return st;
}

static GPState gp_mean(flare_ctx *ctx, flare_sgp *gp, const flare_b2_params *params,
                       const flare_nlist *nl, const int *d_species, int form,
                       double *d_mean_efs, bool want_k) {
  REQUIRE(form >= FLARE_FORM_PLAIN_FP64 && form <= FLARE_FORM_ADJOINT_MIXED,
          "unknown form");
  flare::B2Dev p = make_b2(params, 0);
  check_species(nl, p.S);
  REQUIRE(p.S == gp->S && p.F == gp->F, "descriptor settings differ from the sparse set");
  REQUIRE(gp->have_alpha, "alpha not set");
  if (form == FLARE_FORM_ADJOINT_FP32 || form == FLARE_FORM_ADJOINT_MIXED)
    return gp_mean_lowp(ctx, gp, p, nl, d_species, form, d_mean_efs, want_k);
  const int n = nl->n, F = p.F;
  const bool mcore = gp_mcore(gp, false);
  sgp_build_pad(ctx, gp);
  GPState st;
  st.fw = run_forward(ctx, p, nl, d_species, flare::kPrecF64, "gp");
  const double *B = (const double *)st.fw.B;
  int off[flare::kMaxSpecies + 1];
  nl->row_offsets(p.S, off);
  size_t ktot = 0;
  st.koff.assign(p.S + 1, 0);
  // e partial sums: one per 64 sparse columns on the matrix cores, the
  // row's sum itself otherwise
  int np = 1;
  for (int s = 0; s < p.S; s++) {
    st.koff[s] = ktot;
    ktot += (size_t)(off[s + 1] - off[s]) * gp->ldk[s];
    if (mcore) np = std::max(np, (gp->m_by_type[s] + 63) / 64);
  }
  st.koff[p.S] = ktot;
  st.kval = ctx->buf_t<double>("gp.kval", std::max<size_t>(ktot, 1));
  double *cm = ctx->buf_t<double>("gp.C", std::max<size_t>(ktot, 1));
  double *e_part = ctx->buf_t<double>("gp.e", std::max<size_t>((size_t)n * np, 1));
  double *G = ctx->buf_t<double>("gp.G", std::max<size_t>((size_t)n * F, 1));
  // species without sparse environments: k, c, e and G rows are zero
  for (int s = 0; s < p.S; s++) {
    const int rows = off[s + 1] - off[s];
    if (!rows || gp->m_by_type[s]) continue;
    FLARE_HIP_CHECK(hipMemsetAsync(e_part + (size_t)off[s] * np, 0,
                                   (size_t)rows * np * sizeof(double), ctx->stream));
    FLARE_HIP_CHECK(hipMemsetAsync(st.kval + st.koff[s], 0,
                                   (size_t)rows * gp->ldk[s] * sizeof(double), ctx->stream));
    FLARE_HIP_CHECK(hipMemsetAsync(cm + st.koff[s], 0,
                                   (size_t)rows * gp->ldk[s] * sizeof(double), ctx->stream));
  }
  ctx->mark_begin("gp_K");
  if (mcore) {
    // K = B P^T with k, c = alpha t^(xi-1) and the partial energies in the
    // GEMM's epilogue (normalized_dot_product.cpp:115-210)
    flare::GemmProduct g[flare::kMaxSpecies];
    flare::DEpilogue ep[flare::kMaxSpecies];
    for (int s = 0; s < p.S; s++) {
      const int o = gp->m_off[s];
      g[s] = flare::GemmProduct{off[s + 1] - off[s], gp->m_by_type[s], F, 0,
                                (int64_t)off[s] * F, (int64_t)o * F, 0};
      ep[s].mode = flare::DEpilogue::kKernel;
      ep[s].norm2_rows = st.fw.norm2 + off[s];
      ep[s].pnorm = gp->pnorm.as<double>() + o;
      ep[s].alpha = gp->alpha.as<double>() + o;
      ep[s].sig2 = gp->sig2;
      ep[s].power = gp->power;
      ep[s].normalized = 0;
      ep[s].kval = want_k ? st.kval + st.koff[s] : nullptr;  // the variance's k rows
      ep[s].cmat = cm + st.koff[s];
      ep[s].ldk = gp->ldk[s];
      ep[s].part = e_part + (size_t)off[s] * np;
      ep[s].nparts = np;
    }
    flare::dgemm_epi_batched(ctx->stream, p.S, g, ep, B, gp->P.as<double>());
  } else {
    // rows of odd F: rocBLAS and the separate epilogue
    double *Kraw = ctx->buf_t<double>("gp.K", std::max<size_t>(ktot, 1));
    for (int s = 0; s < p.S; s++) {
      const int rows = off[s + 1] - off[s], Ms = gp->m_by_type[s], o = gp->m_off[s];
      if (!rows || !Ms) continue;
      gemm_rm(ctx, false, true, rows, Ms, F, 1.0, B + (size_t)off[s] * F, F,
              gp->P.as<double>() + (size_t)o * F, F, 0.0, Kraw + st.koff[s], gp->ldk[s]);
    }
    ctx->mark_end();
    ctx->mark_begin("gp_epilogue");
    for (int s = 0; s < p.S; s++) {
      const int rows = off[s + 1] - off[s], Ms = gp->m_by_type[s], o = gp->m_off[s];
      if (!rows || !Ms) continue;
      flare::launch_gp_epilogue(ctx->stream, rows, Ms, 0, Kraw + st.koff[s],
                                st.fw.norm2 + off[s], gp->pnorm.as<double>() + o,
                                gp->alpha.as<double>() + o, gp->sig2, gp->power,
                                st.kval + st.koff[s], cm + st.koff[s], e_part + off[s],
                                gp->ldk[s]);
    }
  }
  ctx->mark_end();
  ctx->mark_begin("gp_G");
  {
    // G = c P_hat: c rows of ldk against the padded P_hat^T rows
    flare::GemmProduct g[flare::kMaxSpecies];
    for (int s = 0; s < p.S; s++)
      g[s] = flare::GemmProduct{off[s + 1] - off[s], F, gp->ldk[s], F,
                                (int64_t)st.koff[s], gp->toff[s], (int64_t)off[s] * F};
    flare::dgemm_batched(ctx->stream, p.S, g, cm, gp->PhatT_d.as<double>(), G);
  }
  ctx->mark_end();
  double *w = ctx->buf_t<double>("gp.w", std::max<size_t>((size_t)n * F, 1));
  double *E_row = ctx->buf_t<double>("gp.E", std::max(n, 1));
  ctx->mark_begin("gp_w");
  flare::launch_gp_w(ctx->stream, n, F, G, e_part, np, B, st.fw.norm2, gp->sig2,
                     gp->power, w, E_row);
  ctx->mark_end();
  const int mode = form == FLARE_FORM_PLAIN_FP64 ? 0 : 1;
  flare::WInput wi{0, 0, nullptr, nullptr, nullptr, w, st.fw.norm2, E_row};
  run_forces(ctx, p, nl, d_species, flare::kPrecF64, mode, st.fw, wi,
             d_mean_efs + 1, d_mean_efs, nullptr, d_mean_efs + 1 + 3 * (size_t)n,
             "gp");
  return st;
}

int flare_sgp_predict_mean(flare_ctx *ctx, flare_sgp *gp,
                           const flare_b2_params *p, const flare_nlist *nl,
                           const int *d_species, int form, double *d_mean_efs) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_mean_efs, "null argument");
    ctx->use();
    gp_mean(ctx, gp, p, nl, d_species, form, d_mean_efs, false);
    ctx->collect();
    ctx->sync();
  });
}

int flare_sgp_predict_local_uncertainties(flare_ctx *ctx, flare_sgp *gp,
                                          const flare_b2_params *params,
                                          const flare_nlist *nl,
                                          const int *d_species, int form,
                                          double *d_mean_efs,
                                          double *d_variances) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_mean_efs && d_variances, "null argument");
    ctx->use();
    GPState st = gp_mean(ctx, gp, params, nl, d_species, form, d_mean_efs, true);
    const int n = nl->n, F = gp->F;
    int off[flare::kMaxSpecies + 1];
    nl->row_offsets(gp->S, off);
    double *kself = ctx->buf_t<double>("gp.kself", std::max(n, 1));
    ctx->mark_begin("gp_kself");
    if (st.lowp)
      flare::launch_kself_lp(ctx->stream, n, F, (const float *)st.fw.B, st.fw.norm2,
                             gp->sig2, gp->power, kself);
    else
      flare::launch_kself(ctx->stream, n, F, (const double *)st.fw.B, st.fw.norm2,
                          gp->sig2, gp->power, kself);
    ctx->mark_end();
    // |L^-1 k_i|^2 (sparse_gp.cpp:93-133) on the matrix cores: k rows of
    // ldk against the species' L^-1 rows, squares summed in the epilogue
    // per 64 columns, the K loop of each tile cut at its last column (L^-1
    // is lower triangular); FP64 in every form
    int np = 1;
    for (int s = 0; s < gp->S; s++) np = std::max(np, (gp->m_by_type[s] + 63) / 64);
    double *vpart = ctx->buf_t<double>("gp.vpart", std::max<size_t>((size_t)n * np, 1));
    flare::GemmProduct g[flare::kMaxSpecies];
    flare::DEpilogue ep[flare::kMaxSpecies];
    for (int s = 0; s < gp->S; s++) {
      const int rows = off[s + 1] - off[s], Ms = gp->m_by_type[s];
      if (rows && !Ms)
        FLARE_HIP_CHECK(hipMemsetAsync(vpart + (size_t)off[s] * np, 0,
                                       (size_t)rows * np * sizeof(double), ctx->stream));
      g[s] = flare::GemmProduct{rows, Ms, gp->ldk[s], 0, (int64_t)st.koff[s],
                                gp->loff[s], 0};
      ep[s].mode = flare::DEpilogue::kSquares;
      ep[s].part = vpart + (size_t)off[s] * np;
      ep[s].nparts = np;
      ep[s].tri = 1;
    }
    ctx->mark_begin("gp_Q");
    flare::dgemm_epi_batched(ctx->stream, gp->S, g, ep, st.kval, gp->Linv_rm.as<double>());
    ctx->mark_end();
    ctx->mark_begin("gp_var");
    flare::launch_var_from_parts(ctx->stream, n, np, vpart, kself, d_variances);
    ctx->mark_end();
    ctx->collect();
    ctx->sync();
  });
}

int flare_envs_envs(flare_ctx *ctx, int n_species, int n_descriptors,
                    const int *n1_by_type, const double *d1,
                    const int *n2_by_type, const double *d2, double sigma,
                    double power, double *d_out) {
  return guarded([&] {
    REQUIRE(ctx && n1_by_type && n2_by_type && d_out, "null argument");
    REQUIRE(n_species >= 1 && n_species <= flare::kMaxSpecies, "n_species must be in [1, 8]");
    ctx->use();
    const int F = n_descriptors;
    int n1 = 0, n2 = 0;
    std::vector<int> o1(n_species + 1, 0), o2(n_species + 1, 0);
    for (int s = 0; s < n_species; s++) {
      o1[s + 1] = o1[s] + n1_by_type[s];
      o2[s + 1] = o2[s] + n2_by_type[s];
    }
    n1 = o1[n_species];
    n2 = o2[n_species];
    if (!n1 || !n2) return;
    double *nrm1 = ctx->buf_t<double>("ee.n1", n1);
    double *nrm2 = ctx->buf_t<double>("ee.n2", n2);
    double *sq1 = ctx->buf_t<double>("ee.s1", n1);
    flare::launch_row_norms(ctx->stream, n1, F, d1, nrm1, nullptr);
    flare::launch_row_norms(ctx->stream, n2, F, d2, nrm2, nullptr);
    flare::launch_square(ctx->stream, n1, nrm1, sq1);
    FLARE_HIP_CHECK(hipMemsetAsync(d_out, 0, (size_t)n1 * n2 * sizeof(double), ctx->stream));
    double *tmp = ctx->buf_t<double>("ee.tmp", (size_t)n1 * n2);
    for (int s = 0; s < n_species; s++) {
      int r = n1_by_type[s], m = n2_by_type[s];
      if (!r || !m) continue;
      double *blk = tmp;  // r x m contiguous
      gemm_rm(ctx, false, true, r, m, F, 1.0, d1 + (size_t)o1[s] * F, F,
              d2 + (size_t)o2[s] * F, F, 0.0, blk, m);
      double *kv = ctx->buf_t<double>("ee.kv", (size_t)r * m);
      flare::launch_gp_epilogue(ctx->stream, r, m, 0, blk, sq1 + o1[s], nrm2 + o2[s],
                                nullptr, sigma * sigma, power, kv, nullptr, nullptr);
      flare::launch_copy_block(ctx->stream, r, m, kv, m, d_out + (size_t)o1[s] * n2 + o2[s], n2);
    }
    ctx->sync();
  });
}

int flare_sgp_mapping_coefficients(flare_ctx *ctx, const flare_sgp *gp,
                                   double *d_out) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_out, "null argument");
    REQUIRE(gp->power == 1.0 || gp->power == 2.0,
            "mapping coefficients exist for power 1 and 2 only");
    ctx->use();
    const int F = gp->F;
    const int64_t bs = gp->power == 1.0 ? (int64_t)F : (int64_t)F * (F + 1) / 2;
    double *full = ctx->buf_t<double>("map.full", (size_t)F * F);
    for (int s = 0; s < gp->S; s++) {
      int Ms = gp->m_by_type[s], o = gp->m_off[s];
      double *dst = d_out + (size_t)s * bs;
      if (!Ms) {
        FLARE_HIP_CHECK(hipMemsetAsync(dst, 0, bs * sizeof(double), ctx->stream));
        continue;
      }
      const double *Ph = gp->Phat.as<double>() + (size_t)o * F;
      double *Y = ctx->buf_t<double>("map.Y", (size_t)Ms * F);
      flare::launch_scale_rows(ctx->stream, Ms, F, Ph, gp->alpha.as<double>() + o, Y);
      if (gp->power == 2.0) {
        gemm_rm(ctx, true, false, F, F, Ms, gp->sig2, Ph, F, Y, F, 0.0, full, F);
        flare::launch_pack_beta(ctx->stream, F, 2, full, dst);
      } else {
        flare::launch_colsum(ctx->stream, Ms, F, Y, dst);
        flare::launch_scale(ctx->stream, F, gp->sig2, dst);
      }
    }
    ctx->sync();
  });
}


// ---- compute flare/std/atom ------------------------------------------------

int flare_std_read(const char *path1, const char *path2, flare_std_info *info,
                   double *cutoffs, double *values1, double *values2) {
  return guarded([&] {
    REQUIRE(path1 && info, "null argument");
    flare_std sd;
    std::vector<double> v1, v2;
    read_std_files(path1, path2, sd, values1 ? &v1 : nullptr,
                   values2 && path2 ? &v2 : nullptr);
    fill_std_info(sd, info);
    if (cutoffs) std::copy(sd.cutoffs.begin(), sd.cutoffs.end(), cutoffs);
    if (values1) std::copy(v1.begin(), v1.end(), values1);
    if (values2 && path2) std::copy(v2.begin(), v2.end(), values2);
  });
}

int flare_std_load(flare_ctx *ctx, const char *path1, const char *path2,
                   flare_std **out) {
  return guarded([&] {
    REQUIRE(ctx && path1 && out, "null argument");
    *out = nullptr;
    ctx->use();
    std::unique_ptr<flare_std> sd(new flare_std);
    sd->ctx = ctx;
    std::vector<double> v1, v2;
    read_std_files(path1, path2, *sd, &v1, &v2);
    upload_std(*sd, v1.data(), v2.data());
    *out = sd.release();
  });
}

int flare_std_create(flare_ctx *ctx, const flare_std_info *info,
                     const double *cutoffs, const double *values1,
                     const double *values2, flare_std **out) {
  return guarded([&] {
    REQUIRE(ctx && info && cutoffs && out, "null argument");
    *out = nullptr;
    ctx->use();
    std::unique_ptr<flare_std> sd(new flare_std);
    sd->ctx = ctx;
    sd->use_map = info->use_map ? 1 : 0;
    sd->power = sd->use_map ? 2 : info->power;
    sd->n_hyps = info->n_hyps;
    REQUIRE(info->n_hyps >= 1, "at least one hyperparameter (sigma) is needed");
    for (int k = 0; k < 4; k++) sd->hyps[k] = k < info->n_hyps ? info->hyps[k] : 0.0;
    sd->S = info->n_species;
    sd->N = info->n_max;
    sd->L = info->l_max;
    sd->cutoff_name = info->cutoff_function == FLARE_CUTOFF_COSINE ? "cosine" : "quadratic";
    REQUIRE(info->cutoff_function == FLARE_CUTOFF_QUADRATIC ||
                info->cutoff_function == FLARE_CUTOFF_COSINE,
            "unknown cutoff function");
    set_basis(*sd);
    sd->cutoffs.assign(cutoffs, cutoffs + sd->S * sd->S);
    sd->cutoff = -1;
    for (double v : sd->cutoffs) sd->cutoff = std::max(sd->cutoff, v);
    if (!sd->use_map) {
      sd->M = info->n_clusters;
      REQUIRE(sd->M >= 0, "negative sparse count");
      sd->m_by_type.assign(sd->S, 0);
      sd->m_off.assign(sd->S + 1, 0);
      for (int s = 0; s < sd->S; s++) {
        REQUIRE(info->n_clusters_by_type[s] >= 0, "negative cluster count");
        sd->m_by_type[s] = info->n_clusters_by_type[s];
        sd->m_off[s + 1] = sd->m_off[s] + sd->m_by_type[s];
      }
      REQUIRE(sd->m_off[sd->S] == sd->M, "cluster counts do not add up to n_clusters");
      REQUIRE(sd->M == 0 || (values1 && values2), "null L^-1 or sparse rows");
    } else {
      REQUIRE(values1, "null variance map");
    }
    upload_std(*sd, values1, values2);
    *out = sd.release();
  });
}

int flare_std_destroy(flare_std *sd) {
  return guarded([&] {
    if (!sd) return;
    sd->ctx->use();
    (void)hipStreamSynchronize(sd->ctx->stream);
    delete sd;
  });
}

int flare_std_get_info(const flare_std *sd, flare_std_info *out) {
  return guarded([&] {
    REQUIRE(sd && out, "null argument");
    fill_std_info(*sd, out);
  });
}

int flare_std_compute(flare_ctx *ctx, flare_std *sd, const flare_nlist *nl,
                      const int *d_species, double *d_stds) {
  return guarded([&] {
    REQUIRE(ctx && sd && nl && d_stds, "null argument");
    flare_b2_params bp;
    bp.n_species = sd->S;
    bp.n_max = sd->N;
    bp.l_max = sd->L;
    bp.radial_basis = FLARE_RADIAL_CHEBYSHEV;
    bp.cutoff_function = sd->cutoff_kind;
    bp.radial_r1 = 0.0;  // radial_hyps = {0, cutoff} (compute_flare_std_atom.cpp:414-417)
    bp.cutoffs = sd->cutoffs.data();
    flare::B2Dev p = make_b2(&bp, 1);  // rsq < rc^2, as the pair style
    check_species(nl, p.S);
    ctx->use();
    const int n = nl->n, F = p.F;
    Forward fw = run_forward(ctx, p, nl, d_species, flare::kPrecF64, "std");
    const double *B = (const double *)fw.B;
    int off[flare::kMaxSpecies + 1];
    nl->row_offsets(p.S, off);
    const int *aor = nl->b.atom_of_row.as<int>();
    const double sig2 = sd->hyps[0] * sd->hyps[0];
    if (sd->use_map) {
      double *Z = ctx->buf_t<double>("std.Z", (size_t)n * F);
      ctx->mark_begin("gemm");
      for (int s = 0; s < p.S; s++) {
        int rows = off[s + 1] - off[s];
        gemm_rm(ctx, false, true, rows, F, F, 1.0, B + (size_t)off[s] * F, F,
                sd->beta.as<double>() + (size_t)s * F * F, F, 0.0,
                Z + (size_t)off[s] * F, F);
      }
      ctx->mark_end();
      ctx->mark_begin("std");
      flare::launch_std_map(ctx->stream, n, F, B, Z, fw.norm2, sig2, aor, d_stds);
      ctx->mark_end();
    } else {
      const int M = sd->M;
      size_t tmax = 1, qmax = 1;
      for (int s = 0; s < p.S; s++) {
        size_t rows = off[s + 1] - off[s];
        tmax = std::max(tmax, rows * sd->m_by_type[s]);
        qmax = std::max(qmax, rows * (size_t)(M - sd->m_off[s]));
      }
      double *T = ctx->buf_t<double>("std.T", tmax);
      double *Q = ctx->buf_t<double>("std.Q", qmax);
      for (int s = 0; s < p.S; s++) {
        const int rows = off[s + 1] - off[s], Ms = sd->m_by_type[s], o = sd->m_off[s];
        if (!rows) continue;
        const double *Bs = B + (size_t)off[s] * F;
        if (!Ms) {  // no sparse environments of this species: k = 0
          flare::launch_std_linv(ctx->stream, rows, 0, Q, fw.norm2 + off[s], sig2,
                                 aor + off[s], d_stds);
          continue;
        }
        ctx->mark_begin("gemm");
        // t = P_hat_s B (then / |B|, ^power), q = L^-1[:, species block] k
        gemm_rm(ctx, false, true, rows, Ms, F, 1.0, Bs, F,
                sd->Phat.as<double>() + (size_t)o * F, F, 0.0, T, Ms);
        flare::launch_std_kvec(ctx->stream, rows, Ms, T, fw.norm2 + off[s],
                               (double)sd->power);
        gemm_rm(ctx, false, true, rows, M - o, Ms, 1.0, T, Ms,
                sd->Linv.as<double>() + (size_t)o * M + o, M, 0.0, Q, M - o);
        ctx->mark_end();
        ctx->mark_begin("std");
        flare::launch_std_linv(ctx->stream, rows, M - o, Q, fw.norm2 + off[s], sig2,
                               aor + off[s], d_stds);
        ctx->mark_end();
      }
    }
    ctx->finish();
  });
}

int flare_sgp_varmap_coefficients(flare_ctx *ctx, const flare_sgp *gp,
                                  double *d_out) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_out, "null argument");
    if (gp->power != 1.0)
      throw Error{FLARE_EINVAL,
                  "Mapping coefficients of the normalized dot product kernel are "
                  "implemented for power 1 only."};
    ctx->use();
    const int F = gp->F, M = gp->M;
    const size_t FF = (size_t)F * F;
    // first sparse norm of each species decides the sigma^2 I term
    std::vector<double> pn(std::max(M, 1));
    if (M)
      FLARE_HIP_CHECK(hipMemcpyAsync(pn.data(), gp->pnorm.ptr, M * sizeof(double),
                                     hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    for (int s = 0; s < gp->S; s++) {
      const int Ms = gp->m_by_type[s], o = gp->m_off[s];
      double *dst = d_out + (size_t)s * FF;
      FLARE_HIP_CHECK(hipMemsetAsync(dst, 0, FF * sizeof(double), ctx->stream));
      if (!Ms) continue;
      // Y = L^-1_s P_hat_s, beta = -sigma^4 Y^T Y (K_uu^-1 = L^-T L^-1,
      // sparse_gp.cpp:588; empty rows of P_hat are zero)
      double *Y = ctx->buf_t<double>("varmap.Y", (size_t)Ms * F);
      gemm_rm(ctx, true, false, Ms, F, Ms, 1.0,
              gp->Linv_cm.as<double>() + (size_t)o * M + o, M,
              gp->Phat.as<double>() + (size_t)o * F, F, 0.0, Y, F);
      gemm_rm(ctx, true, false, F, F, Ms, -gp->sig2 * gp->sig2, Y, F, Y, F, 0.0,
              dst, F);
      if (!(pn[o] < 1e-8)) flare::launch_add_diag_rm(ctx->stream, F, dst, gp->sig2);
    }
    ctx->sync();
  });
}

int flare_write_varmap_coefficients(const char *path, const char *contributor,
                                    int n_hyps, const double *hyps,
                                    int n_species, int n_max, int l_max,
                                    const char *radial_basis,
                                    const char *cutoff_function,
                                    const double *cutoffs, int64_t beta_size,
                                    const double *beta) {
  return guarded([&] {
    REQUIRE(path && cutoffs && beta && (hyps || n_hyps == 0), "null argument");
    FILE *f = open_out(path);
    // sparse_gp.cpp:1208-1281
    std::fprintf(f, "DATE: %s CONTRIBUTOR: %s\n", date_now().c_str(),
                 contributor ? contributor : "");
    put_hyps(f, n_hyps, hyps);
    put_b2_settings(f, radial_basis, n_species, n_max, l_max, (long long)beta_size,
                    cutoff_function, cutoffs);
    int count = 0;
    for (int s = 0; s < n_species; s++) {
      if (count != 0) {
        std::fprintf(f, "\n");
        count = 0;
      }
      for (int64_t j = 0; j < beta_size; j++) {
        double v = beta[(int64_t)s * beta_size + j];
        std::fprintf(f, v > 0 ? "  %.16e" : " %.16e", v);
        if (++count == 5) {
          count = 0;
          std::fprintf(f, "\n");
        }
      }
    }
    close_out(f, path);
  });
}

int flare_write_L_inverse(const char *path, const char *contributor,
                          double power, int n_hyps, const double *hyps,
                          int n_species, int n_max, int l_max,
                          const char *radial_basis,
                          const char *cutoff_function, const double *cutoffs,
                          int n_clusters, const double *L_inv) {
  return guarded([&] {
    REQUIRE(path && cutoffs && (L_inv || n_clusters == 0) && (hyps || n_hyps == 0),
            "null argument");
    FILE *f = open_out(path);
    // sparse_gp.cpp:1283-1339
    std::fprintf(f, "L_inverse_block file DATE: %s CONTRIBUTOR: %s\n",
                 date_now().c_str(), contributor ? contributor : "");
    std::fprintf(f, "%.0f\n", power);  // NormalizedDotProduct::write_info
    put_hyps(f, n_hyps, hyps);
    put_b2_settings(f, radial_basis, n_species, n_max, l_max, 1, cutoff_function,
                    cutoffs);
    std::fprintf(f, "%d\n", n_clusters);
    int count = 1;
    for (int64_t j = 0; j < n_clusters; j++)
      for (int64_t k = 0; k <= j; k++) put_value(f, L_inv[j * n_clusters + k], count);
    if (count % 5 != 0) std::fprintf(f, "\n");
    close_out(f, path);
  });
}

int flare_write_sparse_descriptors(const char *path, const char *contributor,
                                   int n_species, int n_descriptors,
                                   const int *n_clusters_by_type,
                                   const double *descriptors,
                                   const double *norms) {
  return guarded([&] {
    REQUIRE(path && n_clusters_by_type, "null argument");
    int M = 0;
    for (int s = 0; s < n_species; s++) M += n_clusters_by_type[s];
    REQUIRE(M == 0 || (descriptors && norms), "null argument");
    FILE *f = open_out(path);
    // sparse_gp.cpp:1341-1396, one kernel
    std::fprintf(f, "sparse_descriptors file DATE: %s CONTRIBUTOR: %s\n",
                 date_now().c_str(), contributor ? contributor : "");
    std::fprintf(f, "%d\n", 1);
    std::fprintf(f, "%d %d %d\n", 0, M, n_species);
    int count = 1, row = 0;
    for (int s = 0; s < n_species; s++) {
      std::fprintf(f, "%d\n", n_clusters_by_type[s]);
      for (int j = 0; j < n_clusters_by_type[s]; j++, row++) {
        const double nrm = norms[row];
        for (int k = 0; k < n_descriptors; k++) {
          double v = nrm < 1e-8 ? 0.0 : descriptors[(size_t)row * n_descriptors + k] / nrm;
          put_value(f, v, count);
        }
      }
      if (count % 5 != 1) std::fprintf(f, "\n");
    }
    close_out(f, path);
  });
}


}  // extern "C"

// ---- explicit kernel matrices: SOR / DTC ------------------------------------

// DescriptorValues on the device (b2.cpp:76-179 layout, all species' blocks
// concatenated): rows, per-row pair ranges and the pair tensors.
struct DVDev {
  Forward fw;
  int64_t Q = 0;
  std::vector<int64_t> q0h;  // per row
  std::vector<int> cnth;     // per row
  std::vector<int> atom_of_row;
  int off[flare::kMaxSpecies + 1];
  int64_t qoff[flare::kMaxSpecies + 1];
  int64_t *q0 = nullptr;
  int *cnt = nullptr, *ni = nullptr, *pair_rowg = nullptr;
  double *dB = nullptr, *fd = nullptr, *co = nullptr;
};

static DVDev dv_build(flare_ctx *ctx, const flare::B2Dev &p, const flare_nlist *nl,
                      const int *d_species, const std::string &tag) {
  DVDev dv;
  const int n = nl->n;
  dv.fw = run_forward(ctx, p, nl, d_species, flare::kPrecF64, tag);
  B2Pairs r = b2_pairs(ctx, p, nl, d_species);
  nl->row_offsets(p.S, dv.off);
  dv.Q = r.total;
  dv.q0h = r.q0;
  dv.atom_of_row = r.atom_of_row;
  dv.cnth.resize(n);
  for (int row = 0; row < n; row++) dv.cnth[row] = r.count[r.atom_of_row[row]];
  // species pair offsets from the row ranges
  {
    int64_t q = 0;
    for (int s = 0; s < p.S; s++) {
      dv.qoff[s] = q;
      for (int row = dv.off[s]; row < dv.off[s + 1]; row++) q += dv.cnth[row];
    }
    dv.qoff[p.S] = q;
  }
  const size_t Qa = (size_t)std::max<int64_t>(dv.Q, 1);
  dv.q0 = ctx->buf_t<int64_t>(tag + ".q0", n);
  dv.cnt = ctx->buf_t<int>(tag + ".cnt", n);
  if (n) {
    FLARE_HIP_CHECK(hipMemcpyAsync(dv.q0, dv.q0h.data(), n * sizeof(int64_t),
                                   hipMemcpyHostToDevice, ctx->stream));
    FLARE_HIP_CHECK(hipMemcpyAsync(dv.cnt, dv.cnth.data(), n * sizeof(int),
                                   hipMemcpyHostToDevice, ctx->stream));
  }
  std::vector<int> prow(Qa, 0);
  for (int row = 0; row < n; row++)
    for (int k = 0; k < dv.cnth[row]; k++) prow[dv.q0h[row] + k] = row;
  dv.pair_rowg = ctx->buf_t<int>(tag + ".prow", Qa);
  FLARE_HIP_CHECK(hipMemcpyAsync(dv.pair_rowg, prow.data(), Qa * sizeof(int),
                                 hipMemcpyHostToDevice, ctx->stream));
  dv.dB = ctx->buf_t<double>(tag + ".dB", Qa * 3 * p.F);
  dv.fd = ctx->buf_t<double>(tag + ".fd", Qa * 3);
  dv.co = ctx->buf_t<double>(tag + ".co", Qa * 3);
  dv.ni = ctx->buf_t<int>(tag + ".ni", Qa);
  if (dv.Q > 0) {
    int64_t *fpair = ctx->buf_t<int64_t>(tag + ".fpair", dv.Q);
    flare::launch_b2_fill_fpair(ctx->stream, p, nl->dev(), d_species,
                                nl->b.row_of.as<int>(), dv.q0, fpair);
    flare::launch_b2_dervs(ctx->stream, p, nl->dev(), d_species, nl->b.row_of.as<int>(),
                           ctx->ynorm_d.as<double>(), ctx->b2_index(p.S, p.N, p.L),
                           dv.Q, fpair, (const double *)dv.fw.A,
                           (const double *)dv.fw.B, dv.dB, dv.fd, dv.co, dv.ni);
  }
  // keep the host vectors alive until the uploads are done
  ctx->sync();
  return dv;
}

// envs_struc of the sparse set against the structure: d_K M x (1+3n+6)
// (normalized_dot_product.cpp:115-210).
static void sgp_envs_struc(flare_ctx *ctx, flare_sgp *gp, const flare::B2Dev &p,
                           const flare_nlist *nl, const DVDev &dv, double *d_K) {
  const int n = nl->n, F = p.F, M = gp->M;
  const int64_t n_out = 1 + 3 * (int64_t)n + 6;
  FLARE_HIP_CHECK(hipMemsetAsync(d_K, 0, std::max<int64_t>(M * n_out, 1) * sizeof(double),
                                 ctx->stream));
  const double *B = (const double *)dv.fw.B;
  const double vol_inv = 1.0 / nl->wh.volume;
  std::vector<int> own0(std::max(n, 1)), ownc(std::max(n, 1));
  for (int s = 0; s < p.S; s++) {
    const int Ms = gp->m_by_type[s], o = gp->m_off[s];
    const int ns = dv.off[s + 1] - dv.off[s];
    const int64_t Qs = dv.qoff[s + 1] - dv.qoff[s];
    if (!Ms || !ns) continue;
    const double *Ps = gp->P.as<double>() + (size_t)o * F;
    double *DOT = ctx->buf_t<double>("es.dot", (size_t)Ms * ns);
    gemm_rm(ctx, false, true, Ms, ns, F, 1.0, Ps, F, B + (size_t)dv.off[s] * F, F, 0.0,
            DOT, ns);
    double *FD = ctx->buf_t<double>("es.fd", (size_t)Ms * 3 * std::max<int64_t>(Qs, 1));
    if (Qs > 0)
      gemm_rm(ctx, false, true, Ms, (int)(3 * Qs), F, 1.0, Ps, F,
              dv.dB + (size_t)dv.qoff[s] * 3 * F, F, 0.0, FD, (int)(3 * Qs));
    // pair -> row within the species; own pair ranges per atom
    std::vector<int> prow(std::max<int64_t>(Qs, 1));
    std::fill(ownc.begin(), ownc.end(), 0);
    std::fill(own0.begin(), own0.end(), 0);
    for (int row = dv.off[s]; row < dv.off[s + 1]; row++) {
      const int64_t b = dv.q0h[row] - dv.qoff[s];
      for (int k = 0; k < dv.cnth[row]; k++) prow[b + k] = row - dv.off[s];
      own0[dv.atom_of_row[row]] = (int)b;
      ownc[dv.atom_of_row[row]] = dv.cnth[row];
    }
    int *d_prow = ctx->buf_t<int>("es.prow", prow.size());
    int *d_own0 = ctx->buf_t<int>("es.own0", own0.size());
    int *d_ownc = ctx->buf_t<int>("es.ownc", ownc.size());
    FLARE_HIP_CHECK(hipMemcpyAsync(d_prow, prow.data(), prow.size() * sizeof(int),
                                   hipMemcpyHostToDevice, ctx->stream));
    FLARE_HIP_CHECK(hipMemcpyAsync(d_own0, own0.data(), own0.size() * sizeof(int),
                                   hipMemcpyHostToDevice, ctx->stream));
    FLARE_HIP_CHECK(hipMemcpyAsync(d_ownc, ownc.data(), ownc.size() * sizeof(int),
                                   hipMemcpyHostToDevice, ctx->stream));
    double *Ks = d_K + (size_t)o * n_out;
    flare::launch_envs_struc_epilogue(ctx->stream, Ms, ns, Qs, DOT, FD,
                                      gp->pnorm.as<double>() + o, dv.fw.norm2 + dv.off[s],
                                      dv.fd + 3 * dv.qoff[s], dv.co + 3 * dv.qoff[s],
                                      d_prow, gp->sig2, gp->power, vol_inv, Ks, n_out, n);
    int *in_cum = ctx->buf_t<int>("es.in_cum", (size_t)n + 1);
    int *in_pairs = ctx->buf_t<int>("es.in_pairs", std::max<int64_t>(Qs, 1));
    flare::index_by_target(ctx->stream, *ctx->ws_buf("es.w1"), *ctx->ws_buf("es.w2"),
                           *ctx->ws_buf("es.w3"), dv.ni + dv.qoff[s], Qs, n, in_cum,
                           in_pairs);
    flare::launch_envs_struc_gather(ctx->stream, Ms, n, Qs, FD, d_own0, d_ownc, in_cum,
                                    in_pairs, Ks, n_out);
    ctx->sync();  // host vectors are reused by the next species
  }
}

// self_kernel_struc (normalized_dot_product.cpp:446-578) into d_out (1+3n+6)
static void self_kernel(flare_ctx *ctx, flare_sgp *gp, const flare::B2Dev &p,
                        const flare_nlist *nl, const DVDev &dv, double *d_out) {
  const int n = nl->n, F = p.F;
  const int64_t n_out = 1 + 3 * (int64_t)n + 6;
  FLARE_HIP_CHECK(hipMemsetAsync(d_out, 0, n_out * sizeof(double), ctx->stream));
  const double *B = (const double *)dv.fw.B;
  const double vol_inv = 1.0 / nl->wh.volume;
  double *sums = ctx->buf_t<double>("sk.sums", (size_t)p.S * 7);
  FLARE_HIP_CHECK(hipMemsetAsync(sums, 0, (size_t)p.S * 7 * sizeof(double), ctx->stream));
  const int combos[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  for (int s = 0; s < p.S; s++) {
    const int ns = dv.off[s + 1] - dv.off[s];
    if (!ns) continue;
    const double *Bs = B + (size_t)dv.off[s] * F;
    const double *n2 = dv.fw.norm2 + dv.off[s];
    double *T = ctx->buf_t<double>("sk.T", (size_t)ns * ns);
    double *U = ctx->buf_t<double>("sk.U", (size_t)ns * ns);
    double *V = ctx->buf_t<double>("sk.V", (size_t)ns * ns);
    double *G = ctx->buf_t<double>("sk.G", (size_t)ns * F);
    double *g = ctx->buf_t<double>("sk.g", ns);
    double *rows = ctx->buf_t<double>("sk.rows", ns);
    gemm_rm(ctx, false, true, ns, ns, F, 1.0, Bs, F, Bs, F, 0.0, T, ns);
    flare::launch_self_rows(ctx->stream, ns, T, nullptr, nullptr, nullptr, n2, gp->sig2,
                            gp->power, 1, rows);
    flare::launch_colsum(ctx->stream, ns, 1, rows, sums + s * 7);
    for (int q = 0; q < 6; q++) {
      flare::launch_self_g(ctx->stream, ns, F, combos[q][0], combos[q][1],
                           dv.q0 + dv.off[s], dv.cnt + dv.off[s], dv.dB, dv.fd, dv.co, G,
                           g);
      gemm_rm(ctx, false, true, ns, ns, F, 1.0, G, F, Bs, F, 0.0, U, ns);
      gemm_rm(ctx, false, true, ns, ns, F, 1.0, G, F, G, F, 0.0, V, ns);
      flare::launch_self_rows(ctx->stream, ns, T, U, V, g, n2, gp->sig2, gp->power, 0,
                              rows);
      flare::launch_colsum(ctx->stream, ns, 1, rows, sums + s * 7 + 1 + q);
    }
  }
  // force entries
  int *in_cum = ctx->buf_t<int>("sk.in_cum", (size_t)n + 1);
  int *in_pairs = ctx->buf_t<int>("sk.in_pairs", std::max<int64_t>(dv.Q, 1));
  flare::index_by_target(ctx->stream, *ctx->ws_buf("sk.w1"), *ctx->ws_buf("sk.w2"),
                         *ctx->ws_buf("sk.w3"), dv.ni, dv.Q, n, in_cum, in_pairs);
  int *ovf = ctx->buf_t<int>("sk.ovf", 1);
  FLARE_HIP_CHECK(hipMemsetAsync(ovf, 0, sizeof(int), ctx->stream));
  flare::launch_self_force(ctx->stream, n, F, nl->b.row_of.as<int>(), dv.q0, dv.cnt,
                           in_cum, in_pairs, dv.pair_rowg,
                           nl->b.species_of_row.as<int>(), B, dv.fw.norm2, dv.dB, dv.fd,
                           gp->sig2, gp->power, d_out, ovf);
  std::vector<double> hs((size_t)p.S * 7);
  int hovf = 0;
  FLARE_HIP_CHECK(hipMemcpyAsync(hs.data(), sums, hs.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
  FLARE_HIP_CHECK(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost,
                                 ctx->stream));
  ctx->sync();
  REQUIRE(hovf == 0, "an atom touches more pairs than the self-kernel workgroup holds (" +
                         std::to_string(flare::kSelfMaxPairs) + ")");
  double tail[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int s = 0; s < p.S; s++)
    for (int q = 0; q < 7; q++) tail[q] += hs[(size_t)s * 7 + q];
  for (int q = 1; q < 7; q++) tail[q] *= vol_inv * vol_inv;
  FLARE_HIP_CHECK(hipMemcpyAsync(d_out, tail, sizeof(double), hipMemcpyHostToDevice,
                                 ctx->stream));
  FLARE_HIP_CHECK(hipMemcpyAsync(d_out + 1 + 3 * (int64_t)n, tail + 1, 6 * sizeof(double),
                                 hipMemcpyHostToDevice, ctx->stream));
  ctx->sync();
}

enum { kVarSOR = 0, kVarDTC = 1 };

static void sgp_predict_var(flare_ctx *ctx, flare_sgp *gp, const flare_b2_params *params,
                            const flare_nlist *nl, const int *d_species,
                            double *d_mean_efs, double *d_var, int kind) {
  flare::B2Dev p = make_b2(params, 0);
  check_species(nl, p.S);
  REQUIRE(p.S == gp->S && p.F == gp->F, "descriptor settings differ from the sparse set");
  REQUIRE(gp->have_alpha, "alpha not set");
  REQUIRE(gp->have_R_inv, "R_inv not set (flare_sgp_set_R_inv)");
  const int n = nl->n, M = gp->M;
  const int64_t n_out = 1 + 3 * (int64_t)n + 6;
  DVDev dv = dv_build(ctx, p, nl, d_species, "sv");
  double *K = ctx->buf_t<double>("sv.K", (size_t)std::max(M, 1) * n_out);
  sgp_envs_struc(ctx, gp, p, nl, dv, K);
  // mean_efs = K^T alpha
  if (M)
    gemm_rm(ctx, false, false, 1, (int)n_out, M, 1.0, gp->alpha.as<double>(), M, K,
            (int)n_out, 0.0, d_mean_efs, (int)n_out);
  else
    FLARE_HIP_CHECK(hipMemsetAsync(d_mean_efs, 0, n_out * sizeof(double), ctx->stream));
  double *Y = ctx->buf_t<double>("sv.Y", (size_t)std::max(M, 1) * n_out);
  if (kind == kVarDTC) {
    self_kernel(ctx, gp, p, nl, dv, d_var);
    if (M) {
      // - diag(K^T K_uu^-1 K) = - |L^-1 K|^2 (K_uu^-1 = L^-T L^-1)
      gemm_rm(ctx, true, false, M, (int)n_out, M, 1.0, gp->Linv_cm.as<double>(), M, K,
              (int)n_out, 0.0, Y, (int)n_out);
      flare::launch_col_sqsum(ctx->stream, M, n_out, Y, -1.0, 1, d_var);
    }
  } else {
    FLARE_HIP_CHECK(hipMemsetAsync(d_var, 0, n_out * sizeof(double), ctx->stream));
  }
  if (M) {
    // + diag(K^T Sigma K) = |R_inv^T K|^2 (Sigma = R_inv R_inv^T)
    gemm_rm(ctx, true, false, M, (int)n_out, M, 1.0, gp->R_inv.as<double>(), M, K,
            (int)n_out, 0.0, Y, (int)n_out);
    flare::launch_col_sqsum(ctx->stream, M, n_out, Y, 1.0, 1, d_var);
  }
}

extern "C" {

int flare_sgp_envs_struc(flare_ctx *ctx, flare_sgp *gp, const flare_b2_params *params,
                         const flare_nlist *nl, const int *d_species, double *d_K) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_K, "null argument");
    flare::B2Dev p = make_b2(params, 0);
    check_species(nl, p.S);
    REQUIRE(p.S == gp->S && p.F == gp->F, "descriptor settings differ from the sparse set");
    ctx->use();
    DVDev dv = dv_build(ctx, p, nl, d_species, "sv");
    sgp_envs_struc(ctx, gp, p, nl, dv, d_K);
    ctx->sync();
  });
}

int flare_sgp_self_kernel_struc(flare_ctx *ctx, flare_sgp *gp,
                                const flare_b2_params *params, const flare_nlist *nl,
                                const int *d_species, double *d_out) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_out, "null argument");
    flare::B2Dev p = make_b2(params, 0);
    check_species(nl, p.S);
    REQUIRE(p.S == gp->S && p.F == gp->F, "descriptor settings differ from the sparse set");
    ctx->use();
    DVDev dv = dv_build(ctx, p, nl, d_species, "sv");
    self_kernel(ctx, gp, p, nl, dv, d_out);
  });
}

int flare_sgp_set_R_inv(flare_ctx *ctx, flare_sgp *gp, const double *d_R_inv) {
  return guarded([&] {
    REQUIRE(ctx && gp, "null argument");
    REQUIRE(gp->M == 0 || d_R_inv, "null R_inv");
    ctx->use();
    const size_t bytes = (size_t)std::max(gp->M, 1) * std::max(gp->M, 1) * sizeof(double);
    gp->R_inv.ensure(bytes);
    if (gp->M)
      FLARE_HIP_CHECK(hipMemcpyAsync(gp->R_inv.ptr, d_R_inv,
                                     (size_t)gp->M * gp->M * sizeof(double),
                                     hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sync();
    gp->have_R_inv = true;
  });
}

int flare_sgp_predict_SOR(flare_ctx *ctx, flare_sgp *gp, const flare_b2_params *params,
                          const flare_nlist *nl, const int *d_species,
                          double *d_mean_efs, double *d_variance_efs) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_mean_efs && d_variance_efs, "null argument");
    ctx->use();
    sgp_predict_var(ctx, gp, params, nl, d_species, d_mean_efs, d_variance_efs, kVarSOR);
    ctx->sync();
  });
}

int flare_sgp_predict_DTC(flare_ctx *ctx, flare_sgp *gp, const flare_b2_params *params,
                          const flare_nlist *nl, const int *d_species,
                          double *d_mean_efs, double *d_variance_efs) {
  return guarded([&] {
    REQUIRE(ctx && gp && d_mean_efs && d_variance_efs, "null argument");
    ctx->use();
    sgp_predict_var(ctx, gp, params, nl, d_species, d_mean_efs, d_variance_efs, kVarDTC);
    ctx->sync();
  });
}

}  // extern "C"
