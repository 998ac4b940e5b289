// neighbors.hip -- Structure neighbor list on the GPU.
//
// Produces the list of Structure::compute_neighbors (structure.cpp:52-118)
// with the same set and the same order: for atom i, neighbors sorted by
// (j, s1, s2, s3) with s in [-sweep, sweep]^3, predicate 0 < |r| < cutoff on
// r = (w_j - w_i) + s1 a1 + s2 a2 + s3 a3 (structure.cpp:68-77), evaluated
// with the reference's operation order and no FMA contraction so that the
// relative vectors are bit-identical.
//
// Build: wrap + bin (fractional cell list) -> per-atom candidate count ->
// scan -> fill -> per-atom rank sort by key -> reverse-pair index.  One
// wavefront per atom enumerates the candidates of the 27 (or, along thin
// axes, all) neighboring bins cooperatively.
#include <hipcub/hipcub.hpp>

#include <atomic>
#include <chrono>
#include <cmath>

#include "common.h"
#include "kernels.h"

namespace flare {

// 3x3 matrices used by the wrap (structure.cpp:120-133), row-major.
struct WrapMats {
  double ct[9], cdi[9], cd[9], cti[9], cell[9];
  int nb[3];
  int sweep;
  double cutoff;
};

__device__ inline void rowmul3(const double *v, const double *b, double *o) {
#pragma clang fp contract(off)
  for (int j = 0; j < 3; j++) {
    double s = v[0] * b[j];
    s = s + v[1] * b[3 + j];
    s = s + v[2] * b[6 + j];
    o[j] = s;
  }
}

// Per atom: wrapped position and bin, counted into bin_count[bin + 1]; per
// block of kRowBlock atoms the species counts (bcount, bad code -> *bad).
// next_flags: the accumulated summary ints (bad code, largest over-capacity
// count, pair total, bin overflow) of the other summary block, zeroed for the
// next build.
constexpr int kRowBlock = 256;

__global__ void __launch_bounds__(kRowBlock)
wrap_bin_kernel(int n, const double *__restrict__ pos, WrapMats m,
                const int *__restrict__ species, double *__restrict__ wrapped,
                int *__restrict__ bin_of, int *__restrict__ bin_count,
                int *__restrict__ bcount, int *__restrict__ bad,
                int *__restrict__ next_flags) {
#pragma clang fp contract(off)
  __shared__ int cnt[kMaxSpecies];
  const int t = threadIdx.x;
  if (blockIdx.x == 0 && t < 4) next_flags[t] = 0;
  if (t < kMaxSpecies) cnt[t] = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + t;
  if (i < n) {
    double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    double tt[3], rel[3], rw[3], t2[3], w[3];
    rowmul3(p, m.ct, tt);
    rowmul3(tt, m.cdi, rel);
    for (int c = 0; c < 3; c++) rw[c] = rel[c] - floor(rel[c]);
    rowmul3(rw, m.cd, t2);
    rowmul3(t2, m.cti, w);
    for (int c = 0; c < 3; c++) wrapped[3 * i + c] = w[c];
    int b[3];
    for (int c = 0; c < 3; c++) {
      int v = (int)floor(rw[c] * m.nb[c]);
      b[c] = v < 0 ? 0 : (v >= m.nb[c] ? m.nb[c] - 1 : v);
    }
    const int bin = (b[0] * m.nb[1] + b[1]) * m.nb[2] + b[2];
    bin_of[i] = bin;
    atomicAdd(&bin_count[bin + 1], 1);
    const int sp = species[i];
    if (sp < 0 || sp >= kMaxSpecies)
      atomicOr(bad, 1);
    else
      atomicAdd(&cnt[sp], 1);
  }
  __syncthreads();
  if (t < kMaxSpecies) bcount[blockIdx.x * kMaxSpecies + t] = cnt[t];
}

// Fixed-capacity bins: as wrap_bin_kernel, and each atom placed into a slot
// of its bin (bin b's slots [b kBinSlots, (b + 1) kBinSlots) in arrival
// order; the enumeration sorts every neighbor row, so the order within a bin
// does not matter).  bin_count[b + 1] ends as the bin's atom count; a bin
// with more atoms than slots sets *bin_ovf, and the build then groups the
// atoms again in the scanned layout.
constexpr int kBinSlots = 32;

__global__ void __launch_bounds__(kRowBlock)
wrap_place_kernel(int n, const double *__restrict__ pos, WrapMats m,
                  const int *__restrict__ species, double *__restrict__ wrapped,
                  int *__restrict__ bin_of, int *__restrict__ bin_count,
                  int *__restrict__ bcount, int *__restrict__ bad,
                  int *__restrict__ next_flags, int *__restrict__ bin_atoms,
                  double *__restrict__ bin_pos, int *__restrict__ bin_ovf) {
#pragma clang fp contract(off)
  __shared__ int cnt[kMaxSpecies];
  const int t = threadIdx.x;
  if (blockIdx.x == 0 && t < 4) next_flags[t] = 0;
  if (t < kMaxSpecies) cnt[t] = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + t;
  if (i < n) {
    double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    double tt[3], rel[3], rw[3], t2[3], w[3];
    rowmul3(p, m.ct, tt);
    rowmul3(tt, m.cdi, rel);
    for (int c = 0; c < 3; c++) rw[c] = rel[c] - floor(rel[c]);
    rowmul3(rw, m.cd, t2);
    rowmul3(t2, m.cti, w);
    for (int c = 0; c < 3; c++) wrapped[3 * i + c] = w[c];
    int b[3];
    for (int c = 0; c < 3; c++) {
      int v = (int)floor(rw[c] * m.nb[c]);
      b[c] = v < 0 ? 0 : (v >= m.nb[c] ? m.nb[c] - 1 : v);
    }
    const int bin = (b[0] * m.nb[1] + b[1]) * m.nb[2] + b[2];
    bin_of[i] = bin;
    const int slot = atomicAdd(&bin_count[bin + 1], 1);
    if (slot < kBinSlots) {
      const int q = bin * kBinSlots + slot;
      bin_atoms[q] = i;
      for (int c = 0; c < 3; c++) bin_pos[3 * q + c] = w[c];
    } else {
      atomicOr(bin_ovf, 1);
    }
    const int sp = species[i];
    if (sp < 0 || sp >= kMaxSpecies)
      atomicOr(bad, 1);
    else
      atomicAdd(&cnt[sp], 1);
  }
  __syncthreads();
  if (t < kMaxSpecies) bcount[blockIdx.x * kMaxSpecies + t] = cnt[t];
}

// bin_count[b + 1] = number of atoms with bin_of == b (bin_count zero on
// entry): the counts the scanned layout needs when fixed-capacity bins
// overflowed.
__global__ void __launch_bounds__(256)
count_bins_kernel(int n, const int *__restrict__ bin_of, int *__restrict__ bin_count,
                  int *__restrict__ clear, int nclear) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nclear) clear[i] = 0;
  if (i < n) atomicAdd(&bin_count[bin_of[i] + 1], 1);
}

// Zero the build's counters in one launch: a[0..na), b[0..nb), c[0..nc),
// d[0..nd).
__global__ void clear4_kernel(int *__restrict__ a, int na, int *__restrict__ b,
                              int nb, int *__restrict__ c, int nc, int *__restrict__ d,
                              int nd) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < na) a[q] = 0;
  if (q < nb) b[q] = 0;
  if (q < nc) c[q] = 0;
  if (q < nd) d[q] = 0;
}

// Species-grouped row map (b2.cpp:104-176 regroup: atoms of species 0 in
// original order, then species 1, ...): per-block species counts (in
// wrap_bin_kernel or ext_species_kernel), their scan (species_block_scan),
// and per-block ranks from wavefront ballots.
// Block body (blockDim kRowBlock, every thread reaches it).
// blk_off: the block's kMaxSpecies offsets (atoms of each species in the
// blocks before it), totals: the species counts of the whole structure.
__device__ inline void species_rows_body(int n, const int *__restrict__ species,
                                         const int *blk_off, const int *totals,
                                         int *__restrict__ row_of,
                                         int *__restrict__ atom_of_row,
                                         int *__restrict__ species_of_row) {
  __shared__ int wcnt[kRowBlock / 64][kMaxSpecies];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int a = blockIdx.x * kRowBlock + t;
  const int s = a < n ? species[a] : -1;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  int my_rank = 0;
  for (int q = 0; q < kMaxSpecies; q++) {
    const uint64_t m = __ballot(s == q);
    if (s == q) my_rank = __popcll(m & lt);
    if (lane == 0) wcnt[wv][q] = __popcll(m);
  }
  __syncthreads();
  // a code outside [0, kMaxSpecies) gets no row (the build reports it)
  if (s >= 0 && s < kMaxSpecies) {
    int base = 0;
    for (int q = 0; q < s; q++) base += totals[q];
    for (int w = 0; w < wv; w++) my_rank += wcnt[w][s];
    const int row = base + blk_off[s] + my_rank;
    row_of[a] = row;
    atom_of_row[row] = a;
    species_of_row[row] = s;
  }
}

// cum[i] = sum of count[k], k < i, from the per-block pair sums
// blk_sum[b kCumStride] (count[i] accumulated into block i / kCumBlock by
// the enumeration; one 128-byte line per block sum, so that the atomics of
// different blocks do not queue on one line): workgroup b adds the sums of
// the blocks before it and scans its own kCumBlock counts; the last
// workgroup writes cum[n] and the pair total.  Integer sums, so the result
// does not depend on the order of the enumeration's atomics.  For n <=
// kCumBlock kCumMaxBlocks (each workgroup reads at most kCumMaxBlocks sums).
// Copies the finished summary counts[0..kNlistStats) into the pinned host
// block, then its sequence number (one thread; system-scope release so the
// host sees the counts once it sees seq).
__device__ inline void publish_stats(const int *counts, int *host, unsigned seq,
                                     const int *species_totals = nullptr) {
  for (int k = 0; k < kNlistStats; k++)
    host[k] = species_totals && k < kMaxSpecies ? species_totals[k] : counts[k];
  __threadfence_system();
  __hip_atomic_store(host + kNlistSeq, (int)seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void publish_stats_kernel(const int *__restrict__ counts, int *host, unsigned seq) {
  if (threadIdx.x == 0) publish_stats(counts, host, seq);
}

// With rows.species set (the fixed-capacity bin path, where no bin scan
// ran), the same workgroups also form the species-grouped row map of their
// 256 atoms: each sums the per-block species counts of the blocks before it
// and of all blocks, and the last one's totals go into the summary.
struct CumRows {
  const int *species = nullptr;
  const int *bcount = nullptr;
  int *row_of = nullptr, *atom_of_row = nullptr, *species_of_row = nullptr;
};

constexpr int kCumBlock = 256, kCumMaxBlocks = 1024, kCumStride = 32;
static_assert(kCumBlock == kRowBlock, "the species blocks are the CSR blocks");
__global__ void __launch_bounds__(kCumBlock)
cum_blocks_kernel(int n, const int *__restrict__ count, const int *__restrict__ blk_sum,
                  int *__restrict__ cum, int *__restrict__ counts, int *host,
                  unsigned seq, CumRows rows) {
  constexpr int kW = kCumBlock / kWave;
  __shared__ int before_w[kW], total_w[kW];
  __shared__ int sp_before[kMaxSpecies], sp_total[kMaxSpecies];
  const int b = blockIdx.x, t = threadIdx.x, lane = t & (kWave - 1), wv = t / kWave;
  if (rows.species) {
    for (int sp = wv; sp < kMaxSpecies; sp += kW) {
      int bef = 0, tot = 0;
      for (int k = lane; k < (int)gridDim.x; k += kWave) {
        const int v = rows.bcount[k * kMaxSpecies + sp];
        tot += v;
        if (k < b) bef += v;
      }
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        bef += __shfl_xor(bef, off, kWave);
        tot += __shfl_xor(tot, off, kWave);
      }
      if (lane == 0) {
        sp_before[sp] = bef;
        sp_total[sp] = tot;
      }
    }
    __syncthreads();
    species_rows_body(n, rows.species, sp_before, sp_total, rows.row_of,
                      rows.atom_of_row, rows.species_of_row);
  }
  int s = 0;
  for (int k = t; k < b; k += kCumBlock) s += blk_sum[k * kCumStride];
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
  const int i = b * kCumBlock + t;
  const int c = i < n ? count[i] : 0;
  int incl = c;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int v = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += v;
  }
  if (lane == 0) before_w[wv] = s;
  if (lane == kWave - 1) total_w[wv] = incl;
  __syncthreads();
  int base = 0;
#pragma unroll
  for (int w = 0; w < kW; w++) base += before_w[w] + (w < wv ? total_w[w] : 0);
  const int ex = base + incl - c;
  if (i < n) cum[i] = ex;
  if (b == gridDim.x - 1 && t == kCumBlock - 1) {
    cum[n] = ex + c;
    counts[kMaxSpecies + 2] = ex + c;
    if (rows.species)
      for (int sp = 0; sp < kMaxSpecies; sp++) counts[sp] = sp_total[sp];
    publish_stats(counts, host, seq, rows.species ? sp_total : nullptr);
  }
}

// Single-workgroup exclusive scan: out[0..n] with out[n] = total, for
// n <= C * kScanChunk.  Chunk c holds elements [c kScanChunk, (c + 1)
// kScanChunk), thread t its four elements 4t..4t+3, so every load and store
// instruction covers 1 KiB per wavefront contiguously (16-byte vectors where
// in / out are 16-byte aligned).  All chunks are loaded first; per chunk a
// wavefront scan of the four-element sums (shuffles), then one wavefront
// scans the C x 16 wavefront totals in element order.  Two barriers in all.
// Used for the bin counts and the per-atom counts, where one workgroup beats
// a device-wide scan's two launches.
constexpr int kScanThreads = 1024;
constexpr int kScanChunk = 4 * kScanThreads;
constexpr int kScanMaxC = 8;  // 16 would spill at 1,024 threads

template <int C>
__device__ inline int block_scan_body(int n, const int *__restrict__ in,
                                      int *__restrict__ out) {
  constexpr int NW = kScanThreads / kWave;
  static_assert(C * NW <= 4 * kWave, "one wavefront scans the totals");
  __shared__ int wtot[C * NW];
  __shared__ int total_sh;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const bool vin = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
  const bool vout = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  int4 v[C];
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = c * kScanChunk + 4 * t;
    if (vin && e + 3 < n) {
      v[c] = *reinterpret_cast<const int4 *>(in + e);
    } else {
      v[c].x = e < n ? in[e] : 0;
      v[c].y = e + 1 < n ? in[e + 1] : 0;
      v[c].z = e + 2 < n ? in[e + 2] : 0;
      v[c].w = e + 3 < n ? in[e + 3] : 0;
    }
  }
  int s[C], incl[C];
#pragma unroll
  for (int c = 0; c < C; c++) {
    s[c] = v[c].x + v[c].y + v[c].z + v[c].w;
    incl[c] = s[c];
  }
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
#pragma unroll
    for (int c = 0; c < C; c++) {
      const int y = __shfl_up(incl[c], off, kWave);
      if (lane >= off) incl[c] += y;
    }
  }
  if (lane == kWave - 1)
#pragma unroll
    for (int c = 0; c < C; c++) wtot[c * NW + w] = incl[c];
  __syncthreads();
  if (w == 0) {
    // exclusive offsets of the C * NW wavefront totals, four per lane
    int x[4], ls = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = 4 * lane + k;
      x[k] = e < C * NW ? wtot[e] : 0;
      ls += x[k];
    }
    int li = ls;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(li, off, kWave);
      if (lane >= off) li += y;
    }
    int run = li - ls;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = 4 * lane + k;
      if (e < C * NW) wtot[e] = run;
      run += x[k];
    }
    if (lane == kWave - 1) total_sh = li;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int e = c * kScanChunk + 4 * t;
    int4 o;
    o.x = wtot[c * NW + w] + incl[c] - s[c];
    o.y = o.x + v[c].x;
    o.z = o.y + v[c].y;
    o.w = o.z + v[c].z;
    if (vout && e + 3 < n) {
      *reinterpret_cast<int4 *>(out + e) = o;
    } else {
      if (e < n) out[e] = o.x;
      if (e + 1 < n) out[e + 1] = o.y;
      if (e + 2 < n) out[e + 2] = o.z;
      if (e + 3 < n) out[e + 3] = o.w;
    }
  }
  if (t == 0) out[n] = total_sh;
  return total_sh;
}

template <int C>
__global__ void __launch_bounds__(kScanThreads)
scan_block_kernel(int n, const int *__restrict__ in, int *__restrict__ out) {
  block_scan_body<C>(n, in, out);
}

// After the count scan: the species scan of the per-block counts (boff,
// counts[s]) and the pair total cum[n] into stat_P, so that every number
// the host needs sits in one small buffer.
// One wavefront per species slot sp (threads 64 sp .. 64 sp + 63).
__device__ inline void species_block_scan(int nblk, const int *__restrict__ bcount,
                                          int *__restrict__ boff, int *__restrict__ counts) {
  const int lane = threadIdx.x & 63, sp = threadIdx.x >> 6;
  int run = 0;
  for (int b0 = 0; b0 < nblk; b0 += 64) {
    const int b = b0 + lane;
    const int v = b < nblk ? bcount[b * kMaxSpecies + sp] : 0;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      int u = __shfl_up(incl, off, 64);
      if (lane >= off) incl += u;
    }
    if (b < nblk) boff[b * kMaxSpecies + sp] = run + incl - v;
    run += __shfl(incl, 63, 64);
  }
  if (lane == 0) counts[sp] = run;
}

__global__ void __launch_bounds__(64 * kMaxSpecies)
nbr_stats_kernel(int n, const int *__restrict__ cum, int nblk,
                 const int *__restrict__ bcount, int *__restrict__ boff,
                 int *__restrict__ counts, int *__restrict__ stat_P) {
  if (threadIdx.x == 0) *stat_P = cum[n];
  if (bcount) species_block_scan(nblk, bcount, boff, counts);
}

// The species scan alone (its counts are the bin scan's when that runs on
// the device-wide path).
__global__ void __launch_bounds__(64 * kMaxSpecies)
species_scan_kernel(int nblk, const int *__restrict__ bcount, int *__restrict__ boff,
                    int *__restrict__ counts) {
  species_block_scan(nblk, bcount, boff, counts);
}

// The count scan (n <= kScanMaxC kScanChunk) with nbr_stats_kernel's work as
// its epilogue: one launch instead of two.  bcount null: the pair total only
// (the species scan ran with the bin scan).
template <int C>
__global__ void __launch_bounds__(kScanThreads)
count_scan_stats_kernel(int n, const int *__restrict__ count, int *__restrict__ cum,
                        int nblk, const int *__restrict__ bcount, int *__restrict__ boff,
                        int *__restrict__ counts, int *__restrict__ stat_P) {
  static_assert(kScanThreads >= 64 * kMaxSpecies, "a wavefront per species slot");
  const int total = block_scan_body<C>(n, count, cum);
  if (threadIdx.x == 0) *stat_P = total;
  if (bcount && threadIdx.x < 64 * kMaxSpecies) species_block_scan(nblk, bcount, boff, counts);
}

// The bin-count scan (nbins <= kScanMaxC kScanChunk) with the species scan of
// the per-block counts as its epilogue, so that the species rows can be
// written by bin_fill_rows_kernel.
template <int C>
__global__ void __launch_bounds__(kScanThreads)
bin_scan_species_kernel(int nbins, const int *__restrict__ bin_count1,
                        int *__restrict__ bin_start, int nblk,
                        const int *__restrict__ bcount, int *__restrict__ boff,
                        int *__restrict__ counts) {
  block_scan_body<C>(nbins, bin_count1, bin_start);
  if (threadIdx.x < 64 * kMaxSpecies) species_block_scan(nblk, bcount, boff, counts);
}


namespace {

// One per-axis candidate: (target bin along axis, image shift)
__device__ inline void axis_entry(int nb, int home, int sw, int e, int &bin,
                                  int &shift) {
  if (nb >= 3) {
    int t = home + e - 1;
    shift = 0;
    if (t < 0) {
      t += nb;
      shift = -1;
    } else if (t >= nb) {
      t -= nb;
      shift = 1;
    }
    bin = t;
  } else {
    int ns = 2 * sw + 1;
    bin = e / ns;
    shift = e % ns - sw;
  }
}

__device__ inline int axis_len(int nb, int sw) {
  return nb >= 3 ? 3 : nb * (2 * sw + 1);
}

// Sort key of a neighbor entry: j, then the image shifts, as bit fields
// (8 bits per shift offset by the sweep, which is at most kMaxSweep), so
// that integer order is the reference's (j, s1, s2, s3) order.
__device__ inline int64_t pair_key(int j, int s1, int s2, int s3, int sw) {
  return ((int64_t)j << 24) | ((int64_t)(s1 + sw) << 16) |
         ((int64_t)(s2 + sw) << 8) | (int64_t)(s3 + sw);
}
__device__ inline void key_parts(int64_t k, int sw, int &j, int &s1, int &s2,
                                 int &s3) {
  j = (int)(k >> 24);
  s1 = (int)((k >> 16) & 0xff) - sw;
  s2 = (int)((k >> 8) & 0xff) - sw;
  s3 = (int)(k & 0xff) - sw;
}

// Reference predicate and relative vector (structure.cpp:68-77)
__device__ inline bool image_test(const double *wi, const double *wj,
                                  const double *cell, int s1, int s2, int s3,
                                  double cutoff, double *o) {
#pragma clang fp contract(off)
  double im[3];
  for (int c = 0; c < 3; c++) {
    double d = wj[c] - wi[c];
    d = d + (double)s1 * cell[c];
    d = d + (double)s2 * cell[3 + c];
    d = d + (double)s3 * cell[6 + c];
    im[c] = d;
  }
  double dist = sqrt(im[0] * im[0] + im[1] * im[1] + im[2] * im[2]);
  o[0] = dist;
  o[1] = im[0];
  o[2] = im[1];
  o[3] = im[2];
  return (dist < cutoff) && (dist != 0);
}

// Atoms grouped by bin: either a scanned layout (bin b's atoms at
// [start[b], start[b + 1])) or fixed-capacity bins of cap slots (bin b's at
// [b cap, b cap + min(cnt[b], cap)); a fuller bin makes the build fall back
// to the scanned layout).
struct BinLayout {
  const int *start;
  const int *cnt;
  int cap;
};
__device__ inline void bin_range(const BinLayout &bl, int b, int &st, int &cs) {
  if (bl.cap) {
    st = b * bl.cap;
    cs = min(bl.cnt[b], bl.cap);
  } else {
    st = bl.start[b];
    cs = bl.start[b + 1] - st;
  }
}

}  // namespace

// MODE 0: count candidates per atom.  MODE 1: write them (unsorted) into
// the atom's CSR segment.  One wavefront per atom.
template <int MODE>
__global__ void __launch_bounds__(256)
nbr_enum_kernel(int n, WrapMats m, const double *__restrict__ wrapped,
                const int *__restrict__ bin_of, BinLayout bl,
                const int *__restrict__ bin_atoms, int *__restrict__ count,
                const int *__restrict__ cum, int64_t *__restrict__ keys,
                double *__restrict__ rel) {
  __shared__ int prefix_sh[4][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const int i = blockIdx.x * 4 + wv;
  if (i >= n) return;  // whole wave exits together
  int *prefix = prefix_sh[wv];
  const int sw = m.sweep;
  const int hb = bin_of[i];
  const int hx = hb / (m.nb[1] * m.nb[2]);
  const int hy = (hb / m.nb[2]) % m.nb[1];
  const int hz = hb % m.nb[2];
  const int lx = axis_len(m.nb[0], sw), ly = axis_len(m.nb[1], sw),
            lz = axis_len(m.nb[2], sw);
  const int n_combo = lx * ly * lz;
  const double wi[3] = {wrapped[3 * i], wrapped[3 * i + 1], wrapped[3 * i + 2]};
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  int found = 0;
  int64_t base = (MODE == 1) ? (int64_t)cum[i] : 0;

  for (int c0 = 0; c0 < n_combo; c0 += kWave) {
    int c = c0 + lane;
    int cb = 0, s1 = 0, s2 = 0, s3 = 0, csize = 0;
    if (c < n_combo) {
      int ex = c / (ly * lz), ey = (c / lz) % ly, ez = c % lz;
      int bx, by, bz, cst;
      axis_entry(m.nb[0], hx, sw, ex, bx, s1);
      axis_entry(m.nb[1], hy, sw, ey, by, s2);
      axis_entry(m.nb[2], hz, sw, ez, bz, s3);
      cb = (bx * m.nb[1] + by) * m.nb[2] + bz;
      bin_range(bl, cb, cst, csize);
    }
    // inclusive scan of combo sizes across the wave
    int incl = csize;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      int v = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += v;
    }
    prefix[lane] = incl;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int total = __shfl(incl, kWave - 1, kWave);
    for (int t0 = 0; t0 < total; t0 += kWave) {
      int t = t0 + lane;
      bool pass = false;
      double o[4];
      int64_t key = 0;
      if (t < total) {
        // first combo whose inclusive prefix exceeds t
        int lo = 0, hi = kWave - 1;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (prefix[mid] > t)
            hi = mid;
          else
            lo = mid + 1;
        }
        int cc = lo;
        int before = cc ? prefix[cc - 1] : 0;
        int cidx = c0 + cc;
        int ex = cidx / (ly * lz), ey = (cidx / lz) % ly, ez = cidx % lz;
        int bx, by, bz, t1, t2, t3;
        axis_entry(m.nb[0], hx, sw, ex, bx, t1);
        axis_entry(m.nb[1], hy, sw, ey, by, t2);
        axis_entry(m.nb[2], hz, sw, ez, bz, t3);
        int b = (bx * m.nb[1] + by) * m.nb[2] + bz, bst, bcs;
        bin_range(bl, b, bst, bcs);
        int j = bin_atoms[bst + (t - before)];
        const double wj[3] = {wrapped[3 * j], wrapped[3 * j + 1],
                              wrapped[3 * j + 2]};
        pass = image_test(wi, wj, m.cell, t1, t2, t3, m.cutoff, o);
        key = pair_key(j, t1, t2, t3, sw);
      }
      uint64_t mask = __ballot(pass);
      if (MODE == 1 && pass) {
        int64_t slot = base + found + __popcll(mask & lt_mask);
        if (slot >= (int64_t)cum[i + 1]) continue;  // count/fill disagree
        keys[slot] = key;
        rel[4 * slot + 0] = o[0];
        rel[4 * slot + 1] = o[1];
        rel[4 * slot + 2] = o[2];
        rel[4 * slot + 3] = o[3];
      }
      found += __popcll(mask);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (MODE == 0 && lane == 0) count[i] = found;
}

// Fused pass: enumerate one atom's neighbors (as nbr_enum_kernel), keep their
// keys in LDS, rank-sort them and write them to the atom's padded slot row
// (kNbrCap entries) of keys_pad; the rel vectors are recomputed from the
// keys by nbr_compact_kernel.  count[i] is always written;
// atoms with more than kNbrCap neighbors are left to the two-pass path
// (stat[0] becomes the largest such count; it stays 0 otherwise).
// THIN: some axis has fewer than three bins (general combination lists);
// otherwise the 27 neighboring bins with fixed per-axis lists of three.
// One wavefront's pass for atom i; prefix (kWave ints) and skey (kNbrCap
// entries) are the wavefront's LDS scratch.
template <bool THIN>
__device__ inline void enum_sort_atom(int i, const WrapMats &m,
                                      const double *__restrict__ wrapped,
                                      const int *__restrict__ bin_of, const BinLayout &bl,
                                      const int *__restrict__ bin_atoms,
                                      const double *__restrict__ bin_pos,
                                      int *__restrict__ count,
                                      int64_t *__restrict__ keys_pad, int *__restrict__ stat,
                                      int *__restrict__ blk_sum, int *prefix, int64_t *skey) {
  const int lane = threadIdx.x & (kWave - 1);
  const int sw = m.sweep;
  const int hb = bin_of[i];
  const int hx = hb / (m.nb[1] * m.nb[2]);
  const int hy = (hb / m.nb[2]) % m.nb[1];
  const int hz = hb % m.nb[2];
  const int lx = THIN ? axis_len(m.nb[0], sw) : 3,
            ly = THIN ? axis_len(m.nb[1], sw) : 3,
            lz = THIN ? axis_len(m.nb[2], sw) : 3;
  const int n_combo = lx * ly * lz;
  const double wi[3] = {wrapped[3 * i], wrapped[3 * i + 1], wrapped[3 * i + 2]};
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  int found = 0;

  for (int c0 = 0; c0 < n_combo; c0 += kWave) {
    int c = c0 + lane;
    int csize = 0;
    if (c < n_combo) {
      int ex = c / (ly * lz), ey = (c / lz) % ly, ez = c % lz;
      int bx, by, bz, s1, s2, s3;
      axis_entry(m.nb[0], hx, sw, ex, bx, s1);
      axis_entry(m.nb[1], hy, sw, ey, by, s2);
      axis_entry(m.nb[2], hz, sw, ez, bz, s3);
      const int cb = (bx * m.nb[1] + by) * m.nb[2] + bz;
      int cst;
      bin_range(bl, cb, cst, csize);
    }
    int incl = csize;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      int v = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += v;
    }
    prefix[lane] = incl;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int total = __shfl(incl, kWave - 1, kWave);
    for (int t0 = 0; t0 < total; t0 += kWave) {
      const int t = t0 + lane;
      bool pass = false;
      double o[4];
      int64_t key = 0;
      if (t < total) {
        int lo = 0, hi = kWave - 1;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (prefix[mid] > t)
            hi = mid;
          else
            lo = mid + 1;
        }
        const int cc = lo;
        const int before = cc ? prefix[cc - 1] : 0;
        const int cidx = c0 + cc;
        int ex = cidx / (ly * lz), ey = (cidx / lz) % ly, ez = cidx % lz;
        int bx, by, bz, t1, t2, t3;
        axis_entry(m.nb[0], hx, sw, ex, bx, t1);
        axis_entry(m.nb[1], hy, sw, ey, by, t2);
        axis_entry(m.nb[2], hz, sw, ez, bz, t3);
        const int b = (bx * m.nb[1] + by) * m.nb[2] + bz;
        int bst, bcs;
        bin_range(bl, b, bst, bcs);
        const int slot = bst + (t - before);
        const int j = bin_atoms[slot];
        const double wj[3] = {bin_pos[3 * slot], bin_pos[3 * slot + 1],
                              bin_pos[3 * slot + 2]};
        pass = image_test(wi, wj, m.cell, t1, t2, t3, m.cutoff, o);
        key = pair_key(j, t1, t2, t3, sw);
      }
      const uint64_t mask = __ballot(pass);
      if (pass) {
        const int slot = found + __popcll(mask & lt_mask);
        if (slot < kNbrCap) skey[slot] = key;
      }
      found += __popcll(mask);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) {
    count[i] = found;
    if (blk_sum) atomicAdd(blk_sum + i / kCumBlock * kCumStride, found);
    if (found > kNbrCap) atomicMax(stat, found);  // rare: no contention
  }
  if (found > kNbrCap) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // rank = number of smaller keys (keys are distinct); every lane reads the
  // same key per step, an LDS broadcast
  int64_t k0 = lane < found ? skey[lane] : INT64_MAX;
  int64_t k1 = lane + kWave < found ? skey[lane + kWave] : INT64_MAX;
  int r0 = 0, r1 = 0;
  for (int q = 0; q < found; q++) {
    const int64_t kq = skey[q];
    r0 += kq < k0;
    r1 += kq < k1;
  }
  const int64_t row = (int64_t)i * kNbrCap;
  if (lane < found) keys_pad[row + r0] = k0;
  if (lane + kWave < found) keys_pad[row + r1] = k1;
}

template <bool THIN>
__global__ void __launch_bounds__(256)
nbr_enum_sort_kernel(int n, WrapMats m, const double *__restrict__ wrapped,
                     const int *__restrict__ bin_of, BinLayout bl,
                     const int *__restrict__ bin_atoms,
                     const double *__restrict__ bin_pos, int *__restrict__ count,
                     int64_t *__restrict__ keys_pad, int *__restrict__ stat,
                     int *__restrict__ blk_sum) {
  __shared__ int prefix_sh[4][kWave];
  __shared__ int64_t skey_sh[4][kNbrCap];
  const int wv = threadIdx.x / kWave;
  const int i = blockIdx.x * 4 + wv;
  if (i >= n) return;  // whole wave exits together
  enum_sort_atom<THIN>(i, m, wrapped, bin_of, bl, bin_atoms, bin_pos, count,
                       keys_pad, stat, blk_sum, prefix_sh[wv], skey_sh[wv]);
}

// Candidates of one home bin's 27 neighboring bins staged in LDS: FP32
// shifted position and the combination (bin, image shift) index.
constexpr int kBinCap = 480;  // keeps the workgroup's LDS under 32 KiB: five per CU
struct __align__(16) BinCand {
  float x, y, z;
  int k;
};
// LDS of nbr_bin_enum_sort_kernel, in bytes.  Staged path: candidate FP64
// positions [3][kBinCap], BinCand [kBinCap], atom ids [kBinCap], then per
// wavefront the found keys (64-bit [kNbrCap], 32-bit [kNbrCap]) and the
// prefilter queue (u16 [kBinCap]).
// Fallback path (enum_sort_atom): prefix [4][kWave], skey [4][kNbrCap], in
// the same bytes.
constexpr int kBinOffPf = 3 * kBinCap * 8;
constexpr int kBinOffJ = kBinOffPf + kBinCap * 16;
constexpr int kBinOffWave = kBinOffJ + kBinCap * 4;
constexpr int kBinWaveBytes = kNbrCap * 8 + kNbrCap * 4 + kBinCap * 2;
constexpr int kBinStaged = kBinOffWave + 4 * kBinWaveBytes;
constexpr int kBinFallback = 4 * kWave * 4 + 4 * kNbrCap * 8;
constexpr int kBinLds = kBinStaged > kBinFallback ? kBinStaged : kBinFallback;
static_assert(kBinOffWave % 16 == 0 && kBinWaveBytes % 16 == 0 && (kNbrCap * 8) % 16 == 0,
              "LDS alignment (the 32-bit key rows are read as 16-byte vectors)");

// The same rows as nbr_enum_sort_kernel<false> (at least three bins along
// every axis), one workgroup per home bin.  The candidates of the 27
// neighboring bins (one image shift each) are staged in LDS once for all
// of the bin's atoms (which are the candidates of the home bin itself, the
// combination 13 = (0, 0, 0) shift): ids, FP64 wrapped positions and FP32
// copies of the shifted positions.  Each atom's wavefront keeps the
// candidates within cut_f of an FP32 distance test, then runs the exact
// FP64 test of structure.cpp:68-77 on those only.  cut_f exceeds the cutoff
// by more than the FP32 rounding of the test (positions rounded to FP32,
// differences and squares in FP32), so no neighbor is lost; the kept set,
// its rel vectors and its order are those of the exact test alone.  Rows are
// rank-sorted by a 32-bit key (j << 6 | the three shifts + 1 in two bits
// each, the (j, s1, s2, s3) order) when key32 is set (n <= 2^26), else by
// the 64-bit key; nbr_compact_kernel recomputes the rel vectors from the
// keys.  A home bin with more than kBinCap candidates falls back to
// enum_sort_atom for its atoms.
__global__ void __launch_bounds__(256)
nbr_bin_enum_sort_kernel(WrapMats m, float cut_f2, int key32,
                         const double *__restrict__ wrapped,
                         const int *__restrict__ bin_of, BinLayout bl,
                         const int *__restrict__ bin_atoms,
                         const double *__restrict__ bin_pos, int *__restrict__ count,
                         int64_t *__restrict__ keys_pad, int *__restrict__ stat,
                         int *__restrict__ blk_sum) {
  __shared__ int c_pre[28], c_start[27], c_shift[27];
  __shared__ __align__(16) unsigned char lds[kBinLds];
  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  int a0, na;
  bin_range(bl, b, a0, na);
  if (na == 0) return;  // the whole workgroup
  const int nb1 = m.nb[1], nb2 = m.nb[2];
  if (wv == 0) {
    const int hx = b / (nb1 * nb2), hy = (b / nb2) % nb1, hz = b % nb2;
    int cs = 0, st = 0, sh = 0;
    if (lane < 27) {
      int bx, by, bz, s1, s2, s3;
      axis_entry(m.nb[0], hx, m.sweep, lane / 9, bx, s1);
      axis_entry(nb1, hy, m.sweep, (lane / 3) % 3, by, s2);
      axis_entry(nb2, hz, m.sweep, lane % 3, bz, s3);
      const int cb = (bx * nb1 + by) * nb2 + bz;
      bin_range(bl, cb, st, cs);
      sh = ((s1 + 1) << 4) | ((s2 + 1) << 2) | (s3 + 1);
    }
    int incl = cs;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
      const int v = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += v;
    }
    if (lane < 27) {
      c_pre[lane + 1] = incl;
      c_start[lane] = st;
      c_shift[lane] = sh;
    }
    if (lane == 0) c_pre[0] = 0;
  }
  __syncthreads();
  const int total = c_pre[27];
  if (total > kBinCap) {
    int *prefix = reinterpret_cast<int *>(lds) + wv * kWave;
    int64_t *skey = reinterpret_cast<int64_t *>(lds + 4 * kWave * 4) + wv * kNbrCap;
    for (int a = wv; a < na; a += 4)
      enum_sort_atom<false>(bin_atoms[a0 + a], m, wrapped, bin_of, bl, bin_atoms,
                            bin_pos, count, keys_pad, stat, blk_sum, prefix, skey);
    return;
  }
  double(*cw)[kBinCap] = reinterpret_cast<double(*)[kBinCap]>(lds);
  BinCand *cpf = reinterpret_cast<BinCand *>(lds + kBinOffPf);
  int *cj = reinterpret_cast<int *>(lds + kBinOffJ);
  for (int t = tid; t < total; t += 256) {
    // combo k: the last with c_pre[k] <= t
    int lo = 0, hi = 26;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (c_pre[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int slot = c_start[lo] + (t - c_pre[lo]);
    const int sh = c_shift[lo];
    const int j = bin_atoms[slot];
    const double x = bin_pos[3 * slot], y = bin_pos[3 * slot + 1], z = bin_pos[3 * slot + 2];
    const double s1 = (sh >> 4) - 1, s2 = ((sh >> 2) & 3) - 1, s3 = (sh & 3) - 1;
    BinCand pf;
    pf.x = (float)(x + s1 * m.cell[0] + s2 * m.cell[3] + s3 * m.cell[6]);
    pf.y = (float)(y + s1 * m.cell[1] + s2 * m.cell[4] + s3 * m.cell[7]);
    pf.z = (float)(z + s1 * m.cell[2] + s2 * m.cell[5] + s3 * m.cell[8]);
    pf.k = lo;
    cw[0][t] = x;
    cw[1][t] = y;
    cw[2][t] = z;
    cj[t] = j;
    cpf[t] = pf;
  }
  __syncthreads();
  unsigned char *wb = lds + kBinOffWave + wv * kBinWaveBytes;
  int64_t *skey = reinterpret_cast<int64_t *>(wb);
  unsigned *skey32 = reinterpret_cast<unsigned *>(wb + kNbrCap * 8);
  unsigned short *queue = reinterpret_cast<unsigned short *>(wb + kNbrCap * 12);
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int home0 = c_pre[13];
  for (int a = wv; a < na; a += 4) {
    const int hc = home0 + a;
    const int i = cj[hc];
    const double wi[3] = {cw[0][hc], cw[1][hc], cw[2][hc]};
    const float wx = cpf[hc].x, wy = cpf[hc].y, wz = cpf[hc].z;
    int nq = 0;
    for (int t0 = 0; t0 < total; t0 += kWave) {
      const int t = t0 + lane;
      bool keep = false;
      if (t < total) {
        const BinCand c = cpf[t];
        const float dx = c.x - wx, dy = c.y - wy, dz = c.z - wz;
        keep = dx * dx + dy * dy + dz * dz < cut_f2;
      }
      const uint64_t mk = __ballot(keep);
      if (keep) queue[nq + __popcll(mk & lt_mask)] = (unsigned short)t;
      nq += __popcll(mk);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int found = 0;
    for (int q0 = 0; q0 < nq; q0 += kWave) {
      const int q = q0 + lane;
      bool pass = false;
      int t = 0;
      int64_t key = 0;
      unsigned k32 = 0;
      if (q < nq) {
        t = queue[q];
        const int sh = c_shift[cpf[t].k];
        const int s1 = (sh >> 4) - 1, s2 = ((sh >> 2) & 3) - 1, s3 = (sh & 3) - 1;
        const int j = cj[t];
        const double wj[3] = {cw[0][t], cw[1][t], cw[2][t]};
        double o[4];
        pass = image_test(wi, wj, m.cell, s1, s2, s3, m.cutoff, o);
        key = pair_key(j, s1, s2, s3, m.sweep);
        k32 = ((unsigned)j << 6) | (unsigned)sh;
      }
      const uint64_t mk = __ballot(pass);
      if (pass) {
        const int sl = found + __popcll(mk & lt_mask);
        if (sl < kNbrCap) {
          skey[sl] = key;
          skey32[sl] = k32;
        }
      }
      found += __popcll(mk);
    }
    if (lane == 0) {
      count[i] = found;
      if (blk_sum) atomicAdd(blk_sum + i / kCumBlock * kCumStride, found);
      if (found > kNbrCap) atomicMax(stat, found);  // rare: no contention
    }
    // pad the 32-bit keys to a multiple of four with a key no entry is below
    // (j < 2^26, so every key is at most 0xffffffff)
    if (lane < 3 && found + lane < ((found + 3) & ~3) && found + lane < kNbrCap)
      skey32[found + lane] = 0xffffffffu;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (found <= kNbrCap) {
      // rank = number of smaller keys (keys are distinct), four keys per
      // LDS broadcast read
      int r0 = 0, r1 = 0;
      if (key32) {
        const unsigned k0 = lane < found ? skey32[lane] : 0xffffffffu;
        const unsigned k1 = lane + kWave < found ? skey32[lane + kWave] : 0xffffffffu;
        const uint4 *sk4 = reinterpret_cast<const uint4 *>(skey32);
        if (found <= kWave) {
          for (int q = 0; q < found; q += 4) {
            const uint4 kk = sk4[q >> 2];
            r0 += (kk.x < k0) + (kk.y < k0) + (kk.z < k0) + (kk.w < k0);
          }
        } else {
          for (int q = 0; q < found; q += 4) {
            const uint4 kk = sk4[q >> 2];
            r0 += (kk.x < k0) + (kk.y < k0) + (kk.z < k0) + (kk.w < k0);
            r1 += (kk.x < k1) + (kk.y < k1) + (kk.z < k1) + (kk.w < k1);
          }
        }
      } else {
        const int64_t k0 = lane < found ? skey[lane] : INT64_MAX;
        const int64_t k1 = lane + kWave < found ? skey[lane + kWave] : INT64_MAX;
        for (int q = 0; q < found; q++) {
          const int64_t kq = skey[q];
          r0 += kq < k0;
          r1 += kq < k1;
        }
      }
      const int64_t row = (int64_t)i * kNbrCap;
      if (lane < found) keys_pad[row + r0] = skey[lane];
      if (lane + kWave < found) keys_pad[row + r1] = skey[lane + kWave];
    }
    // the next atom reuses the wavefront's scratch
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Padded sorted key rows -> final CSR arrays (rel, idx, nspec, center; the
// keys are not needed past this point), with the reverse-pair index found
// by binary search in the neighbor's sorted row.  The rel vector of
// each entry is recomputed from its key by the enumeration's own test
// (image_test on the same wrapped positions and shift: the same bits).  One
// wavefront per atom; does nothing when a row or a bin overflowed.  It also zeroes
// the build's counters z[0..nz) (the bin counts and fill cursors and the
// per-block pair sums, whose last readers ran before it), so that the next
// build of the same size needs no clearing launch.
struct ZeroSpans {
  int *p[3];
  int len[3];
};

__global__ void __launch_bounds__(256)
nbr_compact_kernel(int n, WrapMats m, ZeroSpans zs, const int *__restrict__ flags,
                   const int *__restrict__ cum,
                   const int *__restrict__ count,
                   const int64_t *__restrict__ keys_pad,
                   const double *__restrict__ wrapped,
                   const int *__restrict__ species, double *__restrict__ rel_out,
                   int *__restrict__ idx_out, int *__restrict__ nspec_out,
                   int *__restrict__ center_out, int *__restrict__ rev_out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int i = blockIdx.x * 4 + threadIdx.x / kWave;
  const int gt = blockIdx.x * blockDim.x + threadIdx.x, gn = gridDim.x * blockDim.x;
#pragma unroll
  for (int z = 0; z < 3; z++)
    for (int q = gt; q < zs.len[z]; q += gn) zs.p[z][q] = 0;
  // some atom has more than kNbrCap neighbors (flags[0]): the two-pass path
  // rebuilds the arrays (the output capacity n * kNbrCap may then be too
  // small); a fixed-capacity bin overflowed (flags[2]): atoms left out of
  // the bins have no count, and the build enumerates again
  if (i >= n || flags[0] || flags[2]) return;
  const int sweep = m.sweep;
  const int b = cum[i], cnt = cum[i + 1] - b;
  const int64_t row = (int64_t)i * kNbrCap;
  const double wi[3] = {wrapped[3 * i], wrapped[3 * i + 1], wrapped[3 * i + 2]};
  for (int q = lane; q < cnt; q += kWave) {
    const int64_t k = keys_pad[row + q];
    int j, s1, s2, s3;
    key_parts(k, sweep, j, s1, s2, s3);
    const double wj[3] = {wrapped[3 * j], wrapped[3 * j + 1], wrapped[3 * j + 2]};
    double o[4];
    image_test(wi, wj, m.cell, s1, s2, s3, m.cutoff, o);
    const int64_t want = pair_key(i, -s1, -s2, -s3, sweep);
    const int dst = b + q;
    // the entry of the two partners with the smaller key looks the other
    // one up and writes both reverse indices (keys are never equal: i == j
    // needs a non-zero shift)
    if (k < want) {
      const int64_t jrow = (int64_t)j * kNbrCap;
      const int cj = count[j];
      int lo = 0, hi = cj;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys_pad[jrow + mid] < want)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < cj && keys_pad[jrow + lo] == want) {
        rev_out[dst] = cum[j] + lo;
        rev_out[cum[j] + lo] = dst;
      } else {
        rev_out[dst] = -1;
      }
    }
    reinterpret_cast<double4 *>(rel_out)[dst] = double4{o[0], o[1], o[2], o[3]};
    idx_out[dst] = j;
    nspec_out[dst] = species[j];
    center_out[dst] = i;
  }
}

// Rank sort of each atom's segment by key; writes the final CSR arrays.
__global__ void __launch_bounds__(256)
nbr_sort_kernel(int n, int sweep, const int *__restrict__ cum,
                const int64_t *__restrict__ keys_in,
                const double *__restrict__ rel_in,
                const int *__restrict__ species,
                int64_t *__restrict__ keys_out, double *__restrict__ rel_out,
                int *__restrict__ idx_out, int *__restrict__ nspec_out,
                int *__restrict__ center_out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int i = blockIdx.x * 4 + threadIdx.x / kWave;
  if (i >= n) return;
  const int b = cum[i], e = cum[i + 1];
  for (int q = b + lane; q < e; q += kWave) {
    int64_t k = keys_in[q];
    int rank = 0;
    for (int r = b; r < e; r++) rank += (keys_in[r] < k);
    int dst = b + rank;
    keys_out[dst] = k;
    int j = (int)(k >> 24);
    idx_out[dst] = j;
    nspec_out[dst] = species[j];
    center_out[dst] = i;
    for (int c = 0; c < 4; c++) rel_out[4 * dst + c] = rel_in[4 * q + c];
  }
}

// Reverse pair: (i -> j, s) <-> (j -> i, -s); both exist since the
// predicate is symmetric bitwise (fl(a-b) = -fl(b-a)).
__global__ void nbr_reverse_kernel(int64_t P, int sweep,
                                   const int *__restrict__ cum,
                                   const int64_t *__restrict__ keys,
                                   const int *__restrict__ idx,
                                   const int *__restrict__ center,
                                   int *__restrict__ rev) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  int64_t k = keys[p];
  int j, s1, s2, s3;
  key_parts(k, sweep, j, s1, s2, s3);
  int i = center[p];
  int64_t want = pair_key(i, -s1, -s2, -s3, sweep);
  int lo = cum[j], hi = cum[j + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (keys[mid] < want)
      lo = mid + 1;
    else
      hi = mid;
  }
  rev[p] = (lo < cum[j + 1] && keys[lo] == want) ? lo : -1;
}

__global__ void __launch_bounds__(kRowBlock)
species_rows_kernel(int n, const int *__restrict__ species,
                    const int *__restrict__ boff, const int *__restrict__ counts,
                    int *__restrict__ row_of, int *__restrict__ atom_of_row,
                    int *__restrict__ species_of_row) {
  species_rows_body(n, species, boff + blockIdx.x * kMaxSpecies, counts, row_of,
                    atom_of_row, species_of_row);
}

// bin_atoms / bin_pos: atoms and their wrapped positions grouped by bin
// (order within a bin is arbitrary; the neighbor rows are sorted later),
// with species_rows_kernel's rows in the same launch (both one thread per
// atom in blocks of kRowBlock).
__global__ void __launch_bounds__(kRowBlock)
bin_fill_rows_kernel(int n, const int *__restrict__ bin_of,
                     const int *__restrict__ bin_start, const double *__restrict__ wrapped,
                     int *__restrict__ fillp, int *__restrict__ bin_atoms,
                     double *__restrict__ bin_pos, const int *__restrict__ species,
                     const int *__restrict__ boff, const int *__restrict__ counts,
                     int *__restrict__ row_of, int *__restrict__ atom_of_row,
                     int *__restrict__ species_of_row) {
  const int i = blockIdx.x * kRowBlock + threadIdx.x;
  if (i < n) {
    const int b = bin_of[i];
    const int slot = bin_start[b] + atomicAdd(&fillp[b], 1);
    bin_atoms[slot] = i;
    for (int c = 0; c < 3; c++) bin_pos[3 * slot + c] = wrapped[3 * i + c];
  }
  species_rows_body(n, species, boff + blockIdx.x * kMaxSpecies, counts, row_of,
                    atom_of_row, species_of_row);
}

// exclusive prefix sum of in[0..n) into out[0..n], out[n] = total; in[n]
// must be readable and zero
static void exclusive_scan(hipStream_t st, DevBuf &tmp, const int *in, int *out,
                           int n) {
  if (n <= kScanMaxC * kScanChunk) {
    const int chunks = (n + kScanChunk - 1) / kScanChunk;
#define FLARE_SCAN(C) \
  hipLaunchKernelGGL(scan_block_kernel<C>, dim3(1), dim3(kScanThreads), 0, st, n, in, out)
    if (chunks <= 1)
      FLARE_SCAN(1);
    else if (chunks <= 2)
      FLARE_SCAN(2);
    else if (chunks <= 4)
      FLARE_SCAN(4);
    else
      FLARE_SCAN(8);
#undef FLARE_SCAN
    FLARE_HIP_CHECK(hipGetLastError());
    return;
  }
  size_t bytes = 0;
  FLARE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n + 1, st));
  tmp.ensure(bytes);
  FLARE_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, in, out, n + 1, st));
}

// cum = exclusive scan of count[0..n), then the species scan of the per-block
// counts (boff, counts[s]) and the pair total into counts[kMaxSpecies + 2]
static void count_scan_stats(hipStream_t st, NListBuffers &b, int n, int nblk,
                             const int *bcount, int *boff, int *counts) {
  if (n <= kScanMaxC * kScanChunk) {
    const int chunks = (n + kScanChunk - 1) / kScanChunk;
#define FLARE_SCAN(C)                                                                    \
  hipLaunchKernelGGL(count_scan_stats_kernel<C>, dim3(1), dim3(kScanThreads), 0, st, n,   \
                     b.count.as<int>(), b.cum.as<int>(), nblk, bcount, boff, counts,     \
                     counts + kMaxSpecies + 2)
    if (chunks <= 1)
      FLARE_SCAN(1);
    else if (chunks <= 2)
      FLARE_SCAN(2);
    else if (chunks <= 4)
      FLARE_SCAN(4);
    else
      FLARE_SCAN(8);
#undef FLARE_SCAN
    FLARE_HIP_CHECK(hipGetLastError());
    return;
  }
  exclusive_scan(st, b.scan_tmp, b.count.as<int>(), b.cum.as<int>(), n);
  hipLaunchKernelGGL(nbr_stats_kernel, dim3(1), dim3(64 * kMaxSpecies), 0, st, n,
                     b.cum.as<int>(), nblk, bcount, boff, counts, counts + kMaxSpecies + 2);
}

// ---------------------------------------------------------------------------
// Host driver.  The build is enqueued without waiting: the fused path
// writes the final arrays into capacity n * kNbrCap, and the numbers the
// host needs (species counts, bad-code flag, largest over-capacity count,
// pair total) are stored by the kernel that completes them into the pinned
// host block with the build's sequence number (no copy and no event: an
// event's system-scope release and the copy kernel cost about 10 us before
// the compaction).  nlist_resolve reads them and, in the rare case of an
// over-capacity atom, rebuilds the arrays with the two-pass path.

static WrapMats wrap_mats(const WrapHost &wh) {
  WrapMats m;
  for (int q = 0; q < 9; q++) {
    m.ct[q] = wh.ct[q];
    m.cdi[q] = wh.cdi[q];
    m.cd[q] = wh.cd[q];
    m.cti[q] = wh.cti[q];
    m.cell[q] = wh.cell[q];
  }
  for (int k = 0; k < 3; k++) m.nb[k] = wh.nb[k];
  m.sweep = wh.sweep;
  m.cutoff = wh.cutoff;
  return m;
}

// Squared radius of nbr_bin_enum_sort_kernel's FP32 distance test.  Wrapped
// positions lie in the cell, |w| <= e = |a1| + |a2| + |a3|, and the images
// tested within 2e.  Rounding positions to FP32 moves a component by at most
// 2^-24 (2e + e), the FP32 difference adds 2^-24 |d|, and the FP32 sum of
// squares a relative 3 * 2^-24: a neighbor at distance below rc tests below
// (rc + 3.1e-7 e + 1.1e-7 rc)^2 (1 + 1.8e-7).  The radius used,
// rc (1 + 1e-5) + 1e-6 (2e + rc), exceeds that several times over.
static float prefilter_cut2(const WrapHost &wh) {
  double e = 0;
  for (int k = 0; k < 3; k++)
    e += std::sqrt(wh.cell[3 * k] * wh.cell[3 * k] + wh.cell[3 * k + 1] * wh.cell[3 * k + 1] +
                   wh.cell[3 * k + 2] * wh.cell[3 * k + 2]);
  const double rc = wh.cutoff;
  const double r = rc * (1 + 1e-5) + 1e-6 * (2 * e + rc);
  return std::nextafter((float)(r * r), INFINITY);
}

// The stream is queried only after 20 ms without the summary (a build that
// failed or never ran): a query enqueues a completion marker whose
// system-scope release stalls the next kernel by about 5 us.
void nlist_wait_stats(hipStream_t st, const volatile int *stats_host, unsigned seq) {
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned spin = 1;; spin++) {
    if ((unsigned)stats_host[kNlistSeq] == seq) break;
    if ((spin & 1023) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
      const hipError_t q = hipStreamQuery(st);
      if (q == hipErrorNotReady) continue;
      FLARE_HIP_CHECK(q);
      FLARE_HIP_CHECK(hipStreamSynchronize(st));
      FLARE_REQUIRE((unsigned)stats_host[kNlistSeq] == seq,
                    "the neighbor-list build ended without publishing its summary");
      break;
    }
    __builtin_ia32_pause();
  }
  std::atomic_thread_fence(std::memory_order_acquire);
}

// The scanned bin layout: bin_start from the counts in bin_count[1..nbins],
// the species scan of the per-block counts (boff, counts[s]), atoms and
// positions grouped by bin, and the species-grouped rows.
static void scan_fill_bins(hipStream_t st, NListBuffers &b, int n, int nbins, int nblk,
                           const int *bcount, int *boff, int *counts, const int *d_species) {
  const int nblk_used = n > 0 ? nblk : 0;
  if (nbins <= kScanMaxC * kScanChunk) {
    const int chunks = (nbins + kScanChunk - 1) / kScanChunk;
#define FLARE_SCAN(C)                                                                    \
  hipLaunchKernelGGL(bin_scan_species_kernel<C>, dim3(1), dim3(kScanThreads), 0, st,      \
                     nbins, b.bin_count.as<int>() + 1, b.bin_start.as<int>(), nblk_used, \
                     bcount, boff, counts)
    if (chunks <= 1)
      FLARE_SCAN(1);
    else if (chunks <= 2)
      FLARE_SCAN(2);
    else if (chunks <= 4)
      FLARE_SCAN(4);
    else
      FLARE_SCAN(8);
#undef FLARE_SCAN
  } else {
    exclusive_scan(st, b.scan_tmp, b.bin_count.as<int>() + 1, b.bin_start.as<int>(), nbins);
    hipLaunchKernelGGL(species_scan_kernel, dim3(1), dim3(64 * kMaxSpecies), 0, st,
                       nblk_used, bcount, boff, counts);
  }
  if (n > 0)
    hipLaunchKernelGGL(bin_fill_rows_kernel, dim3((n + kRowBlock - 1) / kRowBlock),
                       dim3(kRowBlock), 0, st, n, b.bin_of.as<int>(), b.bin_start.as<int>(),
                       b.wrapped.as<double>(), b.bin_fill.as<int>(), b.bin_atoms.as<int>(),
                       b.bin_pos.as<double>(), d_species, boff, counts, b.row_of.as<int>(),
                       b.atom_of_row.as<int>(), b.species_of_row.as<int>());
}

// Per-atom neighbor rows (sorted keys in padded rows of kNbrCap, counts and
// the per-block pair sums).
static void enumerate_rows(hipStream_t st, NListBuffers &b, int n, const WrapHost &wh,
                           const WrapMats &m, const BinLayout &bl, int *overflow,
                           int *blk_sum) {
  const int nbins = wh.nb[0] * wh.nb[1] * wh.nb[2];
  const bool thin = wh.nb[0] < 3 || wh.nb[1] < 3 || wh.nb[2] < 3;
  if (thin)
    hipLaunchKernelGGL(nbr_enum_sort_kernel<true>, dim3((n + 3) / 4), dim3(256), 0, st, n,
                       m, b.wrapped.as<double>(), b.bin_of.as<int>(), bl,
                       b.bin_atoms.as<int>(), b.bin_pos.as<double>(), b.count.as<int>(),
                       b.keys_pad.as<int64_t>(), overflow, blk_sum);
  else
    hipLaunchKernelGGL(nbr_bin_enum_sort_kernel, dim3(nbins), dim3(256), 0, st, m,
                       prefilter_cut2(wh), n <= (1 << 26) ? 1 : 0, b.wrapped.as<double>(),
                       b.bin_of.as<int>(), bl, b.bin_atoms.as<int>(),
                       b.bin_pos.as<double>(), b.count.as<int>(), b.keys_pad.as<int64_t>(),
                       overflow, blk_sum);
}

// CSR offsets, the pair total and the summary (with the species rows when
// rows.species is set), then the compaction.
static void cum_compact(hipStream_t st, NListBuffers &b, int n, int nbins, const WrapMats &m,
                        int *blk_sum, int *boff, int *counts, int *stats_dev, unsigned seq,
                        const CumRows &rows, const int *d_species) {
  const int ncb = (n + kCumBlock - 1) / kCumBlock;
  if (blk_sum) {
    hipLaunchKernelGGL(cum_blocks_kernel, dim3(ncb), dim3(kCumBlock), 0, st, n,
                       b.count.as<int>(), blk_sum, b.cum.as<int>(), counts, stats_dev, seq,
                       rows);
  } else {
    count_scan_stats(st, b, n, 0, nullptr, boff, counts);
    hipLaunchKernelGGL(publish_stats_kernel, dim3(1), dim3(64), 0, st, counts, stats_dev,
                       seq);
  }
  if (n > 0) {
    // every entry's reverse index is written by the partner with the
    // smaller key (the predicate is symmetric bitwise, so both exist)
    ZeroSpans zs{{b.bin_count.as<int>(), b.bin_fill.as<int>(), blk_sum},
                 {nbins + 2, nbins + 1, blk_sum ? ncb * kCumStride : 0}};
    hipLaunchKernelGGL(nbr_compact_kernel, dim3((n + 3) / 4), dim3(256), 0, st, n, m, zs,
                       counts + kMaxSpecies + 1, b.cum.as<int>(), b.count.as<int>(),
                       b.keys_pad.as<int64_t>(), b.wrapped.as<double>(), d_species,
                       b.rel.as<double>(), b.idx.as<int>(), b.nspec.as<int>(),
                       b.center.as<int>(), b.rev.as<int>());
  }
}

void nlist_build_async(hipStream_t st, NListBuffers &b, int n, const double *d_pos,
                       const int *d_species, const WrapHost &wh, int *stats_dev,
                       unsigned seq) {
  const WrapMats m = wrap_mats(wh);
  const int nbins = wh.nb[0] * wh.nb[1] * wh.nb[2];
  const int nblk = std::max(1, (n + kRowBlock - 1) / kRowBlock);
  const size_t cap = (size_t)std::max(n, 1) * kNbrCap;
  // cum from per-block pair sums (cum_blocks_kernel) up to kCumMaxBlocks
  // blocks, else the single-workgroup / device-wide count scan
  const int ncb = (n + kCumBlock - 1) / kCumBlock;
  const bool cum_blocks = n > 0 && ncb <= kCumMaxBlocks;
  // Fixed-capacity bins (placed by the wrap kernel, no bin scan or fill
  // launch, rows formed by cum_blocks_kernel) where at least three bins lie
  // along every axis and the slots stay within a few per atom.
  const bool thin = wh.nb[0] < 3 || wh.nb[1] < 3 || wh.nb[2] < 3;
  const bool fcb = cum_blocks && !thin && (int64_t)nbins * kBinSlots <= 8LL * n + 4096;
  const size_t slots = std::max<size_t>(std::max(n, 1),
                                        fcb ? (size_t)nbins * kBinSlots : (size_t)0);

  b.wrapped.ensure((size_t)std::max(n, 1) * 3 * sizeof(double));
  b.bin_of.ensure((size_t)std::max(n, 1) * sizeof(int));
  b.bin_count.ensure((size_t)(nbins + 2) * sizeof(int));
  b.bin_start.ensure((size_t)(nbins + 1) * sizeof(int));
  b.bin_fill.ensure((size_t)(nbins + 1) * sizeof(int));
  b.bin_atoms.ensure(slots * sizeof(int));
  b.bin_pos.ensure(slots * 3 * sizeof(double));
  b.count.ensure((size_t)(n + 1) * sizeof(int));
  b.cum.ensure((size_t)(n + 1) * sizeof(int));
  b.keys_pad.ensure(cap * sizeof(int64_t));
  b.rel.ensure(cap * 4 * sizeof(double));
  b.idx.ensure(cap * sizeof(int));
  b.nspec.ensure(cap * sizeof(int));
  b.center.ensure(cap * sizeof(int));
  b.rev.ensure(cap * sizeof(int));
  b.row_of.ensure((size_t)std::max(n, 1) * sizeof(int));
  b.atom_of_row.ensure((size_t)std::max(n, 1) * sizeof(int));
  b.species_of_row.ensure((size_t)std::max(n, 1) * sizeof(int));
  // counts[0..8): species counts, [8]: bad species code, [9]: largest
  // over-capacity neighbor count, [10]: pair total, [11]: a fixed-capacity
  // bin overflowed.  Two blocks used in turn: a build zeroes the other
  // block's accumulated ints [8..11].
  b.species_count.ensure(2 * kNlistStats * sizeof(int));
  b.species_blk.ensure((size_t)2 * nblk * kMaxSpecies * sizeof(int));
  int *bcount = b.species_blk.as<int>(), *boff = bcount + nblk * kMaxSpecies;
  int *counts = b.species_count.as<int>() + b.parity * kNlistStats;
  int *next_counts = b.species_count.as<int>() + (1 - b.parity) * kNlistStats;
  int *overflow = counts + kMaxSpecies + 1;
  b.blk_sum.ensure((size_t)std::max(ncb, 1) * kCumStride * sizeof(int));
  int *blk_sum = cum_blocks ? b.blk_sum.as<int>() : nullptr;
  // The counters start at zero: either the previous build's compaction left
  // them zeroed (same buffers, at least as many bins and blocks) and the
  // previous build's wrap zeroed this summary block's flags, or one launch
  // clears them.  count[n] is the zero the count scan reads past the end;
  // blk_sum the enumeration's per-block sums.
  const bool ready = b.zero_ready && cum_blocks && n > 0 && b.zr_bins >= nbins &&
                     b.zr_ncb >= ncb && b.zr_ptr[0] == b.bin_count.ptr &&
                     b.zr_ptr[1] == b.bin_fill.ptr && b.zr_ptr[2] == b.blk_sum.ptr &&
                     b.zr_ptr[3] == b.species_count.ptr;
  // set again once the whole build is enqueued: a build that stops part way
  // (an error) leaves the counters in an unknown state
  b.zero_ready = false;
  if (!ready) {
    const int nclear = std::max(std::max(nbins + 2, kNlistStats), ncb * kCumStride);
    hipLaunchKernelGGL(clear4_kernel, dim3((nclear + 255) / 256), dim3(256), 0, st,
                       counts, kNlistStats, b.bin_count.as<int>(), nbins + 2,
                       b.bin_fill.as<int>(), nbins + 1,
                       cum_blocks ? blk_sum : b.count.as<int>() + n,
                       cum_blocks ? ncb * kCumStride : 1);
  }
  const int g1 = (n + kRowBlock - 1) / kRowBlock;
  BinLayout bl{b.bin_start.as<int>(), nullptr, 0};
  CumRows rows;
  if (fcb) {
    hipLaunchKernelGGL(wrap_place_kernel, dim3(g1), dim3(kRowBlock), 0, st, n, d_pos, m,
                       d_species, b.wrapped.as<double>(), b.bin_of.as<int>(),
                       b.bin_count.as<int>(), bcount, counts + kMaxSpecies,
                       next_counts + kMaxSpecies, b.bin_atoms.as<int>(),
                       b.bin_pos.as<double>(), counts + kMaxSpecies + 3);
    bl = BinLayout{nullptr, b.bin_count.as<int>() + 1, kBinSlots};
    rows.species = d_species;
    rows.bcount = bcount;
    rows.row_of = b.row_of.as<int>();
    rows.atom_of_row = b.atom_of_row.as<int>();
    rows.species_of_row = b.species_of_row.as<int>();
  } else {
    if (n > 0)
      hipLaunchKernelGGL(wrap_bin_kernel, dim3(g1), dim3(kRowBlock), 0, st, n, d_pos, m,
                         d_species, b.wrapped.as<double>(), b.bin_of.as<int>(),
                         b.bin_count.as<int>(), bcount, counts + kMaxSpecies,
                         next_counts + kMaxSpecies);
    // bin_count[1..nbins] holds the counts; bin_count has one spare zero
    // entry.  The species scan of the per-block counts rides on the bin
    // scan, so the species rows are written with the bin fill.
    scan_fill_bins(st, b, n, nbins, nblk, bcount, boff, counts, d_species);
  }
  if (n > 0) enumerate_rows(st, b, n, wh, m, bl, overflow, blk_sum);
  cum_compact(st, b, n, nbins, m, blk_sum, boff, counts, stats_dev, seq, rows, d_species);
  FLARE_HIP_CHECK(hipGetLastError());
  // n > 0: the wrap kernel zeroed the other summary block's flags; with the
  // per-block sums the compaction zeroed this build's counters
  b.zero_ready = cum_blocks && n > 0;
  b.zr_bins = nbins;
  b.zr_ncb = ncb;
  b.zr_ptr[0] = b.bin_count.ptr;
  b.zr_ptr[1] = b.bin_fill.ptr;
  b.zr_ptr[2] = b.blk_sum.ptr;
  b.zr_ptr[3] = b.species_count.ptr;
  b.fcb = fcb;
  b.counts_block = counts;
  b.parity ^= 1;
}

int nlist_resolve(hipStream_t st, NListBuffers &b, int n, const int *d_species,
                  const WrapHost &wh, const volatile int *stats_host, int *stats_dev,
                  unsigned *seq, int *host_counts) {
  int stats[kNlistStats];
  for (int k = 0; k < kNlistStats; k++) stats[k] = stats_host[k];
  FLARE_REQUIRE(stats[kMaxSpecies] == 0,
                "species codes must lie in [0, " + std::to_string(kMaxSpecies) + ")");
  if (n > 0 && stats[kMaxSpecies + 3]) {
    // A fixed-capacity bin held more atoms than its slots: group the atoms
    // again in the scanned layout, from the build's wrapped positions and
    // bins (the caller's positions are not read again), and enumerate again.
    // The compaction left the bin counts, fill cursors and block sums zeroed.
    const WrapMats m = wrap_mats(wh);
    const int nbins = wh.nb[0] * wh.nb[1] * wh.nb[2];
    const int nblk = std::max(1, (n + kRowBlock - 1) / kRowBlock);
    int *counts = b.counts_block;
    int *bcount = b.species_blk.as<int>(), *boff = bcount + nblk * kMaxSpecies;
    int *blk_sum = b.blk_sum.as<int>();
    // bin counts again; the summary's overflow, pair-total and bin-overflow
    // ints cleared
    hipLaunchKernelGGL(count_bins_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n,
                       b.bin_of.as<int>(), b.bin_count.as<int>(), counts + kMaxSpecies + 1,
                       3);
    scan_fill_bins(st, b, n, nbins, nblk, bcount, boff, counts, d_species);
    enumerate_rows(st, b, n, wh, m, BinLayout{b.bin_start.as<int>(), nullptr, 0},
                   counts + kMaxSpecies + 1, blk_sum);
    *seq += 1;
    if (*seq == 0) *seq = 1;
    cum_compact(st, b, n, nbins, m, blk_sum, boff, counts, stats_dev, *seq, CumRows(),
                d_species);
    FLARE_HIP_CHECK(hipGetLastError());
    b.fcb = false;
    nlist_wait_stats(st, stats_host, *seq);
    for (int k = 0; k < kNlistStats; k++) stats[k] = stats_host[k];
    FLARE_REQUIRE(stats[kMaxSpecies + 3] == 0, "neighbor build: bin regrouping failed");
  }
  const int P = stats[kMaxSpecies + 2];
  b.P = P;
  if (n > 0 && stats[kMaxSpecies + 1] > kNbrCap) {
    // two-pass path for atoms with more neighbors than the LDS rows hold
    const WrapMats m = wrap_mats(wh);
    const size_t Pa = std::max(P, 1);
    b.keys.ensure(Pa * sizeof(int64_t));
    b.rel.ensure(Pa * 4 * sizeof(double));
    b.idx.ensure(Pa * sizeof(int));
    b.nspec.ensure(Pa * sizeof(int));
    b.center.ensure(Pa * sizeof(int));
    b.rev.ensure(Pa * sizeof(int));
    b.keys_tmp.ensure(Pa * sizeof(int64_t));
    b.rel_tmp.ensure(Pa * 4 * sizeof(double));
    const int gw = (n + 3) / 4;
    const BinLayout bl = b.fcb ? BinLayout{nullptr, b.bin_count.as<int>() + 1, kBinSlots}
                               : BinLayout{b.bin_start.as<int>(), nullptr, 0};
    hipLaunchKernelGGL(nbr_enum_kernel<1>, dim3(gw), dim3(256), 0, st, n, m,
                       b.wrapped.as<double>(), b.bin_of.as<int>(), bl,
                       b.bin_atoms.as<int>(), (int *)nullptr, b.cum.as<int>(),
                       b.keys_tmp.as<int64_t>(), b.rel_tmp.as<double>());
    hipLaunchKernelGGL(nbr_sort_kernel, dim3(gw), dim3(256), 0, st, n,
                       wh.sweep, b.cum.as<int>(), b.keys_tmp.as<int64_t>(),
                       b.rel_tmp.as<double>(), d_species, b.keys.as<int64_t>(),
                       b.rel.as<double>(), b.idx.as<int>(), b.nspec.as<int>(),
                       b.center.as<int>());
    if (P > 0) {
      const int gp = (P + 255) / 256;
      hipLaunchKernelGGL(nbr_reverse_kernel, dim3(gp), dim3(256), 0, st,
                         (int64_t)P, wh.sweep, b.cum.as<int>(),
                         b.keys.as<int64_t>(), b.idx.as<int>(),
                         b.center.as<int>(), b.rev.as<int>());
    }
    FLARE_HIP_CHECK(hipGetLastError());
  }
  int ns = 0;
  for (int sp = 0; sp < kMaxSpecies; sp++) {
    host_counts[sp] = stats[sp];
    if (stats[sp] > 0) ns = sp + 1;
  }
  return ns;
}


// ---------------------------------------------------------------------------
// Caller-provided neighbor lists.  LAMMPS hands pair styles a full list over
// local and ghost atoms; pair_flare keeps the neighbors with
// r^2 < rc(si,sj)^2 in list order (pair_flare.cpp:95-118,
// lammps_descriptor.cpp:51-66; the Kokkos short list of
// pair_flare_kokkos.cpp:633-662).  One wavefront per centre; ballots keep
// the list order when the kept pairs are compacted.

constexpr int kNeighMask = 0x1FFFFFFF;  // LAMMPS NEIGHMASK

struct ExtCut {
  int S;
  double rc2[kMaxSpecies * kMaxSpecies];
};

// neighbor jj of centre atom i: index, type and the pair test
__device__ inline bool ext_pair(const ExtList &e, const ExtCut &c, int i, int ti,
                                int jj, int &j, int &tj, double *o, int *bad) {
#pragma clang fp contract(off)
  j = e.neighbors[(int64_t)i * e.stride_i + (int64_t)jj * e.stride_j] & kNeighMask;
  if (j < 0 || j >= e.n_all) {
    atomicOr(bad, 2);
    return false;
  }
  tj = e.type[j] - 1;
  if (tj < 0 || tj >= c.S) {
    atomicOr(bad, 1);
    return false;
  }
  const double dx = e.x[3 * (int64_t)j] - e.x[3 * (int64_t)i];
  const double dy = e.x[3 * (int64_t)j + 1] - e.x[3 * (int64_t)i + 1];
  const double dz = e.x[3 * (int64_t)j + 2] - e.x[3 * (int64_t)i + 2];
  const double rsq = dx * dx + dy * dy + dz * dz;
  o[0] = sqrt(rsq);
  o[1] = dx;
  o[2] = dy;
  o[3] = dz;
  return rsq < c.rc2[ti * c.S + tj];
}

__global__ void __launch_bounds__(256)
ext_count_kernel(ExtList e, ExtCut c, int *__restrict__ cnt,
                 int *__restrict__ cspec, int *__restrict__ bad) {
  const int lane = threadIdx.x & (kWave - 1);
  const int ii = blockIdx.x * 4 + threadIdx.x / kWave;
  if (ii >= e.inum) return;
  const int i = e.ilist[ii];
  int ti = -1;
  if (i >= 0 && i < e.n_all) ti = e.type[i] - 1;
  if (ti < 0 || ti >= c.S) {
    if (lane == 0) {
      atomicOr(bad, (i >= 0 && i < e.n_all) ? 1 : 2);
      cnt[ii] = 0;
      cspec[ii] = 0;
    }
    return;
  }
  const int jnum = e.numneigh[i];
  int total = 0;
  for (int jj0 = 0; jj0 < jnum; jj0 += kWave) {
    const int jj = jj0 + lane;
    bool in = false;
    if (jj < jnum) {
      int j, tj;
      double o[4];
      in = ext_pair(e, c, i, ti, jj, j, tj, o, bad);
    }
    total += __popcll(__ballot(in));
  }
  if (lane == 0) {
    cnt[ii] = total;
    cspec[ii] = ti;
  }
}

__global__ void __launch_bounds__(256)
ext_fill_kernel(ExtList e, ExtCut c, const int *__restrict__ cum,
                const int *__restrict__ cspec, int *__restrict__ idx,
                int *__restrict__ nspec, double *__restrict__ rel,
                int *__restrict__ center, int *__restrict__ bad) {
  const int lane = threadIdx.x & (kWave - 1);
  const int ii = blockIdx.x * 4 + threadIdx.x / kWave;
  if (ii >= e.inum) return;
  const int i = e.ilist[ii];
  if (i < 0 || i >= e.n_all) return;
  const int ti = cspec[ii];
  const int jnum = e.numneigh[i];
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  int pos = cum[ii];
  for (int jj0 = 0; jj0 < jnum; jj0 += kWave) {
    const int jj = jj0 + lane;
    bool in = false;
    int j = 0, tj = 0;
    double o[4];
    if (jj < jnum) in = ext_pair(e, c, i, ti, jj, j, tj, o, bad);
    const uint64_t m = __ballot(in);
    if (in) {
      const int k = pos + __popcll(m & lt);
      idx[k] = j;
      nspec[k] = tj;
      reinterpret_cast<double4 *>(rel)[k] = double4{o[0], o[1], o[2], o[3]};
      center[k] = ii;
    }
    pos += __popcll(m);
  }
}

// per-block species counts of the centres (the input of nbr_stats_kernel /
// species_rows_kernel)
__global__ void __launch_bounds__(kRowBlock)
ext_species_kernel(int inum, const int *__restrict__ cspec, int *__restrict__ bcount) {
  __shared__ int cnt[kMaxSpecies];
  const int t = threadIdx.x;
  if (t < kMaxSpecies) cnt[t] = 0;
  __syncthreads();
  const int ii = blockIdx.x * kRowBlock + t;
  if (ii < inum) atomicAdd(&cnt[cspec[ii]], 1);
  __syncthreads();
  if (t < kMaxSpecies) bcount[blockIdx.x * kMaxSpecies + t] = cnt[t];
}

int64_t nlist_from_external(hipStream_t st, NListBuffers &b, const ExtList &e,
                            const double *cut_host, int S, int *cspec,
                            int *host_counts) {
  FLARE_REQUIRE(S >= 1 && S <= kMaxSpecies, "n_species must be in [1, 8]");
  ExtCut c;
  c.S = S;
  for (int q = 0; q < kMaxSpecies * kMaxSpecies; q++) c.rc2[q] = 0;
  for (int q = 0; q < S * S; q++) c.rc2[q] = cut_host[q] * cut_host[q];
  const int n = e.inum;
  const int nblk = std::max(1, (n + kRowBlock - 1) / kRowBlock);
  b.count.ensure((size_t)(n + 1) * sizeof(int));
  b.cum.ensure((size_t)(n + 1) * sizeof(int));
  b.row_of.ensure((size_t)std::max(n, 1) * sizeof(int));
  b.atom_of_row.ensure((size_t)std::max(n, 1) * sizeof(int));
  b.species_of_row.ensure((size_t)std::max(n, 1) * sizeof(int));
  b.species_count.ensure(2 * kNlistStats * sizeof(int));
  b.zero_ready = false;  // this path uses the counters (and bin_count as scratch)
  b.species_blk.ensure((size_t)2 * nblk * kMaxSpecies * sizeof(int));
  int *bcount = b.species_blk.as<int>(), *boff = bcount + nblk * kMaxSpecies;
  int *counts = b.species_count.as<int>();
  int *bad = counts + kMaxSpecies;
  hipLaunchKernelGGL(clear4_kernel, dim3(1), dim3(256), 0, st, counts, kNlistStats,
                     b.count.as<int>() + n, 1, (int *)nullptr, 0, (int *)nullptr, 0);
  const int gw = (n + 3) / 4;
  if (n > 0) {
    hipLaunchKernelGGL(ext_count_kernel, dim3(gw), dim3(256), 0, st, e, c,
                       b.count.as<int>(), cspec, bad);
    hipLaunchKernelGGL(ext_species_kernel, dim3(nblk), dim3(kRowBlock), 0, st, n,
                       cspec, bcount);
  }
  count_scan_stats(st, b, n, n > 0 ? nblk : 0, bcount, boff, counts);
  int stats[kNlistStats];
  FLARE_HIP_CHECK(hipMemcpyAsync(stats, counts, kNlistStats * sizeof(int),
                                 hipMemcpyDeviceToHost, st));
  FLARE_HIP_CHECK(hipStreamSynchronize(st));
  FLARE_REQUIRE(!(stats[kMaxSpecies] & 1), "atom types must lie in [1, n_species]");
  FLARE_REQUIRE(!(stats[kMaxSpecies] & 2), "neighbor or centre index outside [0, n_all)");
  const int64_t P = stats[kMaxSpecies + 2];
  b.P = (int)P;
  const size_t Pa = (size_t)std::max<int64_t>(P, 1);
  b.idx.ensure(Pa * sizeof(int));
  b.nspec.ensure(Pa * sizeof(int));
  b.center.ensure(Pa * sizeof(int));
  b.rev.ensure(Pa * sizeof(int));
  b.rel.ensure(Pa * 4 * sizeof(double));
  if (P > 0)
    FLARE_HIP_CHECK(hipMemsetAsync(b.rev.ptr, 0xff, P * sizeof(int), st));
  if (n > 0) {
    hipLaunchKernelGGL(ext_fill_kernel, dim3(gw), dim3(256), 0, st, e, c,
                       b.cum.as<int>(), cspec, b.idx.as<int>(), b.nspec.as<int>(),
                       b.rel.as<double>(), b.center.as<int>(), bad);
    hipLaunchKernelGGL(species_rows_kernel, dim3(nblk), dim3(kRowBlock), 0, st, n,
                       cspec, boff, counts, b.row_of.as<int>(),
                       b.atom_of_row.as<int>(), b.species_of_row.as<int>());
  }
  FLARE_HIP_CHECK(hipGetLastError());
  for (int sp = 0; sp < kMaxSpecies; sp++) host_counts[sp] = stats[sp];
  return P;
}

__global__ void ext_hist_kernel(int64_t P, const int *__restrict__ idx,
                                int *__restrict__ cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < P) atomicAdd(&cnt[idx[k]], 1);
}

__global__ void iota_kernel(int64_t P, int *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < P) out[k] = (int)k;
}

void index_by_target(hipStream_t st, DevBuf &work, DevBuf &cnt_buf, DevBuf &scan_tmp,
                     const int *keys, int64_t P, int n_all, int *in_cum,
                     int *in_pairs) {
  // counts (n_all + 1, last zero) -> in_cum; stable radix sort of the pair
  // indices by key
  cnt_buf.ensure((size_t)(n_all + 1) * sizeof(int));
  int *cnt = cnt_buf.as<int>();
  FLARE_HIP_CHECK(hipMemsetAsync(cnt, 0, (size_t)(n_all + 1) * sizeof(int), st));
  const size_t Pa = (size_t)std::max<int64_t>(P, 1);
  work.ensure(Pa * 2 * sizeof(int));  // iota values, then sorted keys
  int *vals = work.as<int>();
  int *keys_sorted = vals + Pa;
  if (P > 0) {
    const int g = (int)((P + 255) / 256);
    hipLaunchKernelGGL(ext_hist_kernel, dim3(g), dim3(256), 0, st, P, keys, cnt);
    hipLaunchKernelGGL(iota_kernel, dim3(g), dim3(256), 0, st, P, vals);
    int bits = 1;
    while (bits < 31 && (1 << bits) < n_all) bits++;
    size_t bytes = 0;
    FLARE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
        nullptr, bytes, keys, keys_sorted, vals, in_pairs, (int)P, 0, bits, st));
    scan_tmp.ensure(bytes);
    FLARE_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
        scan_tmp.ptr, bytes, keys, keys_sorted, vals, in_pairs, (int)P, 0, bits, st));
  }
  exclusive_scan(st, scan_tmp, cnt, in_cum, n_all);
  FLARE_HIP_CHECK(hipGetLastError());
}

void ext_incoming(hipStream_t st, NListBuffers &b, int n_all, int64_t P,
                  int *in_cum, int *in_pairs) {
  index_by_target(st, b.keys_tmp, b.bin_count, b.scan_tmp, b.idx.as<int>(), P, n_all,
                  in_cum, in_pairs);
}

template <typename PF>
__global__ void __launch_bounds__(256)
ext_incoming_kernel(int n_all, const int *__restrict__ in_cum,
                    const int *__restrict__ in_pairs, const PF *__restrict__ pf,
                    double *__restrict__ forces) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_all) return;
  double f0 = 0, f1 = 0, f2 = 0;
  for (int q = in_cum[a]; q < in_cum[a + 1]; q++) {
    const int64_t k = in_pairs[q];
    f0 += (double)pf[3 * k];
    f1 += (double)pf[3 * k + 1];
    f2 += (double)pf[3 * k + 2];
  }
  forces[3 * (int64_t)a] -= f0;
  forces[3 * (int64_t)a + 1] -= f1;
  forces[3 * (int64_t)a + 2] -= f2;
}

__global__ void __launch_bounds__(256)
ext_own_kernel(int inum, const int *__restrict__ ilist, const double *__restrict__ own,
               const int *__restrict__ row_of, const double *__restrict__ E_row,
               double *__restrict__ forces, double *__restrict__ eatom) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= inum) return;
  const int64_t a = ilist[c];
  forces[3 * a] += own[3 * (int64_t)c];
  forces[3 * a + 1] += own[3 * (int64_t)c + 1];
  forces[3 * a + 2] += own[3 * (int64_t)c + 2];
  if (eatom) eatom[a] += E_row[row_of[c]];
}

void ext_assemble(hipStream_t st, int n_all, int inum, const int *ilist,
                  const int *in_cum, const int *in_pairs, const PairForces &pf,
                  const double *own, const int *row_of, const double *E_row,
                  double *forces, double *eatom) {
  if (n_all > 0) {
    if (pf.f32)
      hipLaunchKernelGGL(ext_incoming_kernel<float>, dim3((n_all + 255) / 256), dim3(256),
                         0, st, n_all, in_cum, in_pairs, (const float *)pf.p, forces);
    else
      hipLaunchKernelGGL(ext_incoming_kernel<double>, dim3((n_all + 255) / 256), dim3(256),
                         0, st, n_all, in_cum, in_pairs, (const double *)pf.p, forces);
  }
  if (inum > 0)
    hipLaunchKernelGGL(ext_own_kernel, dim3((inum + 255) / 256), dim3(256), 0, st,
                       inum, ilist, own, row_of, E_row, forces, eatom);
  FLARE_HIP_CHECK(hipGetLastError());
}

}  // namespace flare
