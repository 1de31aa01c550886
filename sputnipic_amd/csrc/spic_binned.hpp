// spic_binned.hpp -- argument blocks of the binned phase kernel, its
// per-node reduction and the binned sort / store.  Internal to libspic.so.
//
// Binned particle storage of one species (one per buffer of a pair):
//   slots [seg[t], seg[t+1])      tile t's segment;
//   slots [seg[ntiles], +ovf_cap)  overflow region, for particles whose
//                                  destination segment was full.
// Per buffer, one array of 3 * (ntiles + 1) counters:
//   cnt[t]                 slots of range t taken (t == ntiles: the overflow
//                          region); may exceed the range's capacity, the
//                          excess went to the overflow region;
//   cnt[NT1 + t], cnt[2 * NT1 + t]   (NT1 = ntiles + 1) the start and length
//                          of one unused run inside the taken slots: the
//                          rest of the block a tile's workgroup reserved for
//                          the particles that stayed in it.
// Range t holds min(cnt[t], capacity) - length particles, at the taken slots
// outside the unused run.  Both buffers of a pair share seg; a phase reads
// one and writes the other.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "spic_internal.hpp"

namespace spic {

constexpr int kMaxSpecies = 8;
constexpr int kBucketCap = 7;  // LDS bucket slots per cell and batch
constexpr int kListCap = 128;  // LDS list ring entries
constexpr int kMaxInstances = 16;  // engines in one batched launch
constexpr int kMaxTileCells = 128;  // cells of the largest tile shape (8x4x4)

// Blocks of a batched launch: instance i owns blocks [start[i], start[i+1]),
// i < n; start[n] is the grid size.
struct BatchMap {
  int start[kMaxInstances + 1];
  int n;
};

// The argument blocks hold generic pointers.  Read from a device table (the
// batched launches) the compiler cannot tell that they address global
// memory and would emit flat instructions, which also count against the
// LDS wait counter; device code accesses them through gmem(), which states
// the address space, and gatomic_add() (relaxed, device scope, as atomicAdd).
#define SPIC_GLOBAL __attribute__((address_space(1)))

template <typename P>
__device__ __forceinline__ SPIC_GLOBAL P *gmem(P *p) {
  return (SPIC_GLOBAL P *)p;
}

template <typename P, typename V>
__device__ __forceinline__ P gatomic_add(P *p, V v) {
  return __hip_atomic_fetch_add(gmem(p), (P)v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}

// The particles of range t (a tile's segment, or t == ntiles: the overflow
// region) as a dense sequence: count, and the slot of the r-th.
struct RangeView {
  int64_t first;     // slot of the range's start
  int64_t count;     // particles in the range
  uint32_t gap_lo;   // unused run [gap_lo, gap_lo + gap_len) relative to first
  uint32_t gap_len;
  __host__ __device__ int64_t slot(int64_t r) const {
    return first + (r < (int64_t)gap_lo ? r : r + gap_len);
  }
};

template <typename PS, typename PC>
__host__ __device__ inline RangeView range_view(PS seg, PC cnt, int ntiles,
                                                unsigned int ovf_cap, int t) {
  RangeView v;
  v.first = seg[t];
  const int64_t cap = t < ntiles ? seg[t + 1] - seg[t] : (int64_t)ovf_cap;
  const int64_t taken = (int64_t)cnt[t] < cap ? (int64_t)cnt[t] : cap;
  v.gap_lo = cnt[(ntiles + 1) + t];
  v.gap_len = cnt[2 * (ntiles + 1) + t];
  v.count = taken - (int64_t)v.gap_len;
  return v;
}

// Fixed-point particle positions (FP32 engines in production arithmetic):
// the x y z words of a slot hold int32 coordinates in units of 2^-k[a]
// cells from the grid's start, X = (x - xStart) / dx * 2^k, so that the
// cell is X >> k and the offset within it (X - (cell << k)) * 2^-k, exact;
// k is chosen per axis so that every cell of the grid plus 64 cells of
// excursion fit in int32 (k <= 24, so offsets convert to float exactly).
// Wraps and reflections at the box faces are exact integer operations.
struct FixedParams {
  int k[3];
  int lim[3];     // box length per axis: nc << k
  float unit[3];  // 2^-k
  float s[3];     // 2^k / dx: fixed units per unit length
  double to_fixed[3];    // the same in double, for conversions from caller arrays
  double from_fixed[3];  // dx * 2^-k, back to lengths
  double start[3];       // xStart yStart zStart
};

// Storage type of the tile partials: the particle precision.  An FP32
// engine sums its LDS block in FP64 by default and rounds each tile's block
// to float once; the node reduction adds the tiles' floats in FP64.
template <typename T>
using PartialT = T;

template <typename T>
struct BinnedSpecies {
  // x y z u v w q, binned: component c of slot i at in[c * stride + i]
  const T *in;
  T *out;
  int64_t stride;
  const uint32_t *id_in;  // load index per slot; both null when the engine
  uint32_t *id_out;       // keeps no load order
  const int64_t *seg;          // [ntiles + 1]
  const unsigned int *cnt_in;  // [3 * (ntiles + 1)], see above
  unsigned int *cnt_out;       // [3 * (ntiles + 1)], zero on entry
  unsigned int ovf_cap;
  KPush<T> k;
  KMom m;
};

template <typename T>
struct BinnedArgs {
  KGrid<T> g;
  KField<T> f;
  int ncx, ncy, ncz;       // physical cells per axis
  int ntx, nty, ntz, ntiles;
  int tx, ty, tz;          // cells per tile per axis
  int ns;
  int lds_fp64;            // FP32 particles: LDS moment block in FP64
  int n_ovf_blocks;        // workgroups for the overflow regions
  int push_off;            // 1: deposit only (interpP2G); no push, no field read
  int fixed;               // 1: positions in fixed point (FP32 engines), see above
  int ppt;                 // particles per thread and batch of the phase kernel: 1, or
                           // 2 (FP32 fixed point, single engines)
  FixedParams fx;
  BinnedSpecies<T> sp[kMaxSpecies];
  // status words: [0] mover iterate out of the grid, [1] deposit cell out
  // of the grid, [2] overflow region full
  unsigned int *status;
  // cumulative counters: [0] stencil gathers from global memory, [1]
  // deposits with global atomics, [2] stencil gathers from LDS, [3]
  // deposits through LDS (bucket or LDS atomics)
  unsigned long long *counters;
  // per-tile moment partials [tile][species][block node][moment] in the
  // particle precision (PartialT<T>), written for every tile-species whose
  // input segment is not empty
  void *partials;
  // [tile][species]: 1 when the tile wrote partials for the species this
  // phase (its input segment was not empty), else 0
  unsigned char *pflag;
};

struct BinnedReduceArgs {
  int nyn, nzn;
  int ntx, nty, ntz, ntiles;
  int ns;
  int partials_fp32;  // partials stored as float (FP32 engines), else double
  const void *partials;
  const unsigned char *pflag;  // see BinnedArgs::pflag
  KMom mom[kMaxSpecies];
};

struct CountArrays {
  unsigned int *p[kMaxSpecies];
  int n;  // entries per array
  int ns;
  // nn > 0: also zero the ten moment arrays of nn nodes of each species
  KMom mom[kMaxSpecies];
  int64_t nn;
};

// Work list of a batch phase's persistent launch: one 32-bit item per tile
// of an instance or per part of its overflow regions, instance in the top
// four bits.
constexpr int kQueueInstShift = 28;
constexpr uint32_t kQueueOverflowBit = 1u << 27;
constexpr uint32_t kQueueIndexMask = kQueueOverflowBit - 1u;
constexpr int kQueueOverflowParts = 64;  // overflow parts per instance
static_assert(kMaxInstances <= (1 << (32 - kQueueInstShift)), "instance field too narrow");

#define SPIC_DECLARE_BINNED_LAUNCHERS(NS)                                  \
  namespace NS {                                                           \
  template <typename T>                                                    \
  hipError_t launch_binned_phase(const BinnedArgs<T> &a, hipStream_t s);   \
  template <typename T>                                                    \
  hipError_t launch_binned_phase_queue(const BinnedArgs<T> *tab,           \
                                       const uint32_t *items,              \
                                       uint32_t nitems, unsigned int *ctr, \
                                       int tx, int ty, int tz,             \
                                       int lds_fp64, int fixed,            \
                                       hipStream_t s);                     \
  }
SPIC_DECLARE_BINNED_LAUNCHERS(spic_fast)
SPIC_DECLARE_BINNED_LAUNCHERS(spic_strict)

// Staged block nodes per tile: (tx+3)(ty+3)(tz+3).
inline int block_nodes(int tx, int ty, int tz) {
  return (tx + 3) * (ty + 3) * (tz + 3);
}

// Tile geometry of one grid, the sort key's fields and the slack of the
// segments.  Sort key (32 bits, most significant first):
//   tile | rank of the particle within its cell | cell within the tile.
struct BinLayout {
  int ncx, ncy, ncz;  // physical cells
  int tx, ty, tz;     // cells per tile
  int ntx, nty, ntz;  // tiles per axis
  int ntiles;
  int tile_bits, rank_bits, cell_bits;
  int64_t slack;      // free slots per segment after a sort
  int fixed;          // binned positions are fixed point (FixedParams fx)
  FixedParams fx;
};

// Fixed-point parameters of a grid: per axis the largest k <= 24 with
// (cells + 64) << k below 2^31.
inline FixedParams fixed_params(int ncx, int ncy, int ncz, double dx, double dy, double dz,
                                double xs, double ys, double zs) {
  FixedParams f;
  const int nc[3] = {ncx, ncy, ncz};
  const double d[3] = {dx, dy, dz}, st[3] = {xs, ys, zs};
  for (int a = 0; a < 3; ++a) {
    int k = 24;
    while (k > 0 && ((int64_t)nc[a] + 64) << k > (int64_t)0x7fffffff) --k;
    f.k[a] = k;
    f.lim[a] = nc[a] << k;
    const double sc = (double)(1 << k);
    f.unit[a] = (float)(1.0 / sc);
    f.to_fixed[a] = sc / d[a];
    f.s[a] = (float)f.to_fixed[a];
    f.from_fixed[a] = d[a] / sc;
    f.start[a] = st[a];
  }
  return f;
}

// Sums the tile partials of every species into its node moments (+=).
hipError_t launch_binned_reduce(const BinnedReduceArgs &a, int tx, int ty, int tz,
                                int nn, hipStream_t s);
// Zeroes ns count arrays of n entries each (and the moments, c.nn > 0).
hipError_t zero_counts(const CountArrays &c, hipStream_t s);

// The reduce and the count reset of a batch, from the device-resident
// argument blocks tab[i] of its instances: per instance, the node blocks
// nodes.start[i].. of 256 nodes each (reduce) and the reset blocks
// counts.start[i].. of 256 threads each, thread j resetting counter j of
// every species' cnt_out (3 * (ntiles + 1) entries) and, with `moments`,
// node j of its ten moment arrays; max_ns species per launch row.
template <typename T>
hipError_t launch_batch_reduce(const BinnedArgs<T> *tab, const BatchMap &nodes,
                               int max_ns, int tx, int ty, int tz, hipStream_t s);
template <typename T>
hipError_t launch_batch_zero_counts(const BinnedArgs<T> *tab, const BatchMap &counts,
                                    int max_ns, int moments, hipStream_t s);

// Sort of n particles of one species into binned storage (spic_binsort.hip).
// The source is dense (src_seg == nullptr: particles 0..n-1 of the caller's
// arrays, FP64 when src_f64 (an FP32 engine rounds them on the way in),
// ids assigned as their index; with fixed-point storage the positions are
// converted) or binned (the engine's own storage: src_seg, src_cnt,
// src_ovf_cap; ids from src_id).  The destination gets segments of count +
// slack slots:
// dst_seg[t] = (particles in tiles < t) + t * slack, dst_cnt the counts
// (no unused runs, an empty overflow region).  With workspace == nullptr
// only *needed is set.
template <typename T>
hipError_t bin_sort(const KGrid<T> &g, const BinLayout &L, int64_t n,
                    const void *const src[7], int src_f64, const uint32_t *src_id,
                    const int64_t *src_seg, const unsigned int *src_cnt,
                    unsigned int src_ovf_cap, T *const dst[7], uint32_t *dst_id,
                    int64_t *dst_seg, unsigned int *dst_cnt, void *workspace,
                    size_t workspace_bytes, size_t *needed, hipStream_t s);

// dst[id[i]] = src[i] for every particle in binned storage, seven arrays,
// into FP64 arrays (dst_f64) or FP32 ones (FP32 engines only); fixed-point
// positions are converted back to lengths.  Without ids (id == nullptr) the
// particles are written densely in storage order, range t from base[t]
// (range_bases).
template <typename T>
hipError_t bin_store(const BinLayout &L, T *const src[7], const uint32_t *id,
                     const uint32_t *base, const int64_t *seg, const unsigned int *cnt,
                     unsigned int ovf_cap, void *const dst[7], int dst_f64, hipStream_t s);

// The first dense position of every range (tiles, then the overflow
// region) of a binned species: an exclusive scan of its range counts into
// the workspace (*bases points there).  With workspace == nullptr only
// *needed is set.
hipError_t range_bases(const int64_t *seg, const unsigned int *cnt, unsigned int ovf_cap,
                       int ntiles, void *workspace, size_t workspace_bytes, size_t *needed,
                       const uint32_t **bases, hipStream_t s);

// Tile sort of one binned species (a job) from buffer src (layout
// src_seg) to buffer dst (layout dst_seg): the binned source's alternative
// to bin_sort, by counting within tiles instead of a radix sort over all
// particles.  The destination gets bin_sort's segments (count + L.slack
// slots each, an empty overflow region); each tile holds its own particles
// (those that were in its segment and are in its cells) in round-robin
// cell order -- bin_sort's order -- followed by the particles that came
// from other segments or the overflow region.  Component c of slot i at
// base[c * stride + i].  count, own and fill: workspace of ntiles + 1 words
// each, zeroed by the launchers.
template <typename T>
struct TileSortJob {
  KGrid<T> g;
  BinLayout L;
  const T *src;
  T *dst;
  int64_t stride;
  const uint32_t *src_id;
  uint32_t *dst_id;
  const int64_t *src_seg;
  int64_t *dst_seg;
  const unsigned int *src_cnt;
  unsigned int *dst_cnt;
  unsigned int ovf_cap;
  unsigned int *count, *own, *fill;
};

// Workgroups of the count and sort launches per job: one per tile, plus
// kTileOvfBlocks over the overflow region.
constexpr int kTileOvfBlocks = 64;

// Jobs of a batched tile sort: job j owns blocks start[j] .. start[j+1]-1
// of the count and sort launches.
struct TileSortMap {
  int start[kMaxInstances * kMaxSpecies + 1];
  int n;
};

// One job; count, own and fill must be one block of 3 * (ntiles + 1)
// words starting at count.
template <typename T>
hipError_t tile_sort(const TileSortJob<T> &job, hipStream_t s);

// map.n jobs from the device table `jobs`, all of one tile shape, whose
// count / own / fill arrays all lie in ws[0 .. ws_words).
template <typename T>
hipError_t tile_sort_batch(const TileSortJob<T> *jobs, const TileSortMap &map, int tx,
                           int ty, int tz, unsigned int *ws, size_t ws_words,
                           hipStream_t s);

// bin_store of x y z u v w only, as big-endian IEEE float32 words (values
// rounded to float): the particle datasets of a snapshot.
template <typename T>
hipError_t bin_store_be32(const BinLayout &L, T *const src[7], const uint32_t *id,
                          const uint32_t *base, const int64_t *seg, const unsigned int *cnt,
                          unsigned int ovf_cap, uint32_t *const dst[6], hipStream_t s);

}  // namespace spic
