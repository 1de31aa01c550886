// spic_binned.hpp -- argument blocks of the binned phase kernel, its
// per-node reduction and the binned sort / store.  Internal to libspic.so.
//
// Binned particle storage of one species (one per buffer of a pair):
//   slots [seg[t], seg[t+1])      tile t's segment: its particles first,
//                                  then free slots (slack);
//   slots [seg[ntiles], +ovf_cap)  overflow region, for particles whose
//                                  destination segment was full;
//   cnt[t], t < ntiles             particles written to tile t (may exceed
//                                  the segment's capacity; the excess went
//                                  to the overflow region);
//   cnt[ntiles]                    particles in the overflow region.
// Both buffers of a pair share seg; the phase reads one and writes the other.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "spic_internal.hpp"

namespace spic {

constexpr int kMaxSpecies = 8;
constexpr int kBucketCap = 7;  // LDS bucket slots per cell and batch
constexpr int kListCap = 128;  // LDS list ring entries

template <typename T>
struct BinnedSpecies {
  const T *in[7];  // x y z u v w q, binned
  T *out[7];
  const uint32_t *id_in;
  uint32_t *id_out;
  const int64_t *seg;          // [ntiles + 1]
  const unsigned int *cnt_in;  // [ntiles + 1]
  unsigned int *cnt_out;       // [ntiles + 1], zero on entry
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
  BinnedSpecies<T> sp[kMaxSpecies];
  // status words: [0] mover iterate out of the grid, [1] deposit cell out
  // of the grid, [2] overflow region full
  unsigned int *status;
  // cumulative counters: [0] stencil gathers from global memory, [1]
  // deposits with global atomics, [2] stencil gathers from LDS, [3]
  // deposits through LDS (bucket or LDS atomics)
  unsigned long long *counters;
  // per-tile moment partials [tile][species][moment][block node], FP64,
  // written for every tile-species whose input segment is not empty
  double *partials;
};

struct BinnedReduceArgs {
  int nyn, nzn;
  int ntx, nty, ntz;
  int ns;
  const double *partials;
  const unsigned int *cnt[kMaxSpecies];
  KMom mom[kMaxSpecies];
};

struct CountArrays {
  unsigned int *p[kMaxSpecies];
  int n;  // entries per array
  int ns;
};

#define SPIC_DECLARE_BINNED_LAUNCHERS(NS)                                  \
  namespace NS {                                                           \
  template <typename T>                                                    \
  hipError_t launch_binned_phase(const BinnedArgs<T> &a, hipStream_t s);   \
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
};

// Sums the tile partials of every species into its node moments (+=).
hipError_t launch_binned_reduce(const BinnedReduceArgs &a, int tx, int ty, int tz,
                                int nn, hipStream_t s);
// Zeroes ns count arrays of n entries each.
hipError_t zero_counts(const CountArrays &c, hipStream_t s);

// Sort of n particles of one species into binned storage (spic_binsort.hip).
// The source is dense (src_seg == nullptr: particles 0..n-1, ids assigned
// as their index) or binned (src_seg, src_cnt, src_ovf_cap; ids from
// src_id).  The destination gets segments of count + slack slots:
// dst_seg[t] = (particles in tiles < t) + t * slack, dst_cnt the counts,
// dst_cnt[ntiles] = 0.  With workspace == nullptr only *needed is set.
template <typename T>
hipError_t bin_sort(const KGrid<T> &g, const BinLayout &L, int64_t n,
                    T *const src[7], const uint32_t *src_id,
                    const int64_t *src_seg, const unsigned int *src_cnt,
                    unsigned int src_ovf_cap, T *const dst[7], uint32_t *dst_id,
                    int64_t *dst_seg, unsigned int *dst_cnt, void *workspace,
                    size_t workspace_bytes, size_t *needed, hipStream_t s);

// dst[id[i]] = src[i] for every particle in binned storage, seven arrays.
template <typename T>
hipError_t bin_store(const BinLayout &L, T *const src[7], const uint32_t *id,
                     const int64_t *seg, const unsigned int *cnt,
                     unsigned int ovf_cap, T *const dst[7], hipStream_t s);

}  // namespace spic
