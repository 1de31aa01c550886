// spic_tiled.hpp -- argument blocks of the tile-resident phase kernel and
// the tile sort.  Internal to libspic.so.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "spic_internal.hpp"

namespace spic {

constexpr int kMaxSpecies = 8;

// One species of the tiled phase: particle arrays in tile order, the tile
// segment table (tile t holds particles [tile_start[t], tile_start[t+1])),
// its push constants and its FP64 node moments.
template <typename T>
struct TiledSpecies {
  T *x, *y, *z, *u, *v, *w;
  const T *q;
  const int64_t *tile_start;
  KPush<T> k;
  KMom m;
};

template <typename T>
struct TiledArgs {
  KGrid<T> g;
  KField<T> f;
  int ntx, nty, ntz;  // tiles per axis
  int tx, ty, tz;     // cells per tile per axis
  int ns;
  int lds_fp64;       // FP32 particles: accumulate LDS partials in FP64
  TiledSpecies<T> sp[kMaxSpecies];
  unsigned int *status;
  // cumulative counters: [0] predictor-corrector gathers served from global
  // memory (stencil outside the staged block), [1] deposits added with
  // global atomics, [2] gathers served from LDS, [3] deposits into LDS
  unsigned long long *counters;
  // per-tile moment partials: [tile][species][moment][block node], FP64;
  // written by the phase kernel for every tile-species with particles and
  // summed per node by the reduction kernel
  double *partials;
};

#define SPIC_DECLARE_TILED_LAUNCHERS(NS)                                   \
  namespace NS {                                                           \
  template <typename T>                                                    \
  hipError_t launch_tiled_phase(const TiledArgs<T> &a, int tiles,          \
                                hipStream_t s);                            \
  template <typename T>                                                    \
  hipError_t launch_tile_reduce(const TiledArgs<T> &a, int nn,             \
                                hipStream_t s);                            \
  }
SPIC_DECLARE_TILED_LAUNCHERS(spic_fast)
SPIC_DECLARE_TILED_LAUNCHERS(spic_strict)

// Arguments of the per-node reduction of the tile partials.
struct TiledArgsBase {
  int nyn, nzn;
  int ntx, nty, ntz;
  int ns;
  const double *partials;
  const int64_t *tile_start[kMaxSpecies];
  KMom mom[kMaxSpecies];
};

// Staged block nodes per tile: (tx+3)(ty+3)(tz+3).
inline int block_nodes(int tx, int ty, int tz) {
  return (tx + 3) * (ty + 3) * (tz + 3);
}

// Tile geometry of one grid.
struct TileLayout {
  int ncx, ncy, ncz;  // physical cells
  int tx, ty, tz;     // cells per tile
  int ntx, nty, ntz;  // tiles per axis
  int ntiles;
  int cell_bits, rank_bits, tile_bits;
};

// Sort of one species into tile order with cells interleaved inside each
// tile (spic_tilesort.hip).  Particles in `src` are permuted into `dst`
// (8 arrays: x y z u v w q id); tile_start receives ntiles + 1 offsets.
// `cell_count` is scratch of ncx*ncy*ncz ints.  With workspace == nullptr
// only *needed is set.
template <typename T>
hipError_t tile_sort(const KGrid<T> &g, const TileLayout &L, int n,
                     T *const src[7], const uint32_t *src_id, T *const dst[7],
                     uint32_t *dst_id, int64_t *tile_start, void *workspace,
                     size_t workspace_bytes, size_t *needed, hipStream_t s);

hipError_t scatter_by_id_f64(const double *src, const uint32_t *id, double *dst,
                             int n, hipStream_t s);
hipError_t scatter_by_id_f32(const float *src, const uint32_t *id, float *dst,
                             int n, hipStream_t s);
hipError_t iota_u32(uint32_t *p, int n, hipStream_t s);

}  // namespace spic
