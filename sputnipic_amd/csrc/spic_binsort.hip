// spic_binsort.hip -- binned particle storage: the sort that (re)builds it,
// the store back to the caller's order, the per-node reduction of the
// phase kernel's tile partials, and the count reset before each phase.
//
// Sort.  A particle's cell is its deposit cell 2 + floor((p - start) * inv)
// minus the two guard offsets, clamped to the physical cells; its rank is
// a per-cell counter taken with an atomic.  The key
//   tile | rank within the cell | cell within the tile
// orders each tile's particles round-robin over its cells (every cell's
// first particle, then every cell's second, ...), so that any run of
// 4 * cells-per-tile consecutive particles holds about four per cell: the
// phase kernel takes one such run per batch into its per-cell buckets, and
// particles that stay in their tile keep this order from phase to phase.
// A stable LSD radix sort (rocPRIM) of (key, source slot) pairs gives the
// permutation; the seven arrays and the particle id are gathered to sorted
// position i + tile * slack, so that every segment gets `slack` free slots
// after its particles.  A binned source is enumerated densely through an
// exclusive scan of its counts.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "spic_binned.hpp"

namespace spic {
namespace {

inline unsigned nblocks(long n, int tpb = 256) {
  long b = (n + tpb - 1) / tpb;
  if (b < 1) b = 1;
  if (b > (1L << 20)) b = 1L << 20;
  return (unsigned)b;
}

inline size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

template <typename TS, typename T>
__device__ __forceinline__ int phys_cell(TS p, double start, T inv, int nc) {
  double c = floor(((double)p - start) * (double)inv);
  c = fmin(fmax(c, 0.0), (double)(nc - 1));
  return (int)c;
}

// Fixed-point positions (FixedParams): a length from the caller's arrays
// to the int32 coordinate along axis a, rounded to nearest and bounded to
// the box plus 32 cells either side (a position farther out is reported
// by the phase as out of the grid); and back.
__device__ __forceinline__ int to_fixed_pos(double p, const FixedParams &f, int a) {
  const double ext = (double)(32 << f.k[a]);
  double X = rint((p - f.start[a]) * f.to_fixed[a]);
  X = fmin(fmax(X, -ext), (double)f.lim[a] + ext);
  return (int)X;
}

__device__ __forceinline__ double from_fixed_pos(int X, const FixedParams &f, int a) {
  return f.start[a] + (double)X * f.from_fixed[a];
}

// Particles per range of a binned source (tiles, then the overflow region).
__global__ void k_range_counts(const int64_t *__restrict__ seg,
                               const unsigned int *__restrict__ cnt,
                               unsigned int ovf_cap, int ntiles,
                               unsigned int *__restrict__ out) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles;
       t += gridDim.x * blockDim.x)
    out[t] = (unsigned int)range_view(seg, cnt, ntiles, ovf_cap, t).count;
}


// Rank of this lane's particle among the active lanes of its wave that
// share `key`, plus the count already in counter[key]: one atomic per
// distinct key in the wave (few when the input is in cell order).
__device__ __forceinline__ uint32_t wave_rank(unsigned int *counter, uint32_t key,
                                              bool active) {
  const int lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long todo = __ballot(active);
  uint32_t rank = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const uint32_t lkey = __shfl(key, leader);
    const unsigned long long same = __ballot(active && key == lkey);
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(&counter[lkey], (unsigned int)__popcll(same));
    base = __shfl(base, leader);
    if (active && key == lkey) rank = base + (uint32_t)__popcll(same & below);
    todo &= ~same;
  }
  return rank;
}

// Key of one particle: tile | rank of the particle within its cell | cell
// within the tile.  `binned`: the position comes from engine storage, in
// fixed point when the layout says so (else from the caller's arrays).
template <typename TS, typename T>
__device__ __forceinline__ void cell_of(const KGrid<T> &g, const BinLayout &L, TS x, TS y,
                                        TS z, int &cx, int &cy, int &cz, uint32_t &tile,
                                        uint32_t &cit, bool binned = true) {
  if constexpr (sizeof(TS) == 4) {
    if (binned && L.fixed) {
      cx = min(max(__float_as_int((float)x) >> L.fx.k[0], 0), L.ncx - 1);
      cy = min(max(__float_as_int((float)y) >> L.fx.k[1], 0), L.ncy - 1);
      cz = min(max(__float_as_int((float)z) >> L.fx.k[2], 0), L.ncz - 1);
    } else {
      cx = phys_cell(x, g.xStart, g.invdx, L.ncx);
      cy = phys_cell(y, g.yStart, g.invdy, L.ncy);
      cz = phys_cell(z, g.zStart, g.invdz, L.ncz);
    }
  } else {
    cx = phys_cell(x, g.xStart, g.invdx, L.ncx);
    cy = phys_cell(y, g.yStart, g.invdy, L.ncy);
    cz = phys_cell(z, g.zStart, g.invdz, L.ncz);
  }
  const int tx = cx / L.tx, ty = cy / L.ty, tz = cz / L.tz;
  tile = (uint32_t)((tx * L.nty + ty) * L.ntz + tz);
  cit = (uint32_t)(((cx - tx * L.tx) * L.ty + (cy - ty * L.ty)) * L.tz + (cz - tz * L.tz));
}

// Keys of a dense source (particles 0..n-1 in the caller's order, element
// type TS), ranks from per-cell counters with wave-aggregated atomics.
template <typename TS, typename T>
__global__ void k_bin_keys(KGrid<T> g, BinLayout L, int64_t n,
                           const TS *__restrict__ x, const TS *__restrict__ y,
                           const TS *__restrict__ z, unsigned int *__restrict__ cell_count,
                           uint32_t *__restrict__ key, uint32_t *__restrict__ idx) {
  const uint32_t rank_max = (1u << L.rank_bits) - 1u;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t n_up = (n + 63) / 64 * 64;  // whole waves stay in the loop
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_up; i += stride) {
    const bool active = i < n;
    int cx = 0, cy = 0, cz = 0;
    uint32_t tile = 0, cit = 0;
    if (active) cell_of(g, L, x[i], y[i], z[i], cx, cy, cz, tile, cit, false);
    const uint32_t cell = (uint32_t)(((int64_t)cx * L.ncy + cy) * L.ncz + cz);
    uint32_t r = wave_rank(cell_count, cell, active);
    if (!active) continue;
    r = r < rank_max ? r : rank_max;
    key[i] = (tile << (L.rank_bits + L.cell_bits)) | (r << L.cell_bits) | cit;
    idx[i] = (uint32_t)i;
  }
}

// Keys of a binned source: block t takes range t (a tile's segment, or
// t == ntiles: the overflow region), whose particles land at dense
// positions P[t] + r.  A tile's own particles take their rank from LDS
// counters of its cells; others (overflow region) from the global ones.
// Equal ranks of one cell from the two sources only make equal keys.
template <typename T>
__global__ __launch_bounds__(256) void k_bin_keys_binned(
    KGrid<T> g, BinLayout L, const T *__restrict__ x, const T *__restrict__ y,
    const T *__restrict__ z, const unsigned int *__restrict__ P,
    const int64_t *__restrict__ seg, const unsigned int *__restrict__ cnt,
    unsigned int ovf_cap, unsigned int *__restrict__ cell_count, uint32_t *__restrict__ key,
    uint32_t *__restrict__ idx) {
  __shared__ unsigned int s_cnt[kMaxTileCells];
  const uint32_t rank_max = (1u << L.rank_bits) - 1u;
  const int t = blockIdx.x;
  const int ncells = L.tx * L.ty * L.tz;
  for (int c = threadIdx.x; c < ncells; c += blockDim.x) s_cnt[c] = 0u;
  __syncthreads();
  const RangeView v = range_view(seg, cnt, L.ntiles, ovf_cap, t);
  const int64_t base = P[t];
  for (int64_t r = threadIdx.x; r < v.count; r += blockDim.x) {
    const int64_t j = v.slot(r);
    int cx, cy, cz;
    uint32_t tile, cit;
    cell_of(g, L, x[j], y[j], z[j], cx, cy, cz, tile, cit);
    uint32_t rk;
    if ((int)tile == t) rk = atomicAdd(&s_cnt[cit], 1u);
    else rk = atomicAdd(&cell_count[((size_t)cx * L.ncy + cy) * L.ncz + cz], 1u);
    rk = rk < rank_max ? rk : rank_max;
    key[base + r] = (tile << (L.rank_bits + L.cell_bits)) | (rk << L.cell_bits) | cit;
    idx[base + r] = (uint32_t)j;
  }
}

__global__ void k_zero_u32(unsigned int *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

template <typename T>
struct Arrays7 {
  T *p[7];
};

struct Words6 {
  uint32_t *p[6];
};

// The sorted gather into binned storage.  A dense source (src_id null)
// holds the caller's values in TS: with fixed-point storage (to_fixed) its
// positions are converted, the rest rounded to T; a binned source is the
// engine's own storage, copied word for word.
template <typename TS, typename T>
__global__ void k_bin_gather(Arrays7<TS> src, const uint32_t *__restrict__ src_id,
                             const uint32_t *__restrict__ perm,
                             const uint32_t *__restrict__ keys, int shift, int64_t n,
                             int64_t slack, Arrays7<T> dst, uint32_t *__restrict__ dst_id,
                             int to_fixed, FixedParams fx) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t j = perm[i];
    const int64_t o = i + (int64_t)(keys[i] >> shift) * slack;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if constexpr (sizeof(T) == 4) {
        if (a < 3 && to_fixed) {
          dst.p[a][o] = __int_as_float(to_fixed_pos((double)src.p[a][j], fx, a));
          continue;
        }
      }
      dst.p[a][o] = (T)src.p[a][j];
    }
    if (dst_id) dst_id[o] = src_id ? src_id[j] : j;
  }
}

// seg[t] = (first sorted position of tile >= t) + t * slack; cnt from the
// differences, no unused runs; the overflow region starts empty.
__global__ void k_bin_bounds(const uint32_t *__restrict__ keys, int shift, int64_t n,
                             int ntiles, int64_t slack, int64_t *__restrict__ seg,
                             unsigned int *__restrict__ cnt) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles;
       t += gridDim.x * blockDim.x) {
    int64_t b[2];
    for (int k = 0; k < 2; ++k) {
      const int64_t tt = t + k;
      int64_t lo = 0, hi = n;
      while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if ((int64_t)(keys[mid] >> shift) < tt) lo = mid + 1;
        else hi = mid;
      }
      b[k] = lo;
    }
    seg[t] = b[0] + (int64_t)t * slack;
    cnt[t] = t < ntiles ? (unsigned int)(b[1] - b[0]) : 0u;
    cnt[(ntiles + 1) + t] = 0u;
    cnt[2 * (ntiles + 1) + t] = 0u;
  }
}

// One workgroup per segment (then the overflow region): dst[id] = src, in
// the caller's element type TD; fixed-point positions converted back.
template <typename T, typename TD>
__global__ void k_bin_store(Arrays7<T> src, const uint32_t *__restrict__ id,
                            const uint32_t *__restrict__ base,
                            const int64_t *__restrict__ seg,
                            const unsigned int *__restrict__ cnt,
                            unsigned int ovf_cap, int ntiles, Arrays7<TD> dst, int fixed,
                            FixedParams fx) {
  const RangeView v = range_view(seg, cnt, ntiles, ovf_cap, (int)blockIdx.x);
  for (int64_t r = threadIdx.x; r < v.count; r += blockDim.x) {
    const int64_t i = v.slot(r);
    const uint32_t k = id ? id[i] : base[blockIdx.x] + (uint32_t)r;
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      if constexpr (sizeof(T) == 4) {
        if (a < 3 && fixed) {
          dst.p[a][k] = (TD)from_fixed_pos(__float_as_int((float)src.p[a][i]), fx, a);
          continue;
        }
      }
      dst.p[a][k] = (TD)src.p[a][i];
    }
  }
}

// Big-endian IEEE float32 bits of v rounded to float (H5T_IEEE_F32BE).
template <typename T>
__device__ __forceinline__ uint32_t be32(T v) {
  return __builtin_bswap32(__float_as_uint((float)v));
}

// bin_store's scatter to load order, writing x y z u v w as big-endian
// float32: the datasets of a particle snapshot.
template <typename T>
__global__ void k_bin_store_be32(Arrays7<T> src, const uint32_t *__restrict__ id,
                                 const uint32_t *__restrict__ base,
                                 const int64_t *__restrict__ seg,
                                 const unsigned int *__restrict__ cnt, unsigned int ovf_cap,
                                 int ntiles, Words6 dst, int fixed, FixedParams fx) {
  const RangeView v = range_view(seg, cnt, ntiles, ovf_cap, (int)blockIdx.x);
  for (int64_t r = threadIdx.x; r < v.count; r += blockDim.x) {
    const int64_t i = v.slot(r);
    const uint32_t k = id ? id[i] : base[blockIdx.x] + (uint32_t)r;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      if constexpr (sizeof(T) == 4) {
        if (a < 3 && fixed) {
          dst.p[a][k] = be32(from_fixed_pos(__float_as_int((float)src.p[a][i]), fx, a));
          continue;
        }
      }
      dst.p[a][k] = be32(src.p[a][i]);
    }
  }
}

// Per species (blockIdx.y): thread i zeroes counter i (i < n) and, with
// c.nn > 0, node i of the ten moment arrays (i < nn).
__global__ void k_zero_counts(CountArrays c) {
  const int s = blockIdx.y;
  if (s >= c.ns) return;
  const int64_t total = c.n > c.nn ? c.n : c.nn;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < c.n) gmem(c.p[s])[i] = 0u;
    if (i < c.nn) {
#pragma unroll
      for (int m = 0; m < 10; ++m) gmem(c.mom[s].m[m])[i] = 0.0;
    }
  }
}

// Per node and species: the sum of the partials of every tile whose staged
// block holds the node (at most two tiles per axis), in tile order, with
// Neumaier's compensated summation, added to the FP64 moment arrays.
// Tiles whose input segment held no particle of the species wrote no
// partials (their flag is 0) and are skipped.
template <typename P, int TX, int TY, int TZ>
__device__ __forceinline__ void reduce_node(int n, int s, int nyn, int nzn, int ntx, int nty,
                                            int ntz, int ntiles, int ns,
                                            const P *__restrict__ partials,
                                            const unsigned char *__restrict__ pflag,
                                            double *const *mom) {
  constexpr int BX = TX + 3, BY = TY + 3, BZ = TZ + 3, NB = BX * BY * BZ;
  const int gz = n % nzn, gy = (n / nzn) % nyn, gx = n / (nzn * nyn);
  const int xl = gx >= BX - 1 ? (gx - (BX - 1) + TX - 1) / TX : 0;
  const int xh = min(gx / TX, ntx - 1);
  const int yl = gy >= BY - 1 ? (gy - (BY - 1) + TY - 1) / TY : 0;
  const int yh = min(gy / TY, nty - 1);
  const int zl = gz >= BZ - 1 ? (gz - (BZ - 1) + TZ - 1) / TZ : 0;
  const int zh = min(gz / TZ, ntz - 1);
  double sum[10], cmp[10];
#pragma unroll
  for (int m = 0; m < 10; ++m) { sum[m] = 0.0; cmp[m] = 0.0; }
  bool any = false;
  for (int tx = xl; tx <= xh; ++tx)
    for (int ty = yl; ty <= yh; ++ty)
      for (int tz = zl; tz <= zh; ++tz) {
        const int tile = (tx * nty + ty) * ntz + tz;
        if (!gmem(pflag)[(size_t)tile * ns + s]) continue;
        any = true;
        const int b = ((gx - tx * TX) * BY + (gy - ty * TY)) * BZ + (gz - tz * TZ);
        const SPIC_GLOBAL P *src =
            gmem(partials) + ((size_t)tile * ns + s) * (10 * NB) + (size_t)b * 10;
#pragma unroll
        for (int m = 0; m < 10; ++m) {
          const double v = (double)src[m];
          const double t = sum[m] + v;
          cmp[m] += fabs(sum[m]) >= fabs(v) ? (sum[m] - t) + v : (v - t) + sum[m];
          sum[m] = t;
        }
      }
  if (!any) return;
#pragma unroll
  for (int m = 0; m < 10; ++m) gmem(mom[m])[n] += sum[m] + cmp[m];
}

template <typename P, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_binned_reduce(BinnedReduceArgs a, int nn) {
  const int s = blockIdx.y;
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= nn) return;
  reduce_node<P, TX, TY, TZ>(n, s, a.nyn, a.nzn, a.ntx, a.nty, a.ntz, a.ntiles, a.ns,
                             (const P *)a.partials, a.pflag, a.mom[s].m);
}

// The reduce of a batch: blockIdx.y is the species.
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_binned_reduce_batch(
    const BinnedArgs<T> *__restrict__ tab, BatchMap map) {
  int i = 0;
  for (int k = 1; k < map.n; ++k) i += (int)blockIdx.x >= map.start[k] ? 1 : 0;
  const BinnedArgs<T> &a = tab[i];
  const int s = blockIdx.y;
  const int n = (blockIdx.x - map.start[i]) * 256 + threadIdx.x;
  if (s >= a.ns || n >= a.g.nxn * a.g.nyn * a.g.nzn) return;
  reduce_node<PartialT<T>, TX, TY, TZ>(n, s, a.g.nyn, a.g.nzn, a.ntx, a.nty, a.ntz, a.ntiles,
                                       a.ns, (const PartialT<T> *)a.partials, a.pflag,
                                       a.sp[s].m.m);
}

// The count reset of a batch (blockIdx.y is the species): thread j of
// instance i zeroes counter j and, when `moments` is set, node j of the ten
// moment arrays.
template <typename T>
__global__ __launch_bounds__(256) void k_zero_counts_batch(
    const BinnedArgs<T> *__restrict__ tab, BatchMap map, int moments) {
  int i = 0;
  for (int k = 1; k < map.n; ++k) i += (int)blockIdx.x >= map.start[k] ? 1 : 0;
  const BinnedArgs<T> &a = tab[i];
  const int s = blockIdx.y;
  if (s >= a.ns) return;
  const int64_t j = (int64_t)(blockIdx.x - map.start[i]) * 256 + threadIdx.x;
  if (j < 3 * (int64_t)(a.ntiles + 1)) gmem(a.sp[s].cnt_out)[j] = 0u;
  if (moments && j < (int64_t)a.g.nxn * a.g.nyn * a.g.nzn) {
#pragma unroll
    for (int m = 0; m < 10; ++m) gmem(a.sp[s].m.m[m])[j] = 0.0;
  }
}

}  // namespace

hipError_t launch_binned_reduce(const BinnedReduceArgs &a, int tx, int ty, int tz,
                                int nn, hipStream_t s) {
  if (nn <= 0 || a.ns <= 0) return hipSuccess;
  dim3 grid((nn + 255) / 256, a.ns);
  const bool f = a.partials_fp32 != 0;
  if (tx == 4 && ty == 4 && tz == 4) {
    if (f) hipLaunchKernelGGL((k_binned_reduce<float, 4, 4, 4>), grid, dim3(256), 0, s, a, nn);
    else hipLaunchKernelGGL((k_binned_reduce<double, 4, 4, 4>), grid, dim3(256), 0, s, a, nn);
  } else if (tx == 8 && ty == 4 && tz == 4) {
    if (f) hipLaunchKernelGGL((k_binned_reduce<float, 8, 4, 4>), grid, dim3(256), 0, s, a, nn);
    else hipLaunchKernelGGL((k_binned_reduce<double, 8, 4, 4>), grid, dim3(256), 0, s, a, nn);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_reduce(const BinnedArgs<T> *tab, const BatchMap &nodes,
                               int max_ns, int tx, int ty, int tz, hipStream_t s) {
  if (nodes.n <= 0 || nodes.start[nodes.n] <= 0 || max_ns <= 0) return hipSuccess;
  dim3 grid(nodes.start[nodes.n], max_ns);
  if (tx == 4 && ty == 4 && tz == 4)
    hipLaunchKernelGGL((k_binned_reduce_batch<T, 4, 4, 4>), grid, dim3(256), 0, s, tab,
                       nodes);
  else if (tx == 8 && ty == 4 && tz == 4)
    hipLaunchKernelGGL((k_binned_reduce_batch<T, 8, 4, 4>), grid, dim3(256), 0, s, tab,
                       nodes);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_zero_counts(const BinnedArgs<T> *tab, const BatchMap &counts,
                                    int max_ns, int moments, hipStream_t s) {
  if (counts.n <= 0 || counts.start[counts.n] <= 0 || max_ns <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_zero_counts_batch<T>), dim3(counts.start[counts.n], max_ns),
                     dim3(256), 0, s, tab, counts, moments);
  return hipGetLastError();
}

template hipError_t launch_batch_reduce<double>(const BinnedArgs<double> *,
                                                const BatchMap &, int, int, int, int,
                                                hipStream_t);
template hipError_t launch_batch_reduce<float>(const BinnedArgs<float> *, const BatchMap &,
                                               int, int, int, int, hipStream_t);
template hipError_t launch_batch_zero_counts<double>(const BinnedArgs<double> *,
                                                     const BatchMap &, int, int,
                                                     hipStream_t);
template hipError_t launch_batch_zero_counts<float>(const BinnedArgs<float> *,
                                                    const BatchMap &, int, int, hipStream_t);

// Copy particle j of the source to slot o of the destination.
template <typename PS, typename PSI, typename PD, typename PDI>
__device__ __forceinline__ void copy_particle(PS src, PSI src_id, int64_t stride, int64_t j,
                                              PD dst, PDI dst_id, int64_t o) {
#pragma unroll
  for (int a = 0; a < 7; ++a) dst[a * stride + o] = src[a * stride + j];
  if (dst_id) dst_id[o] = src_id[j];  // engines that keep no ids have neither
}

constexpr int kTileThreads = 512;
constexpr unsigned int kTileStaged = 12288;  // particles a tile block stages in LDS
constexpr int kTileStageBytes = 32768;       // one output chunk of one component
constexpr int kTileU = 4;                    // particles a thread loads at once
constexpr int kScanThreads = 1024;

// wave_rank on a global counter array (global atomics, not flat ones).
__device__ __forceinline__ uint32_t wave_rank_g(SPIC_GLOBAL unsigned int *counter,
                                                uint32_t key, bool active) {
  const int lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long todo = __ballot(active);
  uint32_t rank = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const uint32_t lkey = __shfl(key, leader);
    const unsigned long long same = __ballot(active && key == lkey);
    unsigned int base = 0;
    if (lane == leader)
      base = __hip_atomic_fetch_add(counter + lkey, (unsigned int)__popcll(same),
                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    base = __shfl(base, leader);
    if (active && key == lkey) rank = base + (uint32_t)__popcll(same & below);
    todo &= ~same;
  }
  return rank;
}

// Job of block b in a tile-sort map (uniform across the workgroup).
__device__ __forceinline__ int sort_job(const TileSortMap &map, int b) {
  int j = 0;
  for (int k = 1; k < map.n; ++k) j += b >= map.start[k] ? 1 : 0;
  return j;
}

// Tile sort, step 1 (block b of nb of the job).  Block t < ntiles reads
// segment t: its particles that lie in tile t are tile t's own (own[t],
// added to count[t]); every other particle adds one to its tile's count.
// Blocks ntiles and up split the overflow region, whose particles all add
// one to their tile's count.
// BATCH: the launch covers every job of a map; the block's job, taken from
// the table, replaces the pointer and layout arguments.
template <typename T, bool BATCH>
__global__ __launch_bounds__(kTileThreads) void k_tile_count(
    KGrid<T> g, BinLayout L, const T *src_a, int64_t stride, const int64_t *seg_a,
    const unsigned int *cnt_a, unsigned int ovf_cap, unsigned int *count_a,
    unsigned int *own_a, const TileSortJob<T> *__restrict__ jobs, TileSortMap map) {
  __shared__ unsigned int s_own;
  int b = blockIdx.x;
  int nb = gridDim.x;
  if constexpr (BATCH) {
    const int j = sort_job(map, blockIdx.x);
    const TileSortJob<T> &J = jobs[j];
    g = J.g; L = J.L; src_a = J.src; stride = J.stride; seg_a = J.src_seg;
    cnt_a = J.src_cnt; ovf_cap = J.ovf_cap; count_a = J.count; own_a = J.own;
    b -= map.start[j];
    nb = map.start[j + 1] - map.start[j];
  }
  const SPIC_GLOBAL T *__restrict__ src = gmem(src_a);
  SPIC_GLOBAL unsigned int *__restrict__ count = gmem(count_a);
  const int tid = threadIdx.x;
  const bool ovf = b >= L.ntiles;
  const RangeView v = range_view(gmem(seg_a), gmem(cnt_a), L.ntiles, ovf_cap,
                                 ovf ? L.ntiles : b);
  const int64_t first = ovf ? (int64_t)(b - L.ntiles) * kTileThreads : 0;
  const int64_t step = ovf ? (int64_t)(nb - L.ntiles) * kTileThreads : (int64_t)kTileThreads;
  if (tid == 0) s_own = 0u;
  __syncthreads();
  unsigned int mine = 0;
  for (int64_t base = first; base < v.count; base += kTileU * step) {
    T px[kTileU], py[kTileU], pz[kTileU];
#pragma unroll
    for (int u = 0; u < kTileU; ++u) {
      const int64_t r = base + u * step + tid;
      if (r < v.count) {
        const int64_t j = v.slot(r);
        px[u] = src[j];
        py[u] = src[stride + j];
        pz[u] = src[2 * stride + j];
      }
    }
#pragma unroll
    for (int u = 0; u < kTileU; ++u) {
      const bool in = base + u * step + tid < v.count;
      uint32_t tile = 0;
      if (in) {
        int cx, cy, cz;
        uint32_t cit;
        cell_of(g, L, px[u], py[u], pz[u], cx, cy, cz, tile, cit);
      }
      const bool home = in && !ovf && (int)tile == b;
      mine += home ? 1u : 0u;
      (void)wave_rank_g(count, tile, in && !home);
    }
  }
  if (ovf) return;
  if (mine) atomicAdd(&s_own, mine);
  __syncthreads();
  if (tid == 0) {
    gmem(own_a)[b] = s_own;
    if (s_own)
      __hip_atomic_fetch_add(count + b, s_own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Tile sort, step 2 (one workgroup of kScanThreads): the exclusive scan P
// of the counts (P[ntiles] = n) and the destination layout,
// seg[t] = P[t] + t * slack, cnt[t] = count[t], no unused runs, an empty
// overflow region.
template <typename T>
__device__ __forceinline__ void tile_bounds_body(const TileSortJob<T> &J) {
  constexpr int NW = kScanThreads / 64;
  __shared__ unsigned int s_wsum[NW];
  __shared__ unsigned int s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nt = J.L.ntiles;
  if (tid == 0) s_carry = 0u;
  __syncthreads();
  for (int base = 0; base <= nt; base += kScanThreads) {
    const int t = base + tid;
    const unsigned int c = t < nt ? gmem(J.count)[t] : 0u;
    unsigned int v = c;  // inclusive scan within the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned int u = __shfl_up(v, d);
      if (lane >= d) v += u;
    }
    if (lane == 63) s_wsum[wid] = v;
    __syncthreads();
    if (wid == 0) {
      unsigned int w = lane < NW ? s_wsum[lane] : 0u;
#pragma unroll
      for (int d = 1; d < NW; d <<= 1) {
        const unsigned int u = __shfl_up(w, d);
        if (lane >= d) w += u;
      }
      if (lane < NW) s_wsum[lane] = w;  // inclusive over the waves
    }
    __syncthreads();
    const unsigned int carry = s_carry;
    const unsigned int excl = carry + (wid ? s_wsum[wid - 1] : 0u) + v - c;
    if (t <= nt) {
      gmem(J.dst_seg)[t] = (int64_t)excl + (int64_t)t * J.L.slack;
      gmem(J.dst_cnt)[t] = c;
      gmem(J.dst_cnt)[(nt + 1) + t] = 0u;
      gmem(J.dst_cnt)[2 * (nt + 1) + t] = 0u;
    }
    __syncthreads();
    if (tid == 0) s_carry = carry + s_wsum[NW - 1];
    __syncthreads();
  }
}

// Tile sort, step 3 (block b of nb of the job).  Block t < ntiles writes
// tile t's own particles to the start of its destination segment in
// round-robin cell order -- rank 0 of every cell in cell order, then rank
// 1, and so on: bin_sort's order -- and every other particle of its source
// segment after the own particles of that particle's tile (slot
// dst_seg[tile] + own[tile] + k, k from the fill counters).  Blocks ntiles
// and up do the latter for the overflow region.  For the own particles:
//  1. each particle's cell within the tile (or "foreign") goes to LDS, and
//     LDS counters count each cell;
//  2. tables give the slot of the k-th particle of cell c as A[k] + T[c][k]
//     (A[k]: particles of rank < k over all cells; T[c][k]: cells before c
//     holding more than k particles); ranks beyond the tables' KT columns
//     take the slot by a loop over the cells; each particle's slot goes to
//     LDS;
//  3. per component and output chunk of the LDS stage, the chunk's
//     particles are read in source order (coalesced) into the stage at their
//     slots, and the stage is written out in slot order (coalesced).
// A segment of more than kTileStaged particles skips the stage and writes
// each particle at its slot directly.
// BATCH: the launch covers every job of a map; the block's job, taken from
// the table, replaces the pointer and layout arguments.
template <typename T, int NC, int KT, bool BATCH>
__global__ __launch_bounds__(kTileThreads) void k_tile_sort(
    KGrid<T> g, BinLayout L, const T *src_a, int64_t stride, const uint32_t *src_id_a,
    const int64_t *src_seg_a, const unsigned int *src_cnt_a, unsigned int ovf_cap, T *dst_a,
    uint32_t *dst_id_a, const int64_t *dst_seg_a, const unsigned int *own_a,
    unsigned int *fill_a, const TileSortJob<T> *__restrict__ jobs, TileSortMap map) {
  static_assert(KT * NC <= kTileStageBytes, "slot table exceeds the stage");
  static_assert(NC <= 256, "counts of the cells before a cell must fit a byte");
  constexpr uint16_t kForeign = 0xFFFF;  // also "no slot": foreign particles take none
  constexpr uint16_t kNoSlot = 0xFFFF;
  constexpr int kChunk = kTileStageBytes / (int)sizeof(T);
  __shared__ unsigned int s_count[NC], s_rank[NC], s_kmax, s_own;
  __shared__ unsigned int s_A[KT];
  // each staged particle's cell (step 1), then its slot (step 2)
  __shared__ uint16_t s_slot[kTileStaged];
  __shared__ __align__(16) unsigned char s_raw[kTileStageBytes];
  unsigned char(*s_T)[KT] = reinterpret_cast<unsigned char(*)[KT]>(s_raw);
  int t = blockIdx.x;
  int nb = gridDim.x;
  if constexpr (BATCH) {
    const int j = sort_job(map, blockIdx.x);
    const TileSortJob<T> &J = jobs[j];
    g = J.g; L = J.L; src_a = J.src; stride = J.stride; src_id_a = J.src_id;
    src_seg_a = J.src_seg; src_cnt_a = J.src_cnt; ovf_cap = J.ovf_cap; dst_a = J.dst;
    dst_id_a = J.dst_id; dst_seg_a = J.dst_seg; own_a = J.own; fill_a = J.fill;
    t -= map.start[j];
    nb = map.start[j + 1] - map.start[j];
  }
  const SPIC_GLOBAL T *__restrict__ src = gmem(src_a);
  const SPIC_GLOBAL uint32_t *__restrict__ src_id = gmem(src_id_a);
  const SPIC_GLOBAL int64_t *__restrict__ src_seg = gmem(src_seg_a);
  const SPIC_GLOBAL unsigned int *__restrict__ src_cnt = gmem(src_cnt_a);
  SPIC_GLOBAL T *__restrict__ dst = gmem(dst_a);
  SPIC_GLOBAL uint32_t *__restrict__ dst_id = gmem(dst_id_a);
  const SPIC_GLOBAL int64_t *__restrict__ dst_seg = gmem(dst_seg_a);
  const SPIC_GLOBAL unsigned int *__restrict__ own = gmem(own_a);
  SPIC_GLOBAL unsigned int *__restrict__ fill = gmem(fill_a);
  const int tid = threadIdx.x;

  if (t >= L.ntiles) {
    const RangeView v = range_view(src_seg, src_cnt, L.ntiles, ovf_cap, L.ntiles);
    const int64_t step = (int64_t)(nb - L.ntiles) * kTileThreads;
    for (int64_t base = (int64_t)(t - L.ntiles) * kTileThreads; base < v.count;
         base += step) {
      const int64_t r = base + tid;
      const bool in = r < v.count;
      int64_t j = 0;
      uint32_t tile = 0;
      if (in) {
        j = v.slot(r);
        int cx, cy, cz;
        uint32_t cit;
        cell_of(g, L, src[j], src[stride + j], src[2 * stride + j], cx, cy, cz, tile,
                   cit);
      }
      const uint32_t k = wave_rank_g(fill, tile, in);
      if (in) copy_particle(src, src_id, stride, j, dst, dst_id, dst_seg[tile] + own[tile] + k);
    }
    return;
  }

  const RangeView v = range_view(src_seg, src_cnt, L.ntiles, ovf_cap, t);
  if (v.count == 0) return;
  const bool staged = v.count <= (int64_t)kTileStaged;
  for (int c = tid; c < NC; c += kTileThreads) {
    s_count[c] = 0u;
    s_rank[c] = 0u;
  }
  if (tid == 0) s_kmax = 0u;
  __syncthreads();

  // 1. cells and counts
  for (int64_t r0 = tid; r0 < v.count; r0 += kTileU * kTileThreads) {
    T px[kTileU], py[kTileU], pz[kTileU];
#pragma unroll
    for (int u = 0; u < kTileU; ++u) {
      const int64_t r = r0 + (int64_t)u * kTileThreads;
      if (r < v.count) {
        const int64_t j = v.slot(r);
        px[u] = src[j];
        py[u] = src[stride + j];
        pz[u] = src[2 * stride + j];
      }
    }
#pragma unroll
    for (int u = 0; u < kTileU; ++u) {
      const int64_t r = r0 + (int64_t)u * kTileThreads;
      if (r < v.count) {
        int cx, cy, cz;
        uint32_t tile, cit;
        cell_of(g, L, px[u], py[u], pz[u], cx, cy, cz, tile, cit);
        const bool home = (int)tile == t;
        if (home) atomicAdd(&s_count[cit], 1u);
        if (staged) s_slot[r] = home ? (uint16_t)cit : kForeign;
      }
    }
  }
  __syncthreads();
  for (int c = tid; c < NC; c += kTileThreads) atomicMax(&s_kmax, s_count[c]);
  __syncthreads();

  // 2. slot tables, then each particle's slot
  const int kmax = (int)(s_kmax < (unsigned)KT ? s_kmax : (unsigned)KT);
  for (int k = tid; k < kmax; k += kTileThreads) {
    unsigned int a = 0, before = 0;
#pragma unroll 8
    for (int c = 0; c < NC; ++c) {
      const unsigned int n = s_count[c];
      a += n < (unsigned)k ? n : (unsigned)k;
      s_T[c][k] = (unsigned char)before;
      before += n > (unsigned)k ? 1u : 0u;
    }
    s_A[k] = a;
  }
  if (tid == 0) {
    unsigned int n_own = 0;
    for (int c = 0; c < NC; ++c) n_own += s_count[c];
    s_own = n_own;
  }
  __syncthreads();
  const int64_t out0 = dst_seg[t];
  for (int64_t base = 0; base < v.count; base += kTileThreads) {
    const int64_t r = base + tid;
    const bool in = r < v.count;
    unsigned int c = kForeign;
    uint32_t tile = 0;
    if (in) {
      if (staged) c = s_slot[r];
      if (!staged || c == kForeign) {
        const int64_t j = v.slot(r);
        int cx, cy, cz;
        uint32_t cit;
        cell_of(g, L, src[j], src[stride + j], src[2 * stride + j], cx, cy, cz, tile,
                   cit);
        c = (int)tile == t ? cit : kForeign;
      }
    }
    const bool foreign = in && c == kForeign;
    const uint32_t kf = wave_rank_g(fill, tile, foreign);
    if (!in) continue;
    if (foreign) {
      copy_particle(src, src_id, stride, v.slot(r), dst, dst_id,
                    dst_seg[tile] + own[tile] + kf);
      if (staged) s_slot[r] = kNoSlot;
      continue;
    }
    const unsigned int k = atomicAdd(&s_rank[c], 1u);
    unsigned int pos;
    if (k < (unsigned)KT) {
      pos = s_A[k] + s_T[c][k];
    } else {
      pos = 0;
#pragma unroll 4
      for (int cc = 0; cc < NC; ++cc) {
        const unsigned int n = s_count[cc];
        pos += n < k ? n : k;
        if (cc < (int)c && n > k) ++pos;
      }
    }
    if (staged) s_slot[r] = (uint16_t)pos;
    else copy_particle(src, src_id, stride, v.slot(r), dst, dst_id, out0 + pos);
  }
  if (!staged) return;
  __syncthreads();  // slots complete; the tables' LDS becomes the stage

  // 3. per component and output chunk: gather into the stage, write out
  const unsigned int n_own = s_own;
  T *stage = reinterpret_cast<T *>(s_raw);
  uint32_t *stage_id = reinterpret_cast<uint32_t *>(s_raw);
  const int ncomp = dst_id ? 8 : 7;  // the ids are the eighth, when kept
  for (int a = 0; a < ncomp; ++a) {
    for (unsigned int c0 = 0; c0 < n_own; c0 += (unsigned)kChunk) {
      for (int64_t r0 = tid; r0 < v.count; r0 += kTileU * kTileThreads) {
        unsigned int p[kTileU];
        T val[kTileU];
        uint32_t idv[kTileU];
#pragma unroll
        for (int u = 0; u < kTileU; ++u) {
          const int64_t r = r0 + (int64_t)u * kTileThreads;
          p[u] = kChunk;
          if (r < v.count && s_slot[r] != kNoSlot) {
            p[u] = (unsigned int)s_slot[r] - c0;  // wraps above kChunk below c0
            if (p[u] < (unsigned)kChunk) {
              const int64_t j = v.slot(r);
              if (a < 7) val[u] = src[a * stride + j];
              else idv[u] = src_id[j];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kTileU; ++u) {
          if (p[u] < (unsigned)kChunk) {
            if (a < 7) stage[p[u]] = val[u];
            else stage_id[p[u]] = idv[u];
          }
        }
      }
      __syncthreads();
      const unsigned int m = n_own - c0 < (unsigned)kChunk ? n_own - c0 : (unsigned)kChunk;
      for (unsigned int p = tid; p < m; p += kTileThreads) {
        if (a < 7) dst[a * stride + out0 + c0 + p] = stage[p];
        else dst_id[out0 + c0 + p] = stage_id[p];
      }
      __syncthreads();
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_tile_bounds(TileSortJob<T> J) {
  tile_bounds_body<T>(J);
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_tile_bounds_batch(
    const TileSortJob<T> *__restrict__ jobs) {
  const TileSortJob<T> &J = jobs[blockIdx.x];
  tile_bounds_body<T>(J);
}

template <typename T>
hipError_t tile_sort(const TileSortJob<T> &J, hipStream_t s) {
  const int nc = J.L.tx * J.L.ty * J.L.tz;
  if (nc != 64 && nc != 128) return hipErrorInvalidValue;
  const int64_t nzero = 3 * ((int64_t)J.L.ntiles + 1);
  hipLaunchKernelGGL(k_zero_u32, dim3(nblocks(nzero)), dim3(256), 0, s, J.count, nzero);
  hipError_t e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const dim3 grid(J.L.ntiles + kTileOvfBlocks);
  hipLaunchKernelGGL((k_tile_count<T, false>), grid, dim3(kTileThreads), 0, s, J.g, J.L, J.src,
                     J.stride, J.src_seg, J.src_cnt, J.ovf_cap, J.count, J.own,
                     (const TileSortJob<T> *)nullptr, TileSortMap{});
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_tile_bounds<T>, dim3(1), dim3(kScanThreads), 0, s, J);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (nc == 64)
    hipLaunchKernelGGL((k_tile_sort<T, 64, 512, false>), grid, dim3(kTileThreads), 0, s, J.g, J.L,
                       J.src, J.stride, J.src_id, J.src_seg, J.src_cnt, J.ovf_cap, J.dst,
                       J.dst_id, (const int64_t *)J.dst_seg, (const unsigned int *)J.own,
                       J.fill, (const TileSortJob<T> *)nullptr, TileSortMap{});
  else
    hipLaunchKernelGGL((k_tile_sort<T, 128, 256, false>), grid, dim3(kTileThreads), 0, s, J.g, J.L,
                       J.src, J.stride, J.src_id, J.src_seg, J.src_cnt, J.ovf_cap, J.dst,
                       J.dst_id, (const int64_t *)J.dst_seg, (const unsigned int *)J.own,
                       J.fill, (const TileSortJob<T> *)nullptr, TileSortMap{});
  return hipGetLastError();
}

template <typename T>
hipError_t tile_sort_batch(const TileSortJob<T> *jobs, const TileSortMap &map, int tx,
                           int ty, int tz, unsigned int *ws, size_t ws_words,
                           hipStream_t s) {
  const int nc = tx * ty * tz;
  if (nc != 64 && nc != 128) return hipErrorInvalidValue;
  if (map.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3(nblocks((long)ws_words)), dim3(256), 0, s, ws,
                     (int64_t)ws_words);
  hipError_t e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const dim3 grid(map.start[map.n]);
  const KGrid<T> g0{};
  const BinLayout L0{};
  hipLaunchKernelGGL((k_tile_count<T, true>), grid, dim3(kTileThreads), 0, s, g0, L0,
                     (const T *)nullptr, (int64_t)0, (const int64_t *)nullptr,
                     (const unsigned int *)nullptr, 0u, (unsigned int *)nullptr,
                     (unsigned int *)nullptr, jobs, map);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_tile_bounds_batch<T>, dim3(map.n), dim3(kScanThreads), 0, s, jobs);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // the batch instances take their arguments from the job table
  if (nc == 64)
    hipLaunchKernelGGL((k_tile_sort<T, 64, 512, true>), grid, dim3(kTileThreads), 0, s, g0, L0,
                       (const T *)nullptr, (int64_t)0, (const uint32_t *)nullptr,
                       (const int64_t *)nullptr, (const unsigned int *)nullptr, 0u,
                       (T *)nullptr, (uint32_t *)nullptr, (const int64_t *)nullptr,
                       (const unsigned int *)nullptr, (unsigned int *)nullptr, jobs, map);
  else
    hipLaunchKernelGGL((k_tile_sort<T, 128, 256, true>), grid, dim3(kTileThreads), 0, s, g0,
                       L0, (const T *)nullptr, (int64_t)0, (const uint32_t *)nullptr,
                       (const int64_t *)nullptr, (const unsigned int *)nullptr, 0u,
                       (T *)nullptr, (uint32_t *)nullptr, (const int64_t *)nullptr,
                       (const unsigned int *)nullptr, (unsigned int *)nullptr, jobs, map);
  return hipGetLastError();
}

template hipError_t tile_sort<double>(const TileSortJob<double> &, hipStream_t);
template hipError_t tile_sort<float>(const TileSortJob<float> &, hipStream_t);
template hipError_t tile_sort_batch<double>(const TileSortJob<double> *, const TileSortMap &,
                                            int, int, int, unsigned int *, size_t,
                                            hipStream_t);
template hipError_t tile_sort_batch<float>(const TileSortJob<float> *, const TileSortMap &,
                                           int, int, int, unsigned int *, size_t,
                                           hipStream_t);

hipError_t zero_counts(const CountArrays &c, hipStream_t s) {
  if (c.ns <= 0 || c.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_counts, dim3(nblocks(c.n > c.nn ? c.n : c.nn), c.ns), dim3(256),
                     0, s, c);
  return hipGetLastError();
}

template <typename T>
hipError_t bin_sort(const KGrid<T> &g, const BinLayout &L, int64_t n,
                    const void *const src[7], int src_f64, const uint32_t *src_id,
                    const int64_t *src_seg, const unsigned int *src_cnt,
                    unsigned int src_ovf_cap, T *const dst[7], uint32_t *dst_id,
                    int64_t *dst_seg, unsigned int *dst_cnt, void *workspace,
                    size_t workspace_bytes, size_t *needed, hipStream_t s) {
  if (n < 0 || n > 0xFFFFFFFFll) return hipErrorInvalidValue;
  const bool binned = src_seg != nullptr;
  // the caller's FP64 arrays into an FP32 engine (a binned source is always
  // the engine's own precision)
  const bool widened = !binned && src_f64 && sizeof(T) == 4;
  if (binned && src_f64 != (sizeof(T) == 8 ? 1 : 0)) return hipErrorInvalidValue;
  const size_t nt1 = (size_t)L.ntiles + 1;
  const size_t a_n = align_up(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  const size_t a_t = align_up(sizeof(uint32_t) * nt1);
  const int64_t ncell = (int64_t)L.ncx * L.ncy * L.ncz;
  const size_t a_c = align_up(sizeof(uint32_t) * (size_t)ncell);
  const int shift = L.rank_bits + L.cell_bits;
  const int key_bits = L.tile_bits + shift;
  size_t sort_bytes = 0, scan_bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(
      nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
      (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, key_bits, s);
  if (e != hipSuccess) return e;
  e = rocprim::exclusive_scan(nullptr, scan_bytes, (const unsigned int *)nullptr,
                              (unsigned int *)nullptr, 0u, nt1,
                              rocprim::plus<unsigned int>(), s);
  if (e != hipSuccess) return e;
  const size_t total =
      4 * a_n + 2 * a_t + a_c + align_up(sort_bytes) + align_up(scan_bytes);
  if (needed) *needed = total;
  if (!workspace) return hipSuccess;
  if (workspace_bytes < total) return hipErrorInvalidValue;
  char *w = (char *)workspace;
  uint32_t *keys_in = (uint32_t *)w;
  uint32_t *keys_out = (uint32_t *)(w + a_n);
  uint32_t *idx_in = (uint32_t *)(w + 2 * a_n);
  uint32_t *perm = (uint32_t *)(w + 3 * a_n);
  unsigned int *clamped = (unsigned int *)(w + 4 * a_n);
  unsigned int *P = (unsigned int *)(w + 4 * a_n + a_t);
  unsigned int *cell_count = (unsigned int *)(w + 4 * a_n + 2 * a_t);
  void *sort_tmp = (void *)(w + 4 * a_n + 2 * a_t + a_c);
  void *scan_tmp = (void *)(w + 4 * a_n + 2 * a_t + a_c + align_up(sort_bytes));

  if (binned && n > 0) {
    hipLaunchKernelGGL(k_range_counts, dim3(nblocks((long)nt1)), dim3(256), 0, s,
                       src_seg, src_cnt, src_ovf_cap, L.ntiles, clamped);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = rocprim::exclusive_scan(scan_tmp, scan_bytes, (const unsigned int *)clamped,
                                P, 0u, nt1, rocprim::plus<unsigned int>(), s);
    if (e != hipSuccess) return e;
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_zero_u32, dim3(nblocks(ncell)), dim3(256), 0, s, cell_count,
                       ncell);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (binned)
      hipLaunchKernelGGL(k_bin_keys_binned<T>, dim3((unsigned)nt1), dim3(256), 0, s, g, L,
                         (const T *)src[0], (const T *)src[1], (const T *)src[2],
                         (const unsigned int *)P, src_seg, src_cnt, src_ovf_cap, cell_count,
                         keys_in, idx_in);
    else if (widened)
      hipLaunchKernelGGL((k_bin_keys<double, T>), dim3(nblocks(n)), dim3(256), 0, s, g, L, n,
                         (const double *)src[0], (const double *)src[1],
                         (const double *)src[2], cell_count, keys_in, idx_in);
    else
      hipLaunchKernelGGL((k_bin_keys<T, T>), dim3(nblocks(n)), dim3(256), 0, s, g, L, n,
                         (const T *)src[0], (const T *)src[1], (const T *)src[2],
                         cell_count, keys_in, idx_in);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys_in, keys_out, idx_in,
                                  perm, (size_t)n, 0, key_bits, s);
    if (e != hipSuccess) return e;
    Arrays7<T> ad;
    for (int a = 0; a < 7; ++a) ad.p[a] = dst[a];
    const int conv = (!binned && L.fixed) ? 1 : 0;
    const uint32_t *sid = binned ? src_id : (const uint32_t *)nullptr;
    if (widened) {
      Arrays7<double> as;
      for (int a = 0; a < 7; ++a) as.p[a] = (double *)src[a];
      hipLaunchKernelGGL((k_bin_gather<double, T>), dim3(nblocks(n)), dim3(256), 0, s, as,
                         sid, (const uint32_t *)perm, (const uint32_t *)keys_out, shift, n,
                         L.slack, ad, dst_id, conv, L.fx);
    } else {
      Arrays7<T> as;
      for (int a = 0; a < 7; ++a) as.p[a] = (T *)src[a];
      hipLaunchKernelGGL((k_bin_gather<T, T>), dim3(nblocks(n)), dim3(256), 0, s, as, sid,
                         (const uint32_t *)perm, (const uint32_t *)keys_out, shift, n,
                         L.slack, ad, dst_id, conv, L.fx);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_bin_bounds, dim3(nblocks((long)nt1)), dim3(256), 0, s,
                     (const uint32_t *)keys_out, shift, n, L.ntiles, L.slack, dst_seg,
                     dst_cnt);
  return hipGetLastError();
}

template hipError_t bin_sort<double>(const KGrid<double> &, const BinLayout &, int64_t,
                                     const void *const[7], int, const uint32_t *,
                                     const int64_t *, const unsigned int *,
                                     unsigned int, double *const[7], uint32_t *,
                                     int64_t *, unsigned int *, void *, size_t,
                                     size_t *, hipStream_t);
template hipError_t bin_sort<float>(const KGrid<float> &, const BinLayout &, int64_t,
                                    const void *const[7], int, const uint32_t *,
                                    const int64_t *, const unsigned int *,
                                    unsigned int, float *const[7], uint32_t *,
                                    int64_t *, unsigned int *, void *, size_t,
                                    size_t *, hipStream_t);

template <typename T>
hipError_t bin_store(const BinLayout &L, T *const src[7], const uint32_t *id,
                     const uint32_t *base, const int64_t *seg, const unsigned int *cnt,
                     unsigned int ovf_cap, void *const dst[7], int dst_f64, hipStream_t s) {
  Arrays7<T> as;
  for (int a = 0; a < 7; ++a) as.p[a] = src[a];
  if (dst_f64) {
    Arrays7<double> ad;
    for (int a = 0; a < 7; ++a) ad.p[a] = (double *)dst[a];
    hipLaunchKernelGGL((k_bin_store<T, double>), dim3(L.ntiles + 1), dim3(256), 0, s, as, id,
                       base, seg, cnt, ovf_cap, L.ntiles, ad, L.fixed, L.fx);
  } else {
    if (sizeof(T) == 8) return hipErrorInvalidValue;  // no narrowing store
    Arrays7<float> ad;
    for (int a = 0; a < 7; ++a) ad.p[a] = (float *)dst[a];
    hipLaunchKernelGGL((k_bin_store<T, float>), dim3(L.ntiles + 1), dim3(256), 0, s, as, id,
                       base, seg, cnt, ovf_cap, L.ntiles, ad, L.fixed, L.fx);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t bin_store_be32(const BinLayout &L, T *const src[7], const uint32_t *id,
                          const uint32_t *base, const int64_t *seg, const unsigned int *cnt,
                          unsigned int ovf_cap, uint32_t *const dst[6], hipStream_t s) {
  Arrays7<T> as;
  Words6 ad;
  for (int a = 0; a < 7; ++a) as.p[a] = src[a];
  for (int a = 0; a < 6; ++a) ad.p[a] = dst[a];
  hipLaunchKernelGGL(k_bin_store_be32<T>, dim3(L.ntiles + 1), dim3(256), 0, s, as, id, base,
                     seg, cnt, ovf_cap, L.ntiles, ad, L.fixed, L.fx);
  return hipGetLastError();
}

template hipError_t bin_store_be32<double>(const BinLayout &, double *const[7],
                                           const uint32_t *, const uint32_t *, const int64_t *,
                                           const unsigned int *, unsigned int,
                                           uint32_t *const[6], hipStream_t);
template hipError_t bin_store_be32<float>(const BinLayout &, float *const[7],
                                          const uint32_t *, const uint32_t *, const int64_t *,
                                          const unsigned int *, unsigned int,
                                          uint32_t *const[6], hipStream_t);

template hipError_t bin_store<double>(const BinLayout &, double *const[7],
                                      const uint32_t *, const uint32_t *, const int64_t *,
                                      const unsigned int *, unsigned int,
                                      void *const[7], int, hipStream_t);
template hipError_t bin_store<float>(const BinLayout &, float *const[7],
                                     const uint32_t *, const uint32_t *, const int64_t *,
                                     const unsigned int *, unsigned int,
                                     void *const[7], int, hipStream_t);

// Dense position of each range's first particle (tiles in order, then the
// overflow region), for stores of an engine that keeps no ids: range counts
// then an exclusive scan.  Workspace: counts | bases | rocPRIM storage; with
// workspace == nullptr only *needed is set.  *bases points into the
// workspace.
hipError_t range_bases(const int64_t *seg, const unsigned int *cnt, unsigned int ovf_cap,
                       int ntiles, void *workspace, size_t workspace_bytes, size_t *needed,
                       const uint32_t **bases, hipStream_t s) {
  const size_t n = (size_t)ntiles + 1;
  const size_t a = align_up(n * sizeof(unsigned int));
  size_t tmp = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, tmp, (const unsigned int *)nullptr,
                                         (unsigned int *)nullptr, 0u, n,
                                         rocprim::plus<unsigned int>(), s);
  if (e != hipSuccess) return e;
  if (needed) *needed = 2 * a + align_up(tmp);
  if (!workspace) return hipSuccess;
  if (workspace_bytes < 2 * a + align_up(tmp)) return hipErrorInvalidValue;
  unsigned int *counts = (unsigned int *)workspace;
  unsigned int *out = (unsigned int *)((char *)workspace + a);
  hipLaunchKernelGGL(k_range_counts, dim3(nblocks((long)n)), dim3(256), 0, s, seg,
                     cnt, ovf_cap, ntiles, counts);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  e = rocprim::exclusive_scan((char *)workspace + 2 * a, tmp, counts, out, 0u, n,
                              rocprim::plus<unsigned int>(), s);
  if (e != hipSuccess) return e;
  *bases = out;
  return hipSuccess;
}

}  // namespace spic
