// spic_tilesort.hip -- sort of one species into tile order for the
// tile-resident phase kernel.
//
// Key of a particle (32 bits, most significant first):
//   tile index | rank of the particle within its cell | cell within tile
// where the cell is that of the deposit (floor) index of the current
// position, clamped to the physical cells, and the rank is a per-cell
// counter taken with an atomic.  Sorting by this key puts each tile's
// particles in one contiguous segment, ordered round-robin over the tile's
// cells (all rank-0 particles of the tile's cells, then all rank-1, ...), so
// that 64 consecutive particles lie in up to 64 different cells.
// A stable LSD radix sort (rocPRIM) of (key, index) pairs gives the
// permutation; x..w, q and the particle id are gathered into the second
// buffer set; the tile segment table is found by binary search.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "spic_tiled.hpp"

namespace spic {
namespace {

inline unsigned nblocks(long n, int tpb = 256) {
  long b = (n + tpb - 1) / tpb;
  if (b < 1) b = 1;
  if (b > (1L << 20)) b = 1L << 20;
  return (unsigned)b;
}

inline size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

template <typename T>
__device__ __forceinline__ int phys_cell(T p, double start, T inv, int nc) {
  // deposit cell 2 + floor((p - start) * inv) minus the two guard offsets
  double c = floor(((double)p - start) * (double)inv);
  c = fmin(fmax(c, 0.0), (double)(nc - 1));
  return (int)c;
}

template <typename T>
__global__ void k_tile_keys(KGrid<T> g, TileLayout L, int n,
                            const T *__restrict__ x, const T *__restrict__ y,
                            const T *__restrict__ z,
                            unsigned int *__restrict__ cell_count,
                            uint32_t *__restrict__ key,
                            uint32_t *__restrict__ idx) {
  const uint32_t rank_max = (1u << L.rank_bits) - 1u;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const int cx = phys_cell<T>(x[i], g.xStart, g.invdx, L.ncx);
    const int cy = phys_cell<T>(y[i], g.yStart, g.invdy, L.ncy);
    const int cz = phys_cell<T>(z[i], g.zStart, g.invdz, L.ncz);
    const int tx = cx / L.tx, ty = cy / L.ty, tz = cz / L.tz;
    const int tile = (tx * L.nty + ty) * L.ntz + tz;
    const int cit = ((cx - tx * L.tx) * L.ty + (cy - ty * L.ty)) * L.tz +
                    (cz - tz * L.tz);
    const int cell = (cx * L.ncy + cy) * L.ncz + cz;
    uint32_t r = atomicAdd(&cell_count[cell], 1u);
    r = r < rank_max ? r : rank_max;
    key[i] = ((uint32_t)tile << (L.rank_bits + L.cell_bits)) |
             (r << L.cell_bits) | (uint32_t)cit;
    idx[i] = (uint32_t)i;
  }
}

template <typename T>
__global__ void k_gather8(const T *__restrict__ x, const T *__restrict__ y,
                          const T *__restrict__ z, const T *__restrict__ u,
                          const T *__restrict__ v, const T *__restrict__ w,
                          const T *__restrict__ q,
                          const uint32_t *__restrict__ id,
                          const uint32_t *__restrict__ perm, int n, T *ox,
                          T *oy, T *oz, T *ou, T *ov, T *ow, T *oq,
                          uint32_t *oid) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t j = perm[i];
    ox[i] = x[j]; oy[i] = y[j]; oz[i] = z[j];
    ou[i] = u[j]; ov[i] = v[j]; ow[i] = w[j];
    oq[i] = q[j];
    oid[i] = id[j];
  }
}

// tile_start[t] = first sorted position whose tile field is >= t
__global__ void k_tile_bounds(const uint32_t *__restrict__ keys, int n,
                              int shift, int ntiles,
                              int64_t *__restrict__ tile_start) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles;
       t += gridDim.x * blockDim.x) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = lo + ((hi - lo) >> 1);
      if ((int)(keys[mid] >> shift) < t) lo = mid + 1;
      else hi = mid;
    }
    tile_start[t] = lo;
  }
}

__global__ void k_zero_u32(unsigned int *p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0u;
}

template <typename T>
__global__ void k_scatter_by_id(const T *__restrict__ src,
                                const uint32_t *__restrict__ id,
                                T *__restrict__ dst, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    dst[id[i]] = src[i];
}

__global__ void k_iota(uint32_t *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

}  // namespace

template <typename T>
hipError_t tile_sort(const KGrid<T> &g, const TileLayout &L, int n,
                     T *const src[7], const uint32_t *src_id, T *const dst[7],
                     uint32_t *dst_id, int64_t *tile_start, void *workspace,
                     size_t workspace_bytes, size_t *needed, hipStream_t s) {
  const long ncell = (long)L.ncx * L.ncy * L.ncz;
  const size_t a_n = align_up(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  const size_t a_c = align_up(sizeof(uint32_t) * (size_t)ncell);
  size_t tmp_bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(
      nullptr, tmp_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
      (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, 32, s);
  if (e != hipSuccess) return e;
  const size_t total = 4 * a_n + a_c + align_up(tmp_bytes);
  if (needed) *needed = total;
  if (!workspace) return hipSuccess;
  if (workspace_bytes < total) return hipErrorInvalidValue;
  const int shift = L.rank_bits + L.cell_bits;
  if (n <= 0) {
    hipLaunchKernelGGL(k_tile_bounds, dim3(nblocks(L.ntiles + 1)), dim3(256), 0,
                       s, (const uint32_t *)nullptr, 0, shift, L.ntiles,
                       tile_start);
    return hipGetLastError();
  }
  char *w = (char *)workspace;
  uint32_t *keys_in = (uint32_t *)w;
  uint32_t *keys_out = (uint32_t *)(w + a_n);
  uint32_t *idx_in = (uint32_t *)(w + 2 * a_n);
  uint32_t *perm = (uint32_t *)(w + 3 * a_n);
  unsigned int *cell_count = (unsigned int *)(w + 4 * a_n);
  void *tmp = (void *)(w + 4 * a_n + a_c);

  hipLaunchKernelGGL(k_zero_u32, dim3(nblocks(ncell)), dim3(256), 0, s,
                     cell_count, ncell);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_tile_keys<T>, dim3(nblocks(n)), dim3(256), 0, s, g, L,
                     n, (const T *)src[0], (const T *)src[1],
                     (const T *)src[2], cell_count, keys_in, idx_in);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, idx_in,
                                perm, (size_t)n, 0, 32, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather8<T>, dim3(nblocks(n)), dim3(256), 0, s,
                     (const T *)src[0], (const T *)src[1], (const T *)src[2],
                     (const T *)src[3], (const T *)src[4], (const T *)src[5],
                     (const T *)src[6], src_id, perm, n, dst[0], dst[1], dst[2],
                     dst[3], dst[4], dst[5], dst[6], dst_id);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_tile_bounds, dim3(nblocks(L.ntiles + 1)), dim3(256), 0,
                     s, keys_out, n, shift, L.ntiles, tile_start);
  return hipGetLastError();
}

template hipError_t tile_sort<double>(const KGrid<double> &, const TileLayout &,
                                      int, double *const[7], const uint32_t *,
                                      double *const[7], uint32_t *, int64_t *,
                                      void *, size_t, size_t *, hipStream_t);
template hipError_t tile_sort<float>(const KGrid<float> &, const TileLayout &,
                                     int, float *const[7], const uint32_t *,
                                     float *const[7], uint32_t *, int64_t *,
                                     void *, size_t, size_t *, hipStream_t);

hipError_t scatter_by_id_f64(const double *src, const uint32_t *id, double *dst,
                             int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_by_id<double>, dim3(nblocks(n)), dim3(256), 0, s,
                     src, id, dst, n);
  return hipGetLastError();
}

hipError_t scatter_by_id_f32(const float *src, const uint32_t *id, float *dst,
                             int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_by_id<float>, dim3(nblocks(n)), dim3(256), 0, s,
                     src, id, dst, n);
  return hipGetLastError();
}

hipError_t iota_u32(uint32_t *p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(nblocks(n)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace spic
