// spic_sort.hip -- particle_sort on the device, in exactly the reference's
// output order (src/Particles.cpp:155-233): a stable sort of all particles
// (over npmax) by bin, bins of sort_cps guard-inclusive cells per side,
// z-fastest bin numbering, out-of-range particles in a trailing "lost" bin.
//
// Bin keys are formed per particle; a stable LSD radix sort (rocPRIM) of
// (key, index) pairs over only the key bits gives the permutation; the
// seven arrays are then gathered through it one at a time into one scratch
// array and copied back.
#include <cstring>
#include <mutex>

#include <rocprim/device/device_radix_sort.hpp>

#include "spic_internal.hpp"

namespace spic {

namespace {

template <typename T>
__global__ void k_bin_keys(const T *__restrict__ x, const T *__restrict__ y,
                           const T *__restrict__ z, int n, T dbx, T dby,
                           T dbz, long nbx, long nby, long nbz,
                           unsigned int *__restrict__ key,
                           unsigned int *__restrict__ idx) {
  const long lost = nbx * nby * nbz;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    // floor of the FPpart quotient, then to long (Particles.cpp:180-182);
    // the quotient is clamped so the conversion is defined for any input.
    double qx = (double)(x[i] / dbx), qy = (double)(y[i] / dby),
           qz = (double)(z[i] / dbz);
    qx = fmin(fmax(floor(qx), -1.0), (double)nbx);
    qy = fmin(fmax(floor(qy), -1.0), (double)nby);
    qz = fmin(fmax(floor(qz), -1.0), (double)nbz);
    long px = (long)qx, py = (long)qy, pz = (long)qz;
    long b;
    if (px < 0 || px >= nbx || py < 0 || py >= nby || pz < 0 || pz >= nbz)
      b = lost;
    else
      b = pz + py * nbz + px * nbz * nby;
    key[i] = (unsigned int)b;
    idx[i] = (unsigned int)i;
  }
}

template <typename T>
__global__ void k_gather(const T *__restrict__ src, T *__restrict__ dst,
                         const unsigned int *__restrict__ perm, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

__global__ void k_fill_f64(double *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ void k_fill_u32(unsigned int *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

inline unsigned blocks_for(size_t n, unsigned tpb) {
  size_t b = (n + tpb - 1) / tpb;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (unsigned)b;
}

inline size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

hipError_t zero_f64(double *p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(n, 256)), dim3(256), 0, s, p,
                     n);
  return hipGetLastError();
}

hipError_t zero_u32(unsigned int *p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p,
                     n);
  return hipGetLastError();
}

// Workspace layout: keys_in | keys_out | idx_in | perm | scratch(T, n) |
// rocPRIM temporary storage.  With workspace == nullptr only *needed is set.
template <typename T>
hipError_t sort_reference_order(const spic_grid &g, int64_t sort_cps,
                                spic_species &sp, void *workspace,
                                size_t workspace_bytes, size_t *needed,
                                hipStream_t s) {
  const int n = (int)sp.npmax;
  const long nbx = 1 + ((g.nxc - 1) / sort_cps);
  const long nby = 1 + ((g.nyc - 1) / sort_cps);
  const long nbz = 1 + ((g.nzc - 1) / sort_cps);
  const long bins_tot = nbx * nby * nbz;
  // dist_bin = grd->L / (FPpart)n_bin, a double quotient narrowed to FPpart
  const T dbx = (T)(g.Lx / (double)(T)nbx);
  const T dby = (T)(g.Ly / (double)(T)nby);
  const T dbz = (T)(g.Lz / (double)(T)nbz);
  unsigned end_bit = 1;
  while (end_bit < 32 && ((unsigned long)bins_tot >> end_bit) != 0) ++end_bit;

  const size_t a_u32 = align_up(sizeof(unsigned int) * (size_t)(n > 0 ? n : 1));
  const size_t a_t = align_up(sizeof(T) * (size_t)(n > 0 ? n : 1));
  size_t tmp_bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(
      nullptr, tmp_bytes, (unsigned int *)nullptr, (unsigned int *)nullptr,
      (unsigned int *)nullptr, (unsigned int *)nullptr, (size_t)n, 0, end_bit,
      s);
  if (e != hipSuccess) return e;
  const size_t total = 4 * a_u32 + a_t + align_up(tmp_bytes);
  if (needed) *needed = total;
  if (workspace == nullptr) return hipSuccess;
  if (workspace_bytes < total) return hipErrorInvalidValue;
  if (n <= 0) return hipSuccess;

  char *w = (char *)workspace;
  unsigned int *keys_in = (unsigned int *)w;
  unsigned int *keys_out = (unsigned int *)(w + a_u32);
  unsigned int *idx_in = (unsigned int *)(w + 2 * a_u32);
  unsigned int *perm = (unsigned int *)(w + 3 * a_u32);
  T *scratch = (T *)(w + 4 * a_u32);
  void *tmp = (void *)(w + 4 * a_u32 + a_t);

  hipLaunchKernelGGL(k_bin_keys<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     (const T *)sp.x, (const T *)sp.y, (const T *)sp.z, n, dbx,
                     dby, dbz, nbx, nby, nbz, keys_in, idx_in);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, idx_in,
                                perm, (size_t)n, 0, end_bit, s);
  if (e != hipSuccess) return e;
  T *arrs[7] = {(T *)sp.x, (T *)sp.y, (T *)sp.z, (T *)sp.u,
                (T *)sp.v, (T *)sp.w, (T *)sp.q};
  for (int a = 0; a < 7; ++a) {
    hipLaunchKernelGGL(k_gather<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                       arrs[a], scratch, perm, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = hipMemcpyAsync(arrs[a], scratch, sizeof(T) * (size_t)n,
                       hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template hipError_t sort_reference_order<double>(const spic_grid &, int64_t,
                                                 spic_species &, void *,
                                                 size_t, size_t *,
                                                 hipStream_t);
template hipError_t sort_reference_order<float>(const spic_grid &, int64_t,
                                                spic_species &, void *, size_t,
                                                size_t *, hipStream_t);

}  // namespace spic

// ---- particle_sort without allocation ------------------------------------
//
// The sort's scratch (bin keys in and out, indices, the permutation, the
// gather staging array and rocPRIM's temporary storage) in a fixed device
// array, one copy per GPU, held until the process ends.  Nothing in it has
// to start at zero: every part is written before it is read within a call.
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(SPIC_SORT_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device, looked up once per device with
// that device current.  The host-side name `scratch` is not a device address
// and is never passed to a kernel or a copy.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_pointer(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return spic::hip_fail(e, "particle_sort_noalloc: current device");
  if (dev < 0 || dev >= kMaxDevices)
    return spic::fail(SPIC_EUNSUPPORTED, "particle_sort_noalloc: device %d beyond %d", dev,
                      kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return spic::hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                            "particle_sort_noalloc: scratch address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SPIC_OK;
}

}  // namespace

extern "C" int spic_particle_sort_noalloc(const spic_grid *grd, int64_t sort_cps,
                                          spic_species *sp, void *stream) {
  int rc;
  if ((rc = spic::check_grid(grd)) || (rc = spic::check_species(sp, true, false))) return rc;
  if (sort_cps < 1) return spic::fail(SPIC_EINVAL, "sort_cps must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  size_t need = 0;
  hipError_t e =
      grd->prec == SPIC_F64
          ? spic::sort_reference_order<double>(*grd, sort_cps, *sp, nullptr, 0, &need, s)
          : spic::sort_reference_order<float>(*grd, sort_cps, *sp, nullptr, 0, &need, s);
  if (e != hipSuccess) return spic::hip_fail(e, "particle_sort_noalloc: sizing");
  if (need > kNoAllocScratchBytes)
    return spic::fail(SPIC_EINVAL,
                      "particle_sort_noalloc: %lld particles need %zu bytes of scratch; the "
                      "fixed scratch holds %zu (use spic_particle_sort)",
                      (long long)sp->npmax, need, kNoAllocScratchBytes);
  void *ws = nullptr;
  if ((rc = scratch_pointer(&ws))) return rc;
  e = grd->prec == SPIC_F64
          ? spic::sort_reference_order<double>(*grd, sort_cps, *sp, ws, need, nullptr, s)
          : spic::sort_reference_order<float>(*grd, sort_cps, *sp, ws, need, nullptr, s);
  if (e != hipSuccess) return spic::hip_fail(e, "particle_sort_noalloc");
  if (!spic::async_mode()) {
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return spic::hip_fail(e, "particle_sort_noalloc: execution");
  }
  return SPIC_OK;
}
