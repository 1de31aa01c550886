// spic_snapshot.hip -- particle snapshot payloads on the device, and a
// device permutation of a species (to measure on input in random order).
// HDF5_Write_Particles (src/RW_IO.cpp:376-460) stores, per species, the
// datasets x y z u v w (no q) as H5T_IEEE_F32BE; this kernel produces those
// words from a species in the caller's order, values rounded to float, so
// that a snapshot is one kernel and one device-to-host copy that can run on
// a side stream while the next phase proceeds.  The binned engine's variant
// (spic_engine_snapshot_be32) scatters to load order in the same pass.
#include <hip/hip_runtime.h>

#include "spic.h"
#include "spic_internal.hpp"

namespace spic {
namespace {

struct Src6 {
  const void *p[6];
};
struct Dst6 {
  uint32_t *p[6];
};

template <typename T>
__global__ void k_be32(Src6 src, Dst6 dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int a = 0; a < 6; ++a)
      dst.p[a][i] = __builtin_bswap32(__float_as_uint((float)((const T *)src.p[a])[i]));
  }
}

// dst[(a * i + b) mod n] = src[i] for the seven arrays (a coprime to n:
// a bijection; a, b < n < 2^31, so a * i + b fits in 64 bits).
template <typename T>
__global__ void k_permute(Src6 src, const void *srcq, void *const *dst, int64_t n, uint64_t a,
                          uint64_t b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = (int64_t)((a * (uint64_t)i + b) % (uint64_t)n);
#pragma unroll
    for (int c = 0; c < 6; ++c) ((T *)dst[c])[j] = ((const T *)src.p[c])[i];
    ((T *)dst[6])[j] = ((const T *)srcq)[i];
  }
}

}  // namespace
}  // namespace spic

extern "C" int spic_snapshot_be32(const spic_grid *grd, const spic_species *sp,
                                  void *const dst[6], void *stream) {
  int rc = spic::check_grid(grd);
  if (rc) return rc;
  if ((rc = spic::check_species(sp, false, false))) return rc;
  if (!dst) return spic::fail(SPIC_EINVAL, "snapshot arrays are NULL");
  spic::Src6 s6 = {{sp->x, sp->y, sp->z, sp->u, sp->v, sp->w}};
  spic::Dst6 d6;
  for (int a = 0; a < 6; ++a) {
    if (!dst[a]) return spic::fail(SPIC_EINVAL, "snapshot array %d is NULL", a);
    d6.p[a] = (uint32_t *)dst[a];
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = sp->nop;
  if (n > 0) {
    long b = (long)((n + 255) / 256);
    if (b > 65536) b = 65536;
    if (grd->prec == SPIC_F64)
      hipLaunchKernelGGL(spic::k_be32<double>, dim3((unsigned)b), dim3(256), 0, s, s6, d6, n);
    else
      hipLaunchKernelGGL(spic::k_be32<float>, dim3((unsigned)b), dim3(256), 0, s, s6, d6, n);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !spic::async_mode()) e = hipStreamSynchronize(s);
  return e == hipSuccess ? SPIC_OK : spic::hip_fail(e, "snapshot");
}

namespace {
uint64_t gcd64(uint64_t x, uint64_t y) {
  while (y) {
    const uint64_t t = x % y;
    x = y;
    y = t;
  }
  return x;
}
}  // namespace

extern "C" int spic_permute_species(const spic_grid *grd, const spic_species *src,
                                    const spic_species *dst, uint64_t a, uint64_t b,
                                    void *stream) {
  int rc = spic::check_grid(grd);
  if (rc) return rc;
  if ((rc = spic::check_species(src, true, false)) || (rc = spic::check_species(dst, true, false)))
    return rc;
  if (dst->nop != src->nop) return spic::fail(SPIC_EINVAL, "permute: nop differs");
  const int64_t n = src->nop;
  if (n == 0) return SPIC_OK;
  if (gcd64(a % (uint64_t)n, (uint64_t)n) != 1)
    return spic::fail(SPIC_EINVAL, "permute: a must be coprime to nop");
  void **d = nullptr;
  hipError_t e = hipMalloc((void **)&d, 7 * sizeof(void *));
  if (e != hipSuccess) return spic::hip_fail(e, "permute");
  void *h[7] = {dst->x, dst->y, dst->z, dst->u, dst->v, dst->w, dst->q};
  hipStream_t s = (hipStream_t)stream;
  e = hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, s);
  spic::Src6 s6 = {{src->x, src->y, src->z, src->u, src->v, src->w}};
  long blocks = (long)((n + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  if (e == hipSuccess) {
    if (grd->prec == SPIC_F64)
      hipLaunchKernelGGL(spic::k_permute<double>, dim3((unsigned)blocks), dim3(256), 0, s, s6,
                         src->q, d, n, a % (uint64_t)n, b % (uint64_t)n);
    else
      hipLaunchKernelGGL(spic::k_permute<float>, dim3((unsigned)blocks), dim3(256), 0, s, s6,
                         src->q, d, n, a % (uint64_t)n, b % (uint64_t)n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  return e == hipSuccess ? SPIC_OK : spic::hip_fail(e, "permute");
}
