// spic_moments.hip -- the steps that follow the deposit, on the device:
// the ghost-node fold of each species' node moments (applyBCids,
// src/BC.cpp:362-619) and the sum over species into the net moments
// (sumOverSpecies, src/InterpDensNet.cpp:120-144).  Both work in place on
// the FP64 moment arrays and add in the reference's order, so FP64 results
// are bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include "spic.h"
#include "spic_internal.hpp"

namespace spic {
namespace {

constexpr int kMaxSets = 8;

// Up to kMaxSets moment sets, ten arrays each: array (set * 10 + moment).
struct MomentSets {
  double *a[kMaxSets * 10];
};

// The ten arrays of one moment set in the reference's order.
inline void moment_arrays(const spic_moments &m, double **out) {
  double *const a[10] = {m.rhon, m.Jx,  m.Jy,  m.Jz,  m.pxx,
                         m.pxy,  m.pxz, m.pyy, m.pyz, m.pzz};
  for (int k = 0; k < 10; ++k) out[k] = a[k];
}

// One fold step along `axis` (0 = x, 1 = y, 2 = z) for every array:
// periodic: a[1] += a[n-2]; a[n-2] = a[1]   (src/BC.cpp:369-494)
// otherwise: a[1] *= 2;     a[n-2] *= 2      (src/BC.cpp:495-618)
// over the interior 1..n-2 of the two other axes.  blockIdx.y selects
// the array; the step's elements are independent.
__global__ void k_fold(MomentSets s, int axis, int periodic, int nxn, int nyn,
                       int nzn) {
  double *a = s.a[blockIdx.y];
  const int n0 = axis == 0 ? nyn : nxn;         // first free axis
  const int n1 = axis == 2 ? nyn : nzn;         // second free axis
  const int m0 = n0 - 2, m1 = n1 - 2;           // interior extents
  const long plane = (long)m0 * m1;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < plane;
       t += (long)gridDim.x * blockDim.x) {
    const int u = 1 + (int)(t / m1), v = 1 + (int)(t % m1);
    long lo, hi;
    if (axis == 0) {
      lo = ((long)1 * nyn + u) * nzn + v;
      hi = ((long)(nxn - 2) * nyn + u) * nzn + v;
    } else if (axis == 1) {
      lo = ((long)u * nyn + 1) * nzn + v;
      hi = ((long)u * nyn + (nyn - 2)) * nzn + v;
    } else {
      lo = ((long)u * nyn + v) * nzn + 1;
      hi = ((long)u * nyn + v) * nzn + (nzn - 2);
    }
    if (periodic) {
      a[lo] += a[hi];
      a[hi] = a[lo];
    } else {
      a[lo] *= 2;
      a[hi] *= 2;
    }
  }
}

// net[m][n] += ids[0][m][n], then += ids[1][m][n], ... per element.
__global__ void k_sum_species(MomentSets net, MomentSets ids, int ns, long nn) {
  const int m = blockIdx.y;
  double *dst = net.a[m];
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn;
       n += (long)gridDim.x * blockDim.x) {
    double v = dst[n];
    for (int s = 0; s < ns; ++s) v += ids.a[s * 10 + m][n];
    dst[n] = v;
  }
}

inline unsigned nblocks(long n) {
  long b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 65535) b = 65535;
  return (unsigned)b;
}

int finish(hipStream_t s, const char *what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !async_mode()) e = hipStreamSynchronize(s);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, what);
}

}  // namespace
}  // namespace spic

extern "C" {

int spic_apply_bc_ids(const spic_grid *grd, const spic_moments *mom, int ns,
                      const spic_push *push, void *stream) {
  int rc;
  if ((rc = spic::check_grid(grd)) || (rc = spic::check_push(push))) return rc;
  if (!mom) return spic::fail(SPIC_EINVAL, "moments are NULL");
  if (ns < 1 || ns > spic::kMaxSets)
    return spic::fail(SPIC_EINVAL, "moment set count must be 1..%d", spic::kMaxSets);
  spic::MomentSets sets;
  for (int i = 0; i < ns; ++i) {
    if ((rc = spic::check_moments(&mom[i]))) return rc;
    spic::moment_arrays(mom[i], &sets.a[i * 10]);
  }
  hipStream_t s = (hipStream_t)stream;
  const int nn[3] = {grd->nxn, grd->nyn, grd->nzn};
  const int per[3] = {push->PERIODICX, push->PERIODICY, push->PERIODICZ};
  // the reference's order: the periodic axes x, y, z, then the others
  for (int pass = 0; pass < 2; ++pass)
    for (int axis = 0; axis < 3; ++axis) {
      const int periodic = per[axis] ? 1 : 0;
      if (periodic != (pass == 0 ? 1 : 0)) continue;
      const long plane = (long)(nn[axis == 0 ? 1 : 0] - 2) * (nn[axis == 2 ? 1 : 2] - 2);
      hipLaunchKernelGGL(spic::k_fold, dim3(spic::nblocks(plane), ns * 10), dim3(256),
                         0, s, sets, axis, periodic, grd->nxn, grd->nyn, grd->nzn);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return spic::hip_fail(e, "applyBCids launch");
    }
  return spic::finish(s, "applyBCids");
}

int spic_sum_species(const spic_grid *grd, const spic_moments *net,
                     const spic_moments *ids, int ns, void *stream) {
  int rc;
  if ((rc = spic::check_grid(grd)) || (rc = spic::check_moments(net))) return rc;
  if (!ids) return spic::fail(SPIC_EINVAL, "species moments are NULL");
  if (ns < 1 || ns > spic::kMaxSets)
    return spic::fail(SPIC_EINVAL, "species count must be 1..%d", spic::kMaxSets);
  spic::MomentSets src, dst;
  for (int i = 0; i < ns; ++i) {
    if ((rc = spic::check_moments(&ids[i]))) return rc;
    spic::moment_arrays(ids[i], &src.a[i * 10]);
  }
  spic::moment_arrays(*net, dst.a);
  const long nn = (long)grd->nxn * grd->nyn * grd->nzn;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(spic::k_sum_species, dim3(spic::nblocks(nn), 10), dim3(256), 0, s,
                     dst, src, ns, nn);
  return spic::finish(s, "sumOverSpecies");
}

}  // extern "C"
