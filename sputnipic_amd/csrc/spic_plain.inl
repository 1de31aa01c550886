// spic_plain.inl -- caller-order ("plain") gfx950 kernels for mover_PC,
// interpP2G and the fused mover_interp.  One lane per particle; SoA loads
// and stores are coalesced; field and node-coordinate gathers go through
// L2; the deposit adds FP64 with hardware global atomics.
//
// Included by two translation units: spic_plain_fast.hip (SPIC_STRICT 0,
// FMA contraction allowed) and spic_plain_strict.hip (SPIC_STRICT 1, built
// with -ffp-contract=off).  In strict mode every operation keeps the
// reference's operand types and order (src/Mover.cpp:4-162, 385-572), so an
// FP64 or FP32 particle state matches the CPU reference bit for bit.
#include "spic_internal.hpp"

#ifndef SPIC_NS
#error "define SPIC_NS and SPIC_STRICT before including spic_plain.inl"
#endif

namespace spic {
namespace SPIC_NS {

constexpr bool kStrict = (SPIC_STRICT != 0);

// Mover cell index: 2 + int((x - xStart) * invdx), truncating toward zero
// (src/Mover.cpp:39-41).  The reference evaluates the product in double
// (xStart is double).  The argument is clamped first so that the int
// conversion is defined for any input; any index outside [1, n-1] is
// flagged by the caller.
template <typename T>
__device__ __forceinline__ int cell_trunc(T p, double start, T inv, int n) {
  if (kStrict || sizeof(T) == 8) {
    double c = ((double)p - start) * (double)inv;
    c = fmin(fmax(c, -8.0), (double)n + 8.0);
    return 2 + (int)c;
  } else {
    float c = ((float)p - (float)start) * (float)inv;
    c = fminf(fmaxf(c, -8.0f), (float)n + 8.0f);
    return 2 + (int)c;
  }
}

// Deposit cell index: 2 + int(floor((x - xStart) * invdx))
// (src/Mover.cpp:399-401).
template <typename T>
__device__ __forceinline__ int cell_floor(T p, double start, T inv, int n) {
  if (kStrict || sizeof(T) == 8) {
    double c = ((double)p - start) * (double)inv;
    c = fmin(fmax(c, -8.0), (double)n + 8.0);
    return 2 + (int)floor(c);
  } else {
    float c = ((float)p - (float)start) * (float)inv;
    c = fminf(fmaxf(c, -8.0f), (float)n + 8.0f);
    return 2 + (int)floorf(c);
  }
}

// Both nodes ix-1 and ix must exist: 1 <= ix <= n-1.  Out-of-range indices
// are clamped (so no access leaves the arrays) and reported.
__device__ __forceinline__ bool clamp_index(int &ix, int n) {
  bool bad = (ix < 1) | (ix > n - 1);
  ix = min(max(ix, 1), n - 1);
  return bad;
}

// One particle's predictor-corrector push and particle BC
// (src/Mover.cpp:33-157).  Returns true if an iterate left the grid by
// more than one cell.
template <typename T>
__device__ __forceinline__ bool push_particle_once(const KGrid<T> &g,
                                              const KField<T> &f,
                                              const KPush<T> &k, T &x, T &y,
                                              T &z, T &u, T &v, T &w) {
  const int nyn = g.nyn, nzn = g.nzn;
  const int sx = nyn * nzn, sy = nzn;
  const T xt = x, yt = y, zt = z;
  T upt = T(0), vpt = T(0), wpt = T(0);
  T xc = x, yc = y, zc = z;
  bool bad = false;
  for (int it = 0; it < k.niter; ++it) {
    int ix = cell_trunc<T>(xc, g.xStart, g.invdx, g.nxn);
    int iy = cell_trunc<T>(yc, g.yStart, g.invdy, g.nyn);
    int iz = cell_trunc<T>(zc, g.zStart, g.invdz, g.nzn);
    bad |= clamp_index(ix, g.nxn);
    bad |= clamp_index(iy, g.nyn);
    bad |= clamp_index(iz, g.nzn);
    const T xi[2] = {xc - g.xn[ix - 1], g.xn[ix] - xc};
    const T eta[2] = {yc - g.yn[iy - 1], g.yn[iy] - yc};
    const T zeta[2] = {zc - g.zn[iz - 1], g.zn[iz] - zc};
    const int base = (ix * nyn + iy) * nzn + iz;
    T Exl = T(0), Eyl = T(0), Ezl = T(0), Bxl = T(0), Byl = T(0), Bzl = T(0);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const T wgt = xi[a] * eta[b] * zeta[c] * g.invVOL;
          const int n = base - a * sx - b * sy - c;
          Exl += wgt * f.Ex[n];
          Eyl += wgt * f.Ey[n];
          Ezl += wgt * f.Ez[n];
          Bxl += wgt * f.Bx[n];
          Byl += wgt * f.By[n];
          Bzl += wgt * f.Bz[n];
        }
    const T q2 = k.qomdt2;
    const T omdtsq = q2 * q2 * (Bxl * Bxl + Byl * Byl + Bzl * Bzl);
    T denom;
    if (kStrict || sizeof(T) == 8)
      denom = (T)(1.0 / (1.0 + (double)omdtsq));  // double, as Mover.cpp:71
    else
      denom = T(1) / (T(1) + omdtsq);
    const T ut = u + q2 * Exl;
    const T vt = v + q2 * Eyl;
    const T wt = w + q2 * Ezl;
    const T udotb = ut * Bxl + vt * Byl + wt * Bzl;
    upt = (ut + q2 * (vt * Bzl - wt * Byl + q2 * udotb * Bxl)) * denom;
    vpt = (vt + q2 * (wt * Bxl - ut * Bzl + q2 * udotb * Byl)) * denom;
    wpt = (wt + q2 * (ut * Byl - vt * Bxl + q2 * udotb * Bzl)) * denom;
    xc = xt + upt * k.dto2;
    yc = yt + vpt * k.dto2;
    zc = zt + wpt * k.dto2;
  }
  if (kStrict || sizeof(T) == 8) {
    // 2.0 * uptilde - u in double, then narrowed (Mover.cpp:91-93)
    u = (T)(2.0 * (double)upt - (double)u);
    v = (T)(2.0 * (double)vpt - (double)v);
    w = (T)(2.0 * (double)wpt - (double)w);
  } else {
    u = T(2) * upt - u;
    v = T(2) * vpt - v;
    w = T(2) * wpt - w;
  }
  x = xt + upt * k.dt_sub;
  y = yt + vpt * k.dt_sub;
  z = zt + wpt * k.dt_sub;
  // Particle BC: two independent tests per axis, compared and updated in
  // double as the reference does (Lx is double; Mover.cpp:103-157).
  if ((double)x > g.Lx) {
    if (k.perx) x = (T)((double)x - g.Lx);
    else { u = -u; x = (T)(2 * g.Lx - (double)x); }
  }
  if (x < T(0)) {
    if (k.perx) x = (T)((double)x + g.Lx);
    else { u = -u; x = -x; }
  }
  if ((double)y > g.Ly) {
    if (k.pery) y = (T)((double)y - g.Ly);
    else { v = -v; y = (T)(2 * g.Ly - (double)y); }
  }
  if (y < T(0)) {
    if (k.pery) y = (T)((double)y + g.Ly);
    else { v = -v; y = -y; }
  }
  if ((double)z > g.Lz) {
    if (k.perz) z = (T)((double)z - g.Lz);
    else { w = -w; z = (T)(2 * g.Lz - (double)z); }
  }
  if (z < T(0)) {
    if (k.perz) z = (T)((double)z + g.Lz);
    else { w = -w; z = -z; }
  }
  return bad;
}

// The push of one particle: once, or k.nloop times with the particle BC
// after each (the reference GPU mover's sub-cycle loop,
// src/gpu/Particles_gpu.cu:336-437).
template <typename T>
__device__ __forceinline__ bool push_particle(const KGrid<T> &g, const KField<T> &f,
                                              const KPush<T> &k, T &x, T &y, T &z, T &u,
                                              T &v, T &w) {
  bool bad = false;
#pragma unroll 1
  for (int isub = 0; isub < k.nloop; ++isub)
    bad |= push_particle_once<T>(g, f, k, x, y, z, u, v, w);
  return bad;
}

// One particle's ten-moment CIC deposit (src/Mover.cpp:399-570): per node
// weight q*xi*eta*zeta*invVOL, every moment term times invVOL again, added
// in FP64.  Returns true if the cell index left the grid.
template <typename T>
__device__ __forceinline__ bool deposit_particle(const KGrid<T> &g,
                                                 const KMom &m, T x, T y,
                                                 T z, T u, T v, T w, T q) {
  const int nyn = g.nyn, nzn = g.nzn;
  const int sx = nyn * nzn, sy = nzn;
  int ix = cell_floor<T>(x, g.xStart, g.invdx, g.nxn);
  int iy = cell_floor<T>(y, g.yStart, g.invdy, g.nyn);
  int iz = cell_floor<T>(z, g.zStart, g.invdz, g.nzn);
  bool bad = clamp_index(ix, g.nxn);
  bad |= clamp_index(iy, g.nyn);
  bad |= clamp_index(iz, g.nzn);
  const T xi[2] = {x - g.xn[ix - 1], g.xn[ix] - x};
  const T eta[2] = {y - g.yn[iy - 1], g.yn[iy] - y};
  const T zeta[2] = {z - g.zn[iz - 1], g.zn[iz] - z};
  const T iv = g.invVOL;
  const T uu = u * u, uv = u * v, uw = u * w, vv = v * v, vw = v * w,
          ww = w * w;
  const int base = (ix * nyn + iy) * nzn + iz;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const T wgt = q * xi[a] * eta[b] * zeta[c] * iv;
        const int n = base - a * sx - b * sy - c;
        unsafeAtomicAdd(m.m[0] + n, (double)(wgt * iv));
        unsafeAtomicAdd(m.m[1] + n, (double)((u * wgt) * iv));
        unsafeAtomicAdd(m.m[2] + n, (double)((v * wgt) * iv));
        unsafeAtomicAdd(m.m[3] + n, (double)((w * wgt) * iv));
        unsafeAtomicAdd(m.m[4] + n, (double)((uu * wgt) * iv));
        unsafeAtomicAdd(m.m[5] + n, (double)((uv * wgt) * iv));
        unsafeAtomicAdd(m.m[6] + n, (double)((uw * wgt) * iv));
        unsafeAtomicAdd(m.m[7] + n, (double)((vv * wgt) * iv));
        unsafeAtomicAdd(m.m[8] + n, (double)((vw * wgt) * iv));
        unsafeAtomicAdd(m.m[9] + n, (double)((ww * wgt) * iv));
      }
  return bad;
}

constexpr int kBlock = 256;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mover(KGrid<T> g, KField<T> f,
                                                  KPart<T> p, KPush<T> k,
                                                  unsigned int *status) {
  bool bad = false;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < p.n;
       i += gridDim.x * kBlock) {
    T x = p.x[i], y = p.y[i], z = p.z[i], u = p.u[i], v = p.v[i], w = p.w[i];
    bad |= push_particle<T>(g, f, k, x, y, z, u, v, w);
    p.x[i] = x; p.y[i] = y; p.z[i] = z;
    p.u[i] = u; p.v[i] = v; p.w[i] = w;
  }
  if (bad) atomicAdd(status, 1u);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_deposit(KGrid<T> g, KPart<T> p,
                                                    KMom m,
                                                    unsigned int *status) {
  bool bad = false;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < p.n;
       i += gridDim.x * kBlock)
    bad |= deposit_particle<T>(g, m, p.x[i], p.y[i], p.z[i], p.u[i], p.v[i],
                               p.w[i], p.q[i]);
  if (bad) atomicAdd(status + 1, 1u);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mover_deposit(
    KGrid<T> g, KField<T> f, KPart<T> p, KPush<T> k, KMom m,
    unsigned int *status) {
  bool bad = false, bad_dep = false;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < p.n;
       i += gridDim.x * kBlock) {
    T x = p.x[i], y = p.y[i], z = p.z[i], u = p.u[i], v = p.v[i], w = p.w[i];
    bad |= push_particle<T>(g, f, k, x, y, z, u, v, w);
    p.x[i] = x; p.y[i] = y; p.z[i] = z;
    p.u[i] = u; p.v[i] = v; p.w[i] = w;
    bad_dep |= deposit_particle<T>(g, m, x, y, z, u, v, w, p.q[i]);
  }
  if (bad) atomicAdd(status, 1u);
  if (bad_dep) atomicAdd(status + 1, 1u);
}

static inline unsigned grid_for(int n) {
  long blocks = ((long)n + kBlock - 1) / kBlock;
  if (blocks < 1) blocks = 1;
  if (blocks > (1L << 20)) blocks = 1L << 20;
  return (unsigned)blocks;
}

template <typename T>
hipError_t launch_mover(const KGrid<T> &g, const KField<T> &f,
                        const KPart<T> &p, const KPush<T> &k,
                        unsigned int *status, hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mover<T>, dim3(grid_for(p.n)), dim3(kBlock), 0, s, g,
                     f, p, k, status);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_deposit(const KGrid<T> &g, const KPart<T> &p, const KMom &m,
                          unsigned int *status, hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_deposit<T>, dim3(grid_for(p.n)), dim3(kBlock), 0, s, g,
                     p, m, status);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_mover_deposit(const KGrid<T> &g, const KField<T> &f,
                                const KPart<T> &p, const KPush<T> &k,
                                const KMom &m, unsigned int *status,
                                hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mover_deposit<T>, dim3(grid_for(p.n)), dim3(kBlock), 0,
                     s, g, f, p, k, m, status);
  return hipGetLastError();
}

#define SPIC_INSTANTIATE(T)                                                   \
  template hipError_t launch_mover<T>(const KGrid<T> &, const KField<T> &,    \
                                      const KPart<T> &, const KPush<T> &,     \
                                      unsigned int *, hipStream_t);           \
  template hipError_t launch_deposit<T>(const KGrid<T> &, const KPart<T> &,   \
                                        const KMom &, unsigned int *,         \
                                        hipStream_t);                         \
  template hipError_t launch_mover_deposit<T>(                                \
      const KGrid<T> &, const KField<T> &, const KPart<T> &,                  \
      const KPush<T> &, const KMom &, unsigned int *, hipStream_t);
SPIC_INSTANTIATE(double)
SPIC_INSTANTIATE(float)
#undef SPIC_INSTANTIATE

}  // namespace SPIC_NS
}  // namespace spic
