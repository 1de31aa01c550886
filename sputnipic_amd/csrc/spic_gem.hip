// spic_gem.hip -- synthetic Harris-sheet (GEM) input on the device, for
// runs far larger than the CPU IC can produce.  Restates the formulas of
// initGEM (src/IC.cpp:38-177) and interpN2Crho
// (src/InterpDensSpecies.cpp:64-77):
//   * fields on every node (ghosts included), E = 0,
//     Bx = B0x tanh((Y-Ly/2)/delta) + the localized X-point perturbation,
//     By = B0y + perturbation, Bz = B0z;
//   * particles on the reference's regular sub-lattice in cell-major order
//     (i, j, k, ii, jj, kk with the kk < npcely bound of IC.cpp:128),
//     q = sign(qom) * rhoc / npcel / invVOL with rhoc taken from a per-y
//     table the host forms exactly as interpN2Crho does, and Maxwellian
//     velocities by the reference's Box-Muller formulas.
// The only difference from initGEM is the random source: instead of
// glibc rand() seeded per species, each draw hashes the counter
// (seed, species, particle, draw) with SplitMix64 (Steele, Lea & Flood,
// OOPSLA 2014), so any particle can be generated independently.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "spic.h"
#include "spic_internal.hpp"

namespace spic {
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform double in [0, 1] with 53 random bits
__device__ __forceinline__ double uniform01(uint64_t key) {
  return (double)(splitmix64(key) >> 11) * (1.0 / 9007199254740991.0);
}

template <typename T>
__global__ void k_gem_fields(int nxn, int nyn, int nzn, const T *__restrict__ xn,
                             const T *__restrict__ yn, double Lx, double Ly,
                             double B0x, double B0y, double B0z, double delta,
                             double pertX, T *Ex, T *Ey, T *Ez, T *Bx, T *By,
                             T *Bz) {
  const long nn = (long)nxn * nyn * nzn;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn;
       n += (long)gridDim.x * blockDim.x) {
    const int i = (int)(n / ((long)nyn * nzn));
    const int j = (int)((n / nzn) % nyn);
    const double X = (double)xn[i], Y = (double)yn[j];
    Ex[n] = T(0);
    Ey[n] = T(0);
    Ez[n] = T(0);
    const T bx0 = (T)(B0x * tanh((Y - Ly / 2) / delta));
    const T by0 = (T)B0y;
    const double xpert = X - Lx / 2, ypert = Y - Ly / 2;
    const double exp_pert = exp(-(xpert / delta) * (xpert / delta) -
                                (ypert / delta) * (ypert / delta));
    const double cx = cos(M_PI * xpert / 10.0 / delta);
    const double sx = sin(M_PI * xpert / 10.0 / delta);
    const double cy = cos(M_PI * ypert / 10.0 / delta);
    const double sy = sin(M_PI * ypert / 10.0 / delta);
    Bx[n] = (T)(bx0 + (B0x * pertX) * exp_pert *
                          (-cx * cy * 2.0 * ypert / delta - cx * sy * M_PI / 10.0));
    By[n] = (T)(by0 + (B0x * pertX) * exp_pert *
                          (cx * cy * 2.0 * xpert / delta + sx * cy * M_PI / 10.0));
    Bz[n] = (T)B0z;
  }
}

struct GemSpecies {
  int ncx, ncy, ncz;        // physical cells
  int npcelx, npcely, npcelz;
  double dx, dy, dz;
  double u0, v0, w0, uth, vth, wth;
  uint64_t key;             // (seed, species) part of the counter
};

// q per y cell row: sign * (rhoc/npcel) in FPinterp, times 1/invVOL in
// double, formed by the host (see spic_gem_particles).
template <typename T>
__global__ void k_gem_particles(GemSpecies s, const T *__restrict__ xn,
                                const T *__restrict__ yn,
                                const T *__restrict__ zn,
                                const double *__restrict__ q_row, long n,
                                T *x, T *y, T *z, T *u, T *v, T *w, T *q) {
  const int pz = s.npcely;  // kk < npcely, IC.cpp:128
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n;
       p += (long)gridDim.x * blockDim.x) {
    long r = p;
    const int kk = (int)(r % pz); r /= pz;
    const int jj = (int)(r % s.npcely); r /= s.npcely;
    const int ii = (int)(r % s.npcelx); r /= s.npcelx;
    const long cell = r;
    const int k = (int)(cell % s.ncz);
    const int j = (int)((cell / s.ncz) % s.ncy);
    const int i = (int)(cell / ((long)s.ncz * s.ncy));
    x[p] = (T)((ii + .5) * (s.dx / s.npcelx) + (double)xn[i + 1]);
    y[p] = (T)((jj + .5) * (s.dy / s.npcely) + (double)yn[j + 1]);
    z[p] = (T)((kk + .5) * (s.dz / s.npcelz) + (double)zn[k + 1]);
    q[p] = (T)q_row[j + 1];
    const uint64_t base = (s.key + (uint64_t)p) * 4ull;
    double h = uniform01(base + 0);
    double prob = sqrt(-2.0 * log(1.0 - .999999 * h));
    double theta = 2.0 * M_PI * uniform01(base + 1);
    u[p] = (T)(s.u0 + s.uth * prob * cos(theta));
    h = uniform01(base + 2);
    prob = sqrt(-2.0 * log(1.0 - .999999 * h));
    theta = 2.0 * M_PI * uniform01(base + 3);
    w[p] = (T)(s.w0 + s.wth * prob * cos(theta));
    v[p] = (T)(s.v0 + s.vth * prob * sin(theta));
  }
}

inline unsigned blocks(long n) {
  long b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > (1L << 20)) b = 1L << 20;
  return (unsigned)b;
}

}  // namespace

hipError_t gem_fields(const spic_grid &g, const spic_gem_params &p,
                      const spic_field &f, hipStream_t s) {
  const long nn = (long)g.nxn * g.nyn * g.nzn;
  if (g.prec == SPIC_F64)
    hipLaunchKernelGGL(k_gem_fields<double>, dim3(blocks(nn)), dim3(256), 0, s,
                       g.nxn, g.nyn, g.nzn, (const double *)g.xn,
                       (const double *)g.yn, g.Lx, g.Ly, p.B0x, p.B0y, p.B0z,
                       p.delta, p.pertX, (double *)f.Ex, (double *)f.Ey,
                       (double *)f.Ez, (double *)f.Bxn, (double *)f.Byn,
                       (double *)f.Bzn);
  else
    hipLaunchKernelGGL(k_gem_fields<float>, dim3(blocks(nn)), dim3(256), 0, s,
                       g.nxn, g.nyn, g.nzn, (const float *)g.xn,
                       (const float *)g.yn, g.Lx, g.Ly, p.B0x, p.B0y, p.B0z,
                       p.delta, p.pertX, (float *)f.Ex, (float *)f.Ey,
                       (float *)f.Ez, (float *)f.Bxn, (float *)f.Byn,
                       (float *)f.Bzn);
  return hipGetLastError();
}

hipError_t gem_particles(const spic_grid &g, const spic_gem_species &sp,
                         const double *q_row_dev, spic_species &out,
                         hipStream_t s) {
  GemSpecies k;
  k.ncx = g.nxc - 2;
  k.ncy = g.nyc - 2;
  k.ncz = g.nzc - 2;
  k.npcelx = sp.npcelx;
  k.npcely = sp.npcely;
  k.npcelz = sp.npcelz;
  k.dx = g.dx;
  k.dy = g.dy;
  k.dz = g.dz;
  k.u0 = sp.u0; k.v0 = sp.v0; k.w0 = sp.w0;
  k.uth = sp.uth; k.vth = sp.vth; k.wth = sp.wth;
  k.key = ((uint64_t)sp.seed << 40) ^ ((uint64_t)sp.species << 36) ^
          0x5DEECE66Dull;
  const long n = (long)out.npmax;
  if (n <= 0) return hipSuccess;
  if (g.prec == SPIC_F64)
    hipLaunchKernelGGL(k_gem_particles<double>, dim3(blocks(n)), dim3(256), 0,
                       s, k, (const double *)g.xn, (const double *)g.yn,
                       (const double *)g.zn, q_row_dev, n, (double *)out.x,
                       (double *)out.y, (double *)out.z, (double *)out.u,
                       (double *)out.v, (double *)out.w, (double *)out.q);
  else
    hipLaunchKernelGGL(k_gem_particles<float>, dim3(blocks(n)), dim3(256), 0,
                       s, k, (const float *)g.xn, (const float *)g.yn,
                       (const float *)g.zn, q_row_dev, n, (float *)out.x,
                       (float *)out.y, (float *)out.z, (float *)out.u,
                       (float *)out.v, (float *)out.w, (float *)out.q);
  return hipGetLastError();
}

}  // namespace spic

// ---------------------------------------------------------------------------
// extern "C" entry points
// ---------------------------------------------------------------------------
namespace {

// Per-y q row of one species, formed as initGEM forms it: rhon on the y
// nodes (IC.cpp:44-53), rhoc as interpN2Crho's eight-node sum in FPinterp
// (InterpDensSpecies.cpp:70-75; the eight terms depend on j only), then
// q = (qom/|qom|) * (rhoc/npcel) * (1.0/invVOL) (IC.cpp:138-140).
template <typename T>
void q_rows(const spic_grid &g, const spic_gem_params &p,
            const spic_gem_species &sp, const T *yn, double *q_row) {
  const int nyn = g.nyn;
  T *r = new T[nyn];
  for (int j = 0; j < nyn; ++j) {
    if (sp.sheet) {
      double ch = cosh(((double)yn[j] - g.Ly / 2) / p.delta);
      r[j] = (T)((T)(sp.rhoINIT / (ch * ch)) / p.fourpi);
    } else {
      r[j] = (T)((T)sp.rhoINIT / p.fourpi);
    }
  }
  const T qom = (T)sp.qom;
  const T sgn = qom / (T)std::fabs((double)qom);
  const int npcel = sp.npcelx * sp.npcely * sp.npcelz;
  const T invVOL = (T)g.invVOL;
  for (int j = 0; j < g.nyc; ++j) {
    if (j == 0 || j == g.nyc - 1) {
      q_row[j] = 0.0;
      continue;
    }
    const T a = r[j], b = r[j + 1];
    T s = a + a;
    s = s + b;
    s = s + a;
    s = s + b;
    s = s + a;
    s = s + b;
    s = s + b;
    const T rhoc = (T)(.125 * (double)s);
    const T rq = rhoc / (T)npcel;
    q_row[j] = (double)(T)((double)(sgn * rq) * (1.0 / (double)invVOL));
  }
  delete[] r;
}

}  // namespace

extern "C" int spic_gem_fields(const spic_grid *grd, const spic_gem_params *gem,
                               spic_field *out, void *stream) {
  if (!grd || !gem || !out || !out->Ex || !out->Ey || !out->Ez || !out->Bxn ||
      !out->Byn || !out->Bzn || !grd->xn || !grd->yn)
    return spic::fail(SPIC_EINVAL, "spic_gem_fields: NULL argument");
  if (grd->prec != SPIC_F64 && grd->prec != SPIC_F32)
    return spic::fail(SPIC_EINVAL, "spic_gem_fields: bad precision");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = spic::gem_fields(*grd, *gem, *out, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? SPIC_OK : spic::hip_fail(e, "spic_gem_fields");
}

extern "C" int spic_gem_particles(const spic_grid *grd,
                                  const spic_gem_params *gem,
                                  const spic_gem_species *sp,
                                  spic_species *out, void *stream) {
  if (!grd || !gem || !sp || !out || !grd->xn || !grd->yn || !grd->zn)
    return spic::fail(SPIC_EINVAL, "spic_gem_particles: NULL argument");
  if (grd->prec != SPIC_F64 && grd->prec != SPIC_F32)
    return spic::fail(SPIC_EINVAL, "spic_gem_particles: bad precision");
  if (sp->npcelx < 1 || sp->npcely < 1 || sp->npcelz < 1)
    return spic::fail(SPIC_EINVAL, "spic_gem_particles: particles per cell < 1");
  const long cells = (long)(grd->nxc - 2) * (grd->nyc - 2) * (grd->nzc - 2);
  const long want = cells * sp->npcelx * sp->npcely * sp->npcely;
  if (out->npmax != want)
    return spic::fail(SPIC_EINVAL,
                      "spic_gem_particles: npmax %lld != cells*npcelx*npcely^2 = %ld",
                      (long long)out->npmax, want);
  if (out->npmax > 0 && (!out->x || !out->y || !out->z || !out->u ||
                         !out->v || !out->w || !out->q))
    return spic::fail(SPIC_EINVAL, "spic_gem_particles: NULL particle array");
  hipStream_t s = (hipStream_t)stream;
  const size_t esz = grd->prec == SPIC_F64 ? 8 : 4;
  std::vector<char> yn_host(esz * grd->nyn);
  hipError_t e = hipMemcpy(yn_host.data(), grd->yn, esz * grd->nyn,
                           hipMemcpyDeviceToHost);
  if (e != hipSuccess) return spic::hip_fail(e, "spic_gem_particles: yn");
  std::vector<double> q_row(grd->nyc);
  if (grd->prec == SPIC_F64)
    q_rows<double>(*grd, *gem, *sp, (const double *)yn_host.data(), q_row.data());
  else
    q_rows<float>(*grd, *gem, *sp, (const float *)yn_host.data(), q_row.data());
  double *d_q = nullptr;
  if ((e = hipMalloc(&d_q, sizeof(double) * grd->nyc)) != hipSuccess)
    return spic::hip_fail(e, "spic_gem_particles: q rows");
  e = hipMemcpy(d_q, q_row.data(), sizeof(double) * grd->nyc,
                hipMemcpyHostToDevice);
  if (e == hipSuccess) e = spic::gem_particles(*grd, *sp, d_q, *out, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_q);
  return e == hipSuccess ? SPIC_OK : spic::hip_fail(e, "spic_gem_particles");
}

extern "C" int spic_event_create(void **event) {
  if (!event) return spic::fail(SPIC_EINVAL, "event is NULL");
  hipEvent_t ev;
  hipError_t e = hipEventCreate(&ev);
  if (e != hipSuccess) return spic::hip_fail(e, "hipEventCreate");
  *event = (void *)ev;
  return SPIC_OK;
}

extern "C" int spic_event_destroy(void *event) {
  hipError_t e = hipEventDestroy((hipEvent_t)event);
  return e == hipSuccess ? SPIC_OK : spic::hip_fail(e, "hipEventDestroy");
}

extern "C" int spic_event_record(void *event, void *stream) {
  hipError_t e = hipEventRecord((hipEvent_t)event, (hipStream_t)stream);
  return e == hipSuccess ? SPIC_OK : spic::hip_fail(e, "hipEventRecord");
}

extern "C" int spic_event_elapsed_ms(float *ms, void *start, void *end) {
  if (!ms) return spic::fail(SPIC_EINVAL, "ms is NULL");
  hipError_t e = hipEventSynchronize((hipEvent_t)end);
  if (e == hipSuccess)
    e = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end);
  return e == hipSuccess ? SPIC_OK : spic::hip_fail(e, "hipEventElapsedTime");
}
