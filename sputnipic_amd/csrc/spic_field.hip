// spic_field.hip -- the field-solve side that consumes the deposit, on the
// device and in FP64 (the field solve stays in the reference's FP64):
//   ghost updates of scalar arrays   applyBCscalarDensN / FieldNzero / DensC /
//                                    FieldC (src/BC.cpp:11-123, 242-352,
//                                    623-735, 737-850)
//   net rhoc from the net rhon       src/sputniPIC_CPU.cpp:276-277
//                                    (applyBCscalarDensN + interpN2Cinterp,
//                                    src/Grid.cpp:164-185)
//   hat densities                    calculateHatDensities
//                                    (src/MaxwellSolver.cpp:465-604)
//   implicit field operator          MaxwellImage (src/MaxwellSolver.cpp:95-257)
// This file is compiled without FMA contraction and keeps the reference's
// operation order in every expression, so results are bit-identical to the
// reference's FP64 build.  Volume work is one thread per interior node or
// cell; where the reference makes a temporary array and then transforms it
// (scale, interpolate, add, negate, scale, sum), the transform is applied
// in the kernel that produces the value, in the same order.
#include <hip/hip_runtime.h>

#include "spic.h"
#include "spic_binned.hpp"
#include "spic_internal.hpp"

namespace spic {
namespace {

constexpr int kMaxBcArrays = 3 * kMaxSpecies;

struct BcArrays {
  double *a[kMaxBcArrays];
};

// Ghost updates, the reference's sequence (periodic faces x, y, z; edges
// along x, y, z; corners; non-periodic faces x, y, z, copied from the
// adjacent plane or zeroed) in two launches with the same result:
//  * k_bc_periodic: the periodic faces, edges and corners.  A node array
//    takes its ghosts from planes n-3 and 2, a cell array from n-2 and 1.
//    Every element these steps read is interior and no two steps write the
//    same element, so they run together.  The reference runs the y-edge
//    loop to the node count (j < nyn - 1), which for a cell array also
//    writes the four corners at j = nyc - 1; the corner step, which runs
//    whenever the y-edge step does, overwrites those, so they are skipped.
//  * k_bc_faces: the non-periodic faces.  Applied in x, y, z order, the
//    plane copies leave at every element with a ghost index on a
//    non-periodic axis the value of the element with those indices moved
//    one plane inward (clamped), read after the periodic steps; zeroed
//    faces leave zero there.
__global__ void k_bc_periodic(BcArrays arrs, int n0, int n1, int n2, int nyn, int nodes,
                              int px, int py, int pz) {
  SPIC_GLOBAL double *a = gmem(arrs.a[blockIdx.y]);
  const int lo = nodes ? 3 : 2, hi = nodes ? 2 : 1;
#define A(i, j, k) a[((int64_t)(i) * n1 + (j)) * n2 + (k)]
  const int64_t fx = px ? (int64_t)(n1 - 2) * (n2 - 2) : 0;
  const int64_t fy = py ? (int64_t)(n0 - 2) * (n2 - 2) : 0;
  const int64_t fz = pz ? (int64_t)(n0 - 2) * (n1 - 2) : 0;
  const int64_t ex = (py || pz) ? n0 - 2 : 0;
  // y-edge rows j = 1 .. nyn - 2, without the cell corner row j = n1 - 1
  const int yend = nyn - 1 < n1 - 1 ? nyn - 1 : n1 - 1;
  const int64_t ey = (px || pz) ? yend - 1 : 0;
  const int64_t ez = (px || py) ? n2 - 2 : 0;
  const int64_t co = (px || py || pz) ? 1 : 0;
  const int64_t total = fx + fy + fz + ex + ey + ez + co;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = e;
    if (r < fx) {
      const int j = 1 + (int)(r / (n2 - 2)), k = 1 + (int)(r % (n2 - 2));
      A(0, j, k) = A(n0 - lo, j, k);
      A(n0 - 1, j, k) = A(hi, j, k);
      continue;
    }
    r -= fx;
    if (r < fy) {
      const int i = 1 + (int)(r / (n2 - 2)), k = 1 + (int)(r % (n2 - 2));
      A(i, 0, k) = A(i, n1 - lo, k);
      A(i, n1 - 1, k) = A(i, hi, k);
      continue;
    }
    r -= fy;
    if (r < fz) {
      const int i = 1 + (int)(r / (n1 - 2)), j = 1 + (int)(r % (n1 - 2));
      A(i, j, 0) = A(i, j, n2 - lo);
      A(i, j, n2 - 1) = A(i, j, hi);
      continue;
    }
    r -= fz;
    if (r < ex) {
      const int i = 1 + (int)r;
      A(i, n1 - 1, n2 - 1) = A(i, hi, hi);
      A(i, 0, 0) = A(i, n1 - lo, n2 - lo);
      A(i, 0, n2 - 1) = A(i, n1 - lo, hi);
      A(i, n1 - 1, 0) = A(i, hi, n2 - lo);
      continue;
    }
    r -= ex;
    if (r < ey) {
      const int j = 1 + (int)r;
      A(n0 - 1, j, n2 - 1) = A(hi, j, hi);
      A(0, j, 0) = A(n0 - lo, j, n2 - lo);
      A(0, j, n2 - 1) = A(n0 - lo, j, hi);
      A(n0 - 1, j, 0) = A(hi, j, n2 - lo);
      continue;
    }
    r -= ey;
    if (r < ez) {
      const int k = 1 + (int)r;
      A(n0 - 1, n1 - 1, k) = A(hi, hi, k);
      A(0, 0, k) = A(n0 - lo, n1 - lo, k);
      A(n0 - 1, 0, k) = A(hi, n1 - lo, k);
      A(0, n1 - 1, k) = A(n0 - lo, hi, k);
      continue;
    }
    A(n0 - 1, n1 - 1, n2 - 1) = A(hi, hi, hi);
    A(0, n1 - 1, n2 - 1) = A(n0 - lo, hi, hi);
    A(n0 - 1, 0, n2 - 1) = A(hi, n1 - lo, hi);
    A(0, 0, n2 - 1) = A(n0 - lo, n1 - lo, hi);
    A(n0 - 1, n1 - 1, 0) = A(hi, hi, n2 - lo);
    A(0, n1 - 1, 0) = A(n0 - lo, hi, n2 - lo);
    A(n0 - 1, 0, 0) = A(hi, n1 - lo, n2 - lo);
    A(0, 0, 0) = A(n0 - lo, n1 - lo, n2 - lo);
  }
}

// Index moved one plane inward if it is a ghost of a non-periodic axis.
__device__ __forceinline__ int clamp_ghost(int v, int n, int periodic) {
  if (periodic) return v;
  return v == 0 ? 1 : (v == n - 1 ? n - 2 : v);
}

// Every element on the ghost planes of the non-periodic axes (an element
// on two such planes is visited twice and written the same value twice).
__global__ void k_bc_faces(BcArrays arrs, int n0, int n1, int n2, int zero_faces, int px,
                           int py, int pz) {
  SPIC_GLOBAL double *a = gmem(arrs.a[blockIdx.y]);
  const int64_t sx = px ? 0 : 2 * (int64_t)n1 * n2;
  const int64_t sy = py ? 0 : 2 * (int64_t)n0 * n2;
  const int64_t sz = pz ? 0 : 2 * (int64_t)n0 * n1;
  const int64_t total = sx + sy + sz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    int64_t r = e;
    if (r < sx) {
      const int64_t plane = (int64_t)n1 * n2;
      i = r < plane ? 0 : n0 - 1;
      r %= plane;
      j = (int)(r / n2);
      k = (int)(r % n2);
    } else if ((r -= sx) < sy) {
      const int64_t plane = (int64_t)n0 * n2;
      j = r < plane ? 0 : n1 - 1;
      r %= plane;
      i = (int)(r / n2);
      k = (int)(r % n2);
    } else {
      r -= sy;
      const int64_t plane = (int64_t)n0 * n1;
      k = r < plane ? 0 : n2 - 1;
      r %= plane;
      i = (int)(r / n1);
      j = (int)(r % n1);
    }
    A(i, j, k) = zero_faces ? 0.0
                            : A(clamp_ghost(i, n0, px), clamp_ghost(j, n1, py),
                                clamp_ghost(k, n2, pz));
  }
#undef A
}

// Interior point (i, j, k) of an n0 x n1 x n2 array for thread index t
// (t < (n0 - 2)(n1 - 2)(n2 - 2)).
__device__ __forceinline__ void interior(int64_t t, int n1, int n2, int &i, int &j, int &k) {
  const int m1 = n1 - 2, m2 = n2 - 2;
  k = 1 + (int)(t % m2);
  j = 1 + (int)((t / m2) % m1);
  i = 1 + (int)(t / ((int64_t)m2 * m1));
}

struct Dims {
  int nxn, nyn, nzn, nxc, nyc, nzc;
  int64_t nn, nc;
  int64_t in_nodes, in_cells;  // interior counts
};

inline Dims dims_of(const spic_grid *g) {
  Dims d;
  d.nxn = g->nxn; d.nyn = g->nyn; d.nzn = g->nzn;
  d.nxc = g->nxc; d.nyc = g->nyc; d.nzc = g->nzc;
  d.nn = (int64_t)d.nxn * d.nyn * d.nzn;
  d.nc = (int64_t)d.nxc * d.nyc * d.nzc;
  d.in_nodes = (int64_t)(d.nxn - 2) * (d.nyn - 2) * (d.nzn - 2);
  d.in_cells = (int64_t)(d.nxc - 2) * (d.nyc - 2) * (d.nzc - 2);
  return d;
}

#define NIDX(i, j, k) (((int64_t)(i) * d.nyn + (j)) * d.nzn + (k))
#define CIDX(i, j, k) (((int64_t)(i) * d.nyc + (j)) * d.nzc + (k))

// interpN2Cinterp (src/Grid.cpp:164-185) on the interior cells.
__global__ void k_n2c(Dims d, const double *__restrict__ rhon, double *__restrict__ rhoc) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_cells) return;
  int i, j, k;
  interior(t, d.nyc, d.nzc, i, j, k);
  const SPIC_GLOBAL double *n = gmem(rhon);
  gmem(rhoc)[CIDX(i, j, k)] =
      .125 * (n[NIDX(i, j, k)] + n[NIDX(i + 1, j, k)] + n[NIDX(i, j + 1, k)] +
              n[NIDX(i, j, k + 1)] + n[NIDX(i + 1, j + 1, k)] + n[NIDX(i + 1, j, k + 1)] +
              n[NIDX(i, j + 1, k + 1)] + n[NIDX(i + 1, j + 1, k + 1)]);
}

struct Tensor {
  const double *pxx, *pxy, *pxz, *pyy, *pyz, *pzz;
};
struct TensorSet {
  Tensor p[kMaxSpecies];
  double *out[3 * kMaxSpecies];  // species s: out[3s .. 3s+2]
};

// .25 * (a - b) * inv, one term of the reference's divergence stencils.
__device__ __forceinline__ double q4(double a, double b, double inv) {
  return .25 * (a - b) * inv;
}

// divSymmTensorN2C (src/Grid.cpp:322-398) on the interior cells, then
// scale by alfa = -dt/2 (src/MaxwellSolver.cpp:518-520); blockIdx.y is the
// species.
__global__ void k_divsymm(Dims d, TensorSet ts, double invdx, double invdy, double invdz,
                          double alfa) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_cells) return;
  int i, j, k;
  interior(t, d.nyc, d.nzc, i, j, k);
  const Tensor &P = ts.p[blockIdx.y];
  const SPIC_GLOBAL double *xx = gmem(P.pxx), *xy = gmem(P.pxy), *xz = gmem(P.pxz),
                           *yy = gmem(P.pyy), *yz = gmem(P.pyz), *zz = gmem(P.pzz);
  // d/dx of a: nodes (i+1, ., .) against (i, ., .) over the four (j, k)
  // corners in the order (j,k), (j,k+1), (j+1,k), (j+1,k+1)
#define DX(a)                                                                       \
  (q4(a[NIDX(i + 1, j, k)], a[NIDX(i, j, k)], invdx) +                              \
   q4(a[NIDX(i + 1, j, k + 1)], a[NIDX(i, j, k + 1)], invdx) +                      \
   q4(a[NIDX(i + 1, j + 1, k)], a[NIDX(i, j + 1, k)], invdx) +                      \
   q4(a[NIDX(i + 1, j + 1, k + 1)], a[NIDX(i, j + 1, k + 1)], invdx))
#define DY(a)                                                                       \
  (q4(a[NIDX(i, j + 1, k)], a[NIDX(i, j, k)], invdy) +                              \
   q4(a[NIDX(i, j + 1, k + 1)], a[NIDX(i, j, k + 1)], invdy) +                      \
   q4(a[NIDX(i + 1, j + 1, k)], a[NIDX(i + 1, j, k)], invdy) +                      \
   q4(a[NIDX(i + 1, j + 1, k + 1)], a[NIDX(i + 1, j, k + 1)], invdy))
#define DZ(a)                                                                       \
  (q4(a[NIDX(i, j, k + 1)], a[NIDX(i, j, k)], invdz) +                              \
   q4(a[NIDX(i + 1, j, k + 1)], a[NIDX(i + 1, j, k)], invdz) +                      \
   q4(a[NIDX(i, j + 1, k + 1)], a[NIDX(i, j + 1, k)], invdz) +                      \
   q4(a[NIDX(i + 1, j + 1, k + 1)], a[NIDX(i + 1, j + 1, k)], invdz))
  const double c1X = DX(xx), c2X = DX(xy), c3X = DX(xz);
  const double c1Y = DY(xy), c2Y = DY(yy), c3Y = DY(yz);
  const double c1Z = DZ(xz), c2Z = DZ(yz), c3Z = DZ(zz);
  const int64_t c = CIDX(i, j, k);
  const int s3 = 3 * blockIdx.y;
  gmem(ts.out[s3])[c] = (c1X + c1Y + c1Z) * alfa;
  gmem(ts.out[s3 + 1])[c] = (c2X + c2Y + c2Z) * alfa;
  gmem(ts.out[s3 + 2])[c] = (c3X + c3Y + c3Z) * alfa;
}

struct PidotArgs {
  const double *tc[3 * kMaxSpecies];  // scaled tensor divergence, cells
  const double *J[3 * kMaxSpecies];   // species currents Jx Jy Jz, nodes
  double beta[kMaxSpecies];           // .5 * qom * dt / c
  const double *Bx, *By, *Bz;
  double *Jh[3];
  int ns;
};

// Per interior node, species in order: interpC2Ninterp of the three
// divergences (src/Grid.cpp:122-141), plus the species current, then the
// PIdot update of the hat current (src/MaxwellSolver.cpp:531-566).  The hat
// current starts at zero, as the reference zeroes it.
__global__ void k_pidot(Dims d, PidotArgs a) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_nodes) return;
  int i, j, k;
  interior(t, d.nyn, d.nzn, i, j, k);
  const int64_t n = NIDX(i, j, k);
  const double bx = gmem(a.Bx)[n], by = gmem(a.By)[n], bz = gmem(a.Bz)[n];
  double hx = 0.0, hy = 0.0, hz = 0.0;
  for (int s = 0; s < a.ns; ++s) {
    double tn[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const SPIC_GLOBAL double *c = gmem(a.tc[3 * s + m]);
      tn[m] = .125 * (c[CIDX(i, j, k)] + c[CIDX(i - 1, j, k)] + c[CIDX(i, j - 1, k)] +
                      c[CIDX(i, j, k - 1)] + c[CIDX(i - 1, j - 1, k)] +
                      c[CIDX(i - 1, j, k - 1)] + c[CIDX(i, j - 1, k - 1)] +
                      c[CIDX(i - 1, j - 1, k - 1)]);
      tn[m] += gmem(a.J[3 * s + m])[n];
    }
    const double beta = a.beta[s];
    const double omcx = beta * bx, omcy = beta * by, omcz = beta * bz;
    const double edotb = tn[0] * omcx + tn[1] * omcy + tn[2] * omcz;
    const double denom = 1 / (1.0 + omcx * omcx + omcy * omcy + omcz * omcz);
    hx += (tn[0] + (tn[1] * omcz - tn[2] * omcy + edotb * omcx)) * denom;
    hy += (tn[1] + (tn[2] * omcx - tn[0] * omcz + edotb * omcy)) * denom;
    hz += (tn[2] + (tn[0] * omcy - tn[1] * omcx + edotb * omcz)) * denom;
  }
  gmem(a.Jh[0])[n] = hx;
  gmem(a.Jh[1])[n] = hy;
  gmem(a.Jh[2])[n] = hz;
}

// divN2C (src/Grid.cpp:267-319) of (X, Y, Z) at interior cell (i, j, k).
__device__ __forceinline__ double div_n2c(const Dims &d, const SPIC_GLOBAL double *X,
                                          const SPIC_GLOBAL double *Y,
                                          const SPIC_GLOBAL double *Z, int i, int j, int k,
                                          double invdx, double invdy, double invdz) {
  const double cx = q4(X[NIDX(i + 1, j, k)], X[NIDX(i, j, k)], invdx) +
                    q4(X[NIDX(i + 1, j, k + 1)], X[NIDX(i, j, k + 1)], invdx) +
                    q4(X[NIDX(i + 1, j + 1, k)], X[NIDX(i, j + 1, k)], invdx) +
                    q4(X[NIDX(i + 1, j + 1, k + 1)], X[NIDX(i, j + 1, k + 1)], invdx);
  const double cy = q4(Y[NIDX(i, j + 1, k)], Y[NIDX(i, j, k)], invdy) +
                    q4(Y[NIDX(i, j + 1, k + 1)], Y[NIDX(i, j, k + 1)], invdy) +
                    q4(Y[NIDX(i + 1, j + 1, k)], Y[NIDX(i + 1, j, k)], invdy) +
                    q4(Y[NIDX(i + 1, j + 1, k + 1)], Y[NIDX(i + 1, j, k + 1)], invdy);
  const double cz = q4(Z[NIDX(i, j, k + 1)], Z[NIDX(i, j, k)], invdz) +
                    q4(Z[NIDX(i + 1, j, k + 1)], Z[NIDX(i + 1, j, k)], invdz) +
                    q4(Z[NIDX(i, j + 1, k + 1)], Z[NIDX(i, j + 1, k)], invdz) +
                    q4(Z[NIDX(i + 1, j + 1, k + 1)], Z[NIDX(i + 1, j + 1, k)], invdz);
  return cx + cy + cz;
}

// rhoh = rhoc - dt * th * divN2C(Jh) on the interior cells
// (src/MaxwellSolver.cpp:586-596).
__global__ void k_rhoh(Dims d, const double *Jx, const double *Jy, const double *Jz,
                       const double *rhoc, double *rhoh, double invdx, double invdy,
                       double invdz, double dt, double th) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_cells) return;
  int i, j, k;
  interior(t, d.nyc, d.nzc, i, j, k);
  const double div = div_n2c(d, gmem(Jx), gmem(Jy), gmem(Jz), i, j, k, invdx, invdy, invdz);
  const int64_t c = CIDX(i, j, k);
  gmem(rhoh)[c] = gmem(rhoc)[c] - dt * th * div;
}

// solver2phys (src/Basic.cpp:365-375): the Krylov vector's interior
// nodes, components interleaved, into three node arrays.
__global__ void k_unpack(Dims d, const double *vec, double *vx, double *vy, double *vz) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_nodes) return;
  int i, j, k;
  interior(t, d.nyn, d.nzn, i, j, k);
  const int64_t n = NIDX(i, j, k);
  gmem(vx)[n] = gmem(vec)[3 * t];
  gmem(vy)[n] = gmem(vec)[3 * t + 1];
  gmem(vz)[n] = gmem(vec)[3 * t + 2];
}

struct ImageArgs {
  const double *v[3];                 // E(n + theta), nodes, ghosts applied
  const double *Bx, *By, *Bz;
  const double *rhon[kMaxSpecies];
  double beta[kMaxSpecies];           // .5 * qom * dt / c
  double coef[kMaxSpecies];           // fourpi / 2 * delt * dt / c * qom
  double *D[3];
  int ns;
};

// D = sum over species of mu . E (src/MaxwellSolver.cpp:169-198) at the
// interior nodes, species in order from zero.
__global__ void k_image_d(Dims d, ImageArgs a) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_nodes) return;
  int i, j, k;
  interior(t, d.nyn, d.nzn, i, j, k);
  const int64_t n = NIDX(i, j, k);
  const double vx = gmem(a.v[0])[n], vy = gmem(a.v[1])[n], vz = gmem(a.v[2])[n];
  const double bx = gmem(a.Bx)[n], by = gmem(a.By)[n], bz = gmem(a.Bz)[n];
  double dx = 0.0, dy = 0.0, dz = 0.0;
  for (int s = 0; s < a.ns; ++s) {
    const double beta = a.beta[s];
    const double omcx = beta * bx, omcy = beta * by, omcz = beta * bz;
    const double edotb = vx * omcx + vy * omcy + vz * omcz;
    const double denom =
        a.coef[s] * gmem(a.rhon[s])[n] / (1.0 + omcx * omcx + omcy * omcy + omcz * omcz);
    dx += (vx + (vy * omcz - vz * omcy + edotb * omcx)) * denom;
    dy += (vy + (vz * omcx - vx * omcz + edotb * omcy)) * denom;
    dz += (vz + (vx * omcy - vy * omcx + edotb * omcz)) * denom;
  }
  gmem(a.D[0])[n] = dx;
  gmem(a.D[1])[n] = dy;
  gmem(a.D[2])[n] = dz;
}

// divN2C of D on the interior cells (its stencil reads interior nodes
// only, so D's ghost update, which the reference applies first, does not
// change it).
__global__ void k_div_d(Dims d, const double *Dx, const double *Dy, const double *Dz,
                        double *divc, double invdx, double invdy, double invdz) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_cells) return;
  int i, j, k;
  interior(t, d.nyc, d.nzc, i, j, k);
  gmem(divc)[CIDX(i, j, k)] =
      div_n2c(d, gmem(Dx), gmem(Dy), gmem(Dz), i, j, k, invdx, invdy, invdz);
}

// The image at an interior node, packed as phys2solver does
// (src/Basic.cpp:385-392): ((-lapN2N_V(v) - gradC2N(divC)) * delt^2 + D) + v
// (src/MaxwellSolver.cpp:168-231, src/Grid.cpp:187-225 and 668-717).
__global__ void k_image(Dims d, ImageArgs a, const double *divc, double *im, double invdx,
                        double invdy, double invdz, double dd) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= d.in_nodes) return;
  int i, j, k;
  interior(t, d.nyn, d.nzn, i, j, k);
  const int64_t n = NIDX(i, j, k);
  const SPIC_GLOBAL double *S = gmem(divc);
#define SC(i, j, k) S[CIDX(i, j, k)]
  double grad[3];
  grad[0] = q4(SC(i, j, k), SC(i - 1, j, k), invdx) +
            q4(SC(i, j, k - 1), SC(i - 1, j, k - 1), invdx) +
            q4(SC(i, j - 1, k), SC(i - 1, j - 1, k), invdx) +
            q4(SC(i, j - 1, k - 1), SC(i - 1, j - 1, k - 1), invdx);
  grad[1] = q4(SC(i, j, k), SC(i, j - 1, k), invdy) +
            q4(SC(i, j, k - 1), SC(i, j - 1, k - 1), invdy) +
            q4(SC(i - 1, j, k), SC(i - 1, j - 1, k), invdy) +
            q4(SC(i - 1, j, k - 1), SC(i - 1, j - 1, k - 1), invdy);
  grad[2] = q4(SC(i, j, k), SC(i, j, k - 1), invdz) +
            q4(SC(i - 1, j, k), SC(i - 1, j, k - 1), invdz) +
            q4(SC(i, j - 1, k), SC(i, j - 1, k - 1), invdz) +
            q4(SC(i - 1, j - 1, k), SC(i - 1, j - 1, k - 1), invdz);
#undef SC
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const SPIC_GLOBAL double *s = gmem(a.v[m]);
    const double c = s[n];
    const double lap = (s[NIDX(i - 1, j, k)] - 2 * c + s[NIDX(i + 1, j, k)]) * invdx * invdx +
                       (s[NIDX(i, j - 1, k)] - 2 * c + s[NIDX(i, j + 1, k)]) * invdy * invdy +
                       (s[NIDX(i, j, k - 1)] - 2 * c + s[NIDX(i, j, k + 1)]) * invdz * invdz;
    double img = -lap;
    img -= grad[m];
    img *= dd;
    img += gmem(a.D[m])[n];
    img += c;
    gmem(im)[3 * t + m] = img;
  }
}

#undef NIDX
#undef CIDX

inline unsigned grid1(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_bc(double *const *arrs, int narr, const Dims &d, bool nodes,
                     bool zero_faces, const spic_solve_params *p, hipStream_t s) {
  BcArrays b;
  for (int i = 0; i < narr; ++i) b.a[i] = arrs[i];
  const int n0 = nodes ? d.nxn : d.nxc, n1 = nodes ? d.nyn : d.nyc,
            n2 = nodes ? d.nzn : d.nzc;
  const int px = p->PERIODICX ? 1 : 0, py = p->PERIODICY ? 1 : 0, pz = p->PERIODICZ ? 1 : 0;
  // enough threads for the largest face pair; grid-stride loops do the rest
  const int64_t faces = 2 * ((int64_t)n0 * n1 + (int64_t)n0 * n2 + (int64_t)n1 * n2);
  unsigned blocks = (unsigned)((faces + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  if (px || py || pz)
    hipLaunchKernelGGL(k_bc_periodic, dim3(blocks, narr), dim3(256), 0, s, b, n0, n1, n2,
                       d.nyn, nodes ? 1 : 0, px, py, pz);
  if (!px || !py || !pz)
    hipLaunchKernelGGL(k_bc_faces, dim3(blocks, narr), dim3(256), 0, s, b, n0, n1, n2,
                       zero_faces ? 1 : 0, px, py, pz);
  return hipGetLastError();
}

int check_solve(const spic_grid *grd, const spic_solve_params *p, int ns) {
  int rc = check_grid(grd);
  if (rc) return rc;
  if (grd->prec != SPIC_F64)
    return fail(SPIC_EINVAL, "the field-solve operators run in FP64 (grid prec must be F64)");
  if (!p) return fail(SPIC_EINVAL, "solve parameters are NULL");
  if (ns < 0 || ns > kMaxSpecies)
    return fail(SPIC_EINVAL, "species count must be 0..%d", kMaxSpecies);
  if (grd->nxc < 3 || grd->nyc < 3 || grd->nzc < 3)
    return fail(SPIC_EINVAL, "grid needs at least one physical cell per axis");
  return SPIC_OK;
}

int finish(hipStream_t s, hipError_t e, const char *what) {
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess && !async_mode()) e = hipStreamSynchronize(s);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, what);
}

}  // namespace
}  // namespace spic

using spic::Dims;

extern "C" {

size_t spic_field_workspace_bytes(const spic_grid *grd, int ns) {
  if (!grd || ns < 0) return 0;
  const Dims d = spic::dims_of(grd);
  const int64_t hat = 3 * (int64_t)ns * d.nc;
  const int64_t image = 6 * d.nn + d.nc;
  return sizeof(double) * (size_t)(hat > image ? hat : image);
}

int spic_bc_scalar(const spic_grid *grd, double *a, int kind, const spic_solve_params *p,
                   void *stream) {
  int rc = spic::check_solve(grd, p, 0);
  if (rc) return rc;
  if (!a) return spic::fail(SPIC_EINVAL, "array is NULL");
  if (kind < 0 || kind > 3) return spic::fail(SPIC_EINVAL, "BC kind must be 0..3");
  hipStream_t s = (hipStream_t)stream;
  const Dims d = spic::dims_of(grd);
  double *arr[1] = {a};
  return spic::finish(s, spic::launch_bc(arr, 1, d, kind <= 1, kind == 1, p, s), "BC");
}

int spic_rhoc_from_rhon(const spic_grid *grd, double *rhon, double *rhoc,
                        const spic_solve_params *p, void *stream) {
  int rc = spic::check_solve(grd, p, 0);
  if (rc) return rc;
  if (!rhon || !rhoc) return spic::fail(SPIC_EINVAL, "rhon or rhoc is NULL");
  hipStream_t s = (hipStream_t)stream;
  const Dims d = spic::dims_of(grd);
  double *arr[1] = {rhon};
  hipError_t e = spic::launch_bc(arr, 1, d, true, false, p, s);
  if (e == hipSuccess && d.in_cells > 0)
    hipLaunchKernelGGL(spic::k_n2c, dim3(spic::grid1(d.in_cells)), dim3(256), 0, s, d,
                       rhon, rhoc);
  return spic::finish(s, e, "rhoc from rhon");
}

int spic_hat_densities(const spic_grid *grd, const spic_field *field,
                       const spic_moments *ids, int ns, const double *rhoc,
                       const spic_solve_params *p, const spic_hat *out, void *work,
                       size_t work_bytes, void *stream) {
  int rc = spic::check_solve(grd, p, ns);
  if (rc) return rc;
  if ((rc = spic::check_field(field))) return rc;
  if (!ids || !rhoc || !out || !out->rhoh || !out->Jxh || !out->Jyh || !out->Jzh)
    return spic::fail(SPIC_EINVAL, "moments, rhoc or hat arrays are NULL");
  if (ns < 1) return spic::fail(SPIC_EINVAL, "species count must be 1..%d", spic::kMaxSpecies);
  for (int is = 0; is < ns; ++is)
    if ((rc = spic::check_moments(&ids[is]))) return rc;
  if (!work || work_bytes < spic_field_workspace_bytes(grd, ns))
    return spic::fail(SPIC_EINVAL, "workspace of %zu bytes needed",
                      spic_field_workspace_bytes(grd, ns));
  hipStream_t s = (hipStream_t)stream;
  const Dims d = spic::dims_of(grd);
  double *tc = (double *)work;  // 3 ns cell arrays
  spic::TensorSet ts;
  spic::PidotArgs pa;
  double *tcs[3 * spic::kMaxSpecies];
  for (int is = 0; is < ns; ++is) {
    ts.p[is] = {ids[is].pxx, ids[is].pxy, ids[is].pxz, ids[is].pyy, ids[is].pyz,
                ids[is].pzz};
    for (int m = 0; m < 3; ++m) {
      tcs[3 * is + m] = tc + (3 * is + m) * d.nc;
      ts.out[3 * is + m] = tcs[3 * is + m];
      pa.tc[3 * is + m] = tcs[3 * is + m];
    }
    pa.J[3 * is] = ids[is].Jx;
    pa.J[3 * is + 1] = ids[is].Jy;
    pa.J[3 * is + 2] = ids[is].Jz;
    pa.beta[is] = .5 * p->qom[is] * p->dt / p->c;
  }
  pa.Bx = (const double *)field->Bxn;
  pa.By = (const double *)field->Byn;
  pa.Bz = (const double *)field->Bzn;
  pa.Jh[0] = out->Jxh;
  pa.Jh[1] = out->Jyh;
  pa.Jh[2] = out->Jzh;
  pa.ns = ns;
  const double alfa = -p->dt / 2.0;
  hipError_t e = hipSuccess;
  // hat current: every species' scaled tensor divergence, their ghost
  // cells, then the per-node sum over species
  hipLaunchKernelGGL(spic::k_divsymm, dim3(spic::grid1(d.in_cells), ns), dim3(256), 0, s, d,
                     ts, grd->invdx, grd->invdy, grd->invdz, alfa);
  if ((e = hipGetLastError()) == hipSuccess)
    e = spic::launch_bc(tcs, 3 * ns, d, false, false, p, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(spic::k_pidot, dim3(spic::grid1(d.in_nodes)), dim3(256), 0, s, d, pa);
    e = hipGetLastError();
  }
  double *jh[3] = {out->Jxh, out->Jyh, out->Jzh};
  if (e == hipSuccess) e = spic::launch_bc(jh, 3, d, true, false, p, s);
  // hat density
  if (e == hipSuccess) {
    hipLaunchKernelGGL(spic::k_rhoh, dim3(spic::grid1(d.in_cells)), dim3(256), 0, s, d,
                       out->Jxh, out->Jyh, out->Jzh, rhoc, out->rhoh, grd->invdx, grd->invdy,
                       grd->invdz, p->dt, p->th);
    e = hipGetLastError();
  }
  double *rh[1] = {out->rhoh};
  if (e == hipSuccess) e = spic::launch_bc(rh, 1, d, false, false, p, s);
  return spic::finish(s, e, "hat densities");
}

int spic_maxwell_image(const spic_grid *grd, const spic_field *field,
                       const spic_moments *ids, int ns, const spic_solve_params *p,
                       const double *vec, double *im, void *work, size_t work_bytes,
                       void *stream) {
  int rc = spic::check_solve(grd, p, ns);
  if (rc) return rc;
  if ((rc = spic::check_field(field))) return rc;
  if (!vec || !im || (ns > 0 && !ids))
    return spic::fail(SPIC_EINVAL, "vector, image or moments are NULL");
  for (int is = 0; is < ns; ++is)
    if (!ids[is].rhon) return spic::fail(SPIC_EINVAL, "species %d rhon is NULL", is);
  if (!work || work_bytes < spic_field_workspace_bytes(grd, ns))
    return spic::fail(SPIC_EINVAL, "workspace of %zu bytes needed",
                      spic_field_workspace_bytes(grd, ns));
  hipStream_t s = (hipStream_t)stream;
  const Dims d = spic::dims_of(grd);
  double *w = (double *)work;
  double *v[3] = {w, w + d.nn, w + 2 * d.nn};
  double *D[3] = {w + 3 * d.nn, w + 4 * d.nn, w + 5 * d.nn};
  double *divc = w + 6 * d.nn;
  const double delt = p->c * p->th * p->dt;
  spic::ImageArgs a;
  for (int m = 0; m < 3; ++m) {
    a.v[m] = v[m];
    a.D[m] = D[m];
  }
  a.Bx = (const double *)field->Bxn;
  a.By = (const double *)field->Byn;
  a.Bz = (const double *)field->Bzn;
  for (int is = 0; is < ns; ++is) {
    a.rhon[is] = ids[is].rhon;
    a.beta[is] = .5 * p->qom[is] * p->dt / p->c;
    a.coef[is] = p->fourpi / 2 * delt * p->dt / p->c * p->qom[is];
  }
  a.ns = ns;
  hipError_t e = hipSuccess;
  if (d.in_nodes > 0) {
    hipLaunchKernelGGL(spic::k_unpack, dim3(spic::grid1(d.in_nodes)), dim3(256), 0, s, d, vec,
                       v[0], v[1], v[2]);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = spic::launch_bc(v, 3, d, true, true, p, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(spic::k_image_d, dim3(spic::grid1(d.in_nodes)), dim3(256), 0, s, d, a);
    hipLaunchKernelGGL(spic::k_div_d, dim3(spic::grid1(d.in_cells)), dim3(256), 0, s, d, D[0],
                       D[1], D[2], divc, grd->invdx, grd->invdy, grd->invdz);
    e = hipGetLastError();
  }
  double *dc[1] = {divc};
  if (e == hipSuccess) e = spic::launch_bc(dc, 1, d, false, false, p, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(spic::k_image, dim3(spic::grid1(d.in_nodes)), dim3(256), 0, s, d, a,
                       divc, im, grd->invdx, grd->invdy, grd->invdz, delt * delt);
    e = hipGetLastError();
  }
  return spic::finish(s, e, "MaxwellImage");
}

}  // extern "C"
