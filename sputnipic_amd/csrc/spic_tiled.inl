// spic_tiled.inl -- the tile-resident gfx950 kernel for a whole particle
// phase (every species' mover_PC followed by its interpP2G) over
// tile-sorted particles.
//
// One workgroup per tile of TX x TY x TZ cells.  The workgroup
//   1. stages the six node fields of its tile plus a one-cell halo into LDS
//      (node-major, six components per node),
//   2. for each species: zeroes an LDS block of the ten moments over the
//      same nodes, pushes the tile's particles (SoA loads/stores, one lane
//      per particle, predictor-corrector iterates gathered from LDS with the
//      eight-node stencil kept in registers while the cell is unchanged),
//      deposits each particle's 80 contributions with LDS atomics, and
//      finally adds every nonzero LDS node moment to the FP64 global arrays
//      with one hardware atomic per node and moment.
// Particles are ordered within a tile so that consecutive lanes sit in
// different cells (round-robin over cells), which keeps the LDS atomics of a
// wave on distinct addresses.  A particle whose iterate or final cell falls
// outside the staged block (it drifted more than one cell out of its tile
// since the last sort) takes the global-memory path of the plain kernels for
// that step, so results never depend on sort staleness.
//
// The arithmetic per particle is exactly that of spic_plain.inl (same
// operation order); with SPIC_STRICT it matches the reference bit for bit.
#include "spic_internal.hpp"
#include "spic_tiled.hpp"

#ifndef SPIC_NS
#error "define SPIC_NS and SPIC_STRICT before including spic_tiled.inl"
#endif

namespace spic {
namespace SPIC_NS {

constexpr bool kStrictT = (SPIC_STRICT != 0);

template <typename T>
__device__ __forceinline__ int tcell_trunc(T p, double start, T inv, int n) {
  if (kStrictT || sizeof(T) == 8) {
    double c = ((double)p - start) * (double)inv;
    c = fmin(fmax(c, -8.0), (double)n + 8.0);
    return 2 + (int)c;
  } else {
    float c = ((float)p - (float)start) * (float)inv;
    c = fminf(fmaxf(c, -8.0f), (float)n + 8.0f);
    return 2 + (int)c;
  }
}

template <typename T>
__device__ __forceinline__ int tcell_floor(T p, double start, T inv, int n) {
  if (kStrictT || sizeof(T) == 8) {
    double c = ((double)p - start) * (double)inv;
    c = fmin(fmax(c, -8.0), (double)n + 8.0);
    return 2 + (int)floor(c);
  } else {
    float c = ((float)p - (float)start) * (float)inv;
    c = fminf(fmaxf(c, -8.0f), (float)n + 8.0f);
    return 2 + (int)floorf(c);
  }
}

__device__ __forceinline__ bool tclamp(int &ix, int n) {
  bool bad = (ix < 1) | (ix > n - 1);
  ix = min(max(ix, 1), n - 1);
  return bad;
}

// Denominator of the implicit velocity solve, 1/(1 + omdtsq): in double
// (as src/Mover.cpp:71) for FP64 and strict FP32; one float division in
// the production FP32 form.
template <typename T>
__device__ __forceinline__ T implicit_denom(T omdtsq) {
  if (kStrictT || sizeof(T) == 8) return (T)(1.0 / (1.0 + (double)omdtsq));
  return T(1) / (T(1) + omdtsq);
}

template <typename T, int TX, int TY, int TZ>
struct TileGeom {
  static constexpr int BX = TX + 3, BY = TY + 3, BZ = TZ + 3;
  static constexpr int NB = BX * BY * BZ;  // staged nodes
};

// A: type of the LDS moment partials (T, or double for FP32 particles).
template <typename T, typename A, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) void k_tiled_phase(TiledArgs<T> a) {
  using G = TileGeom<T, TX, TY, TZ>;
  constexpr int BX = G::BX, BY = G::BY, BZ = G::BZ, NB = G::NB;
  __shared__ T s_fld[NB * 6];
  __shared__ A s_mom[10 * NB];
  __shared__ T s_xn[BX], s_yn[BY], s_zn[BZ];

  const int tile = blockIdx.x;
  const int tz = tile % a.ntz;
  const int ty = (tile / a.ntz) % a.nty;
  const int tx = tile / (a.ntz * a.nty);
  // global node index of block node 0 along each axis; block nodes cover
  // global nodes [c0, c0 + T + 2], i.e. the tile's cells and one halo cell
  // on each side (cell c of the tile uses nodes c+1, c+2).
  const int c0x = tx * TX, c0y = ty * TY, c0z = tz * TZ;
  const int nyn = a.g.nyn, nzn = a.g.nzn;
  const int sxg = nyn * nzn;

  // 1. stage fields and the block's node coordinates
  for (int b = threadIdx.x; b < BX + BY + BZ; b += NT) {
    if (b < BX) s_xn[b] = (c0x + b < a.g.nxn) ? a.g.xn[c0x + b] : T(0);
    else if (b < BX + BY) s_yn[b - BX] = (c0y + b - BX < nyn) ? a.g.yn[c0y + b - BX] : T(0);
    else s_zn[b - BX - BY] = (c0z + b - BX - BY < nzn) ? a.g.zn[c0z + b - BX - BY] : T(0);
  }
  for (int b = threadIdx.x; b < NB; b += NT) {
    const int bz = b % BZ, by = (b / BZ) % BY, bx = b / (BZ * BY);
    const int gx = c0x + bx, gy = c0y + by, gz = c0z + bz;
    T v[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (gx < a.g.nxn && gy < nyn && gz < nzn) {
      const int n = (gx * nyn + gy) * nzn + gz;
      v[0] = a.f.Ex[n]; v[1] = a.f.Ey[n]; v[2] = a.f.Ez[n];
      v[3] = a.f.Bx[n]; v[4] = a.f.By[n]; v[5] = a.f.Bz[n];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) s_fld[b * 6 + c] = v[c];
  }

  bool bad_move = false, bad_dep = false;
  unsigned int n_gg = 0, n_gd = 0, n_lg = 0, n_ld = 0;
  for (int s = 0; s < a.ns; ++s) {
    const TiledSpecies<T> &sp = a.sp[s];
    const int64_t p0 = sp.tile_start[tile], p1 = sp.tile_start[tile + 1];
    if (p0 == p1) continue;  // uniform across the block
    for (int i = threadIdx.x; i < 10 * NB; i += NT) s_mom[i] = A(0);
    __syncthreads();

    // The next particle's seven values are loaded while the current one is
    // pushed, so HBM latency overlaps the arithmetic.
    int64_t i = p0 + threadIdx.x;
    T nx = T(0), ny = T(0), nz = T(0), nu = T(0), nv = T(0), nw = T(0), nq = T(0);
    if (i < p1) {
      nx = sp.x[i]; ny = sp.y[i]; nz = sp.z[i];
      nu = sp.u[i]; nv = sp.v[i]; nw = sp.w[i]; nq = sp.q[i];
    }
    for (; i < p1; i += NT) {
      T x = nx, y = ny, z = nz, u = nu, v = nv, w = nw;
      const T q = nq;
      const int64_t inext = i + NT;
      if (inext < p1) {
        nx = sp.x[inext]; ny = sp.y[inext]; nz = sp.z[inext];
        nu = sp.u[inext]; nv = sp.v[inext]; nw = sp.w[inext]; nq = sp.q[inext];
      }
      // ---- mover (src/Mover.cpp:33-157) ----
      const T xt = x, yt = y, zt = z;
      T upt = T(0), vpt = T(0), wpt = T(0);
      T xc = x, yc = y, zc = z;
      int cix = -100, ciy = -100, ciz = -100;  // cached stencil cell
      T F[2][2][2][6];
      T X0 = T(0), X1 = T(0), Y0 = T(0), Y1 = T(0), Z0 = T(0), Z1 = T(0);
      for (int it = 0; it < sp.k.niter; ++it) {
        int ix = tcell_trunc<T>(xc, a.g.xStart, a.g.invdx, a.g.nxn);
        int iy = tcell_trunc<T>(yc, a.g.yStart, a.g.invdy, a.g.nyn);
        int iz = tcell_trunc<T>(zc, a.g.zStart, a.g.invdz, a.g.nzn);
        bad_move |= tclamp(ix, a.g.nxn);
        bad_move |= tclamp(iy, a.g.nyn);
        bad_move |= tclamp(iz, a.g.nzn);
        if (ix != cix || iy != ciy || iz != ciz) {
          cix = ix; ciy = iy; ciz = iz;
          const int bx = ix - 1 - c0x, by = iy - 1 - c0y, bz = iz - 1 - c0z;
          const bool in_block = (bx >= 0) & (bx <= BX - 2) & (by >= 0) &
                                (by <= BY - 2) & (bz >= 0) & (bz <= BZ - 2);
          if (in_block) {
            ++n_lg;
            X0 = s_xn[bx]; X1 = s_xn[bx + 1];
            Y0 = s_yn[by]; Y1 = s_yn[by + 1];
            Z0 = s_zn[bz]; Z1 = s_zn[bz + 1];
#pragma unroll
            for (int aa = 0; aa < 2; ++aa)
#pragma unroll
              for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                  const int nb = ((bx + 1 - aa) * BY + (by + 1 - bb)) * BZ +
                                 (bz + 1 - cc);
#pragma unroll
                  for (int c = 0; c < 6; ++c) F[aa][bb][cc][c] = s_fld[nb * 6 + c];
                }
          } else {
            ++n_gg;
            X0 = a.g.xn[ix - 1]; X1 = a.g.xn[ix];
            Y0 = a.g.yn[iy - 1]; Y1 = a.g.yn[iy];
            Z0 = a.g.zn[iz - 1]; Z1 = a.g.zn[iz];
            const int base = (ix * nyn + iy) * nzn + iz;
#pragma unroll
            for (int aa = 0; aa < 2; ++aa)
#pragma unroll
              for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                  const int n = base - aa * sxg - bb * nzn - cc;
                  F[aa][bb][cc][0] = a.f.Ex[n];
                  F[aa][bb][cc][1] = a.f.Ey[n];
                  F[aa][bb][cc][2] = a.f.Ez[n];
                  F[aa][bb][cc][3] = a.f.Bx[n];
                  F[aa][bb][cc][4] = a.f.By[n];
                  F[aa][bb][cc][5] = a.f.Bz[n];
                }
          }
        }
        const T xi[2] = {xc - X0, X1 - xc};
        const T eta[2] = {yc - Y0, Y1 - yc};
        const T zeta[2] = {zc - Z0, Z1 - zc};
        T Exl = T(0), Eyl = T(0), Ezl = T(0), Bxl = T(0), Byl = T(0), Bzl = T(0);
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const T wgt = xi[aa] * eta[bb] * zeta[cc] * a.g.invVOL;
              Exl += wgt * F[aa][bb][cc][0];
              Eyl += wgt * F[aa][bb][cc][1];
              Ezl += wgt * F[aa][bb][cc][2];
              Bxl += wgt * F[aa][bb][cc][3];
              Byl += wgt * F[aa][bb][cc][4];
              Bzl += wgt * F[aa][bb][cc][5];
            }
        const T q2 = sp.k.qomdt2;
        const T omdtsq = q2 * q2 * (Bxl * Bxl + Byl * Byl + Bzl * Bzl);
        const T denom = implicit_denom<T>(omdtsq);
        const T ut = u + q2 * Exl;
        const T vt = v + q2 * Eyl;
        const T wt = w + q2 * Ezl;
        const T udotb = ut * Bxl + vt * Byl + wt * Bzl;
        upt = (ut + q2 * (vt * Bzl - wt * Byl + q2 * udotb * Bxl)) * denom;
        vpt = (vt + q2 * (wt * Bxl - ut * Bzl + q2 * udotb * Byl)) * denom;
        wpt = (wt + q2 * (ut * Byl - vt * Bxl + q2 * udotb * Bzl)) * denom;
        xc = xt + upt * sp.k.dto2;
        yc = yt + vpt * sp.k.dto2;
        zc = zt + wpt * sp.k.dto2;
      }
      if (kStrictT || sizeof(T) == 8) {
        u = (T)(2.0 * (double)upt - (double)u);
        v = (T)(2.0 * (double)vpt - (double)v);
        w = (T)(2.0 * (double)wpt - (double)w);
      } else {
        u = T(2) * upt - u;
        v = T(2) * vpt - v;
        w = T(2) * wpt - w;
      }
      x = xt + upt * sp.k.dt_sub;
      y = yt + vpt * sp.k.dt_sub;
      z = zt + wpt * sp.k.dt_sub;
      if ((double)x > a.g.Lx) {
        if (sp.k.perx) x = (T)((double)x - a.g.Lx);
        else { u = -u; x = (T)(2 * a.g.Lx - (double)x); }
      }
      if (x < T(0)) {
        if (sp.k.perx) x = (T)((double)x + a.g.Lx);
        else { u = -u; x = -x; }
      }
      if ((double)y > a.g.Ly) {
        if (sp.k.pery) y = (T)((double)y - a.g.Ly);
        else { v = -v; y = (T)(2 * a.g.Ly - (double)y); }
      }
      if (y < T(0)) {
        if (sp.k.pery) y = (T)((double)y + a.g.Ly);
        else { v = -v; y = -y; }
      }
      if ((double)z > a.g.Lz) {
        if (sp.k.perz) z = (T)((double)z - a.g.Lz);
        else { w = -w; z = (T)(2 * a.g.Lz - (double)z); }
      }
      if (z < T(0)) {
        if (sp.k.perz) z = (T)((double)z + a.g.Lz);
        else { w = -w; z = -z; }
      }
      sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
      sp.u[i] = u; sp.v[i] = v; sp.w[i] = w;

      // ---- deposit (src/Mover.cpp:399-570) ----
      int ix = tcell_floor<T>(x, a.g.xStart, a.g.invdx, a.g.nxn);
      int iy = tcell_floor<T>(y, a.g.yStart, a.g.invdy, a.g.nyn);
      int iz = tcell_floor<T>(z, a.g.zStart, a.g.invdz, a.g.nzn);
      bad_dep |= tclamp(ix, a.g.nxn);
      bad_dep |= tclamp(iy, a.g.nyn);
      bad_dep |= tclamp(iz, a.g.nzn);
      const int bx = ix - 1 - c0x, by = iy - 1 - c0y, bz = iz - 1 - c0z;
      const bool in_block = (bx >= 0) & (bx <= BX - 2) & (by >= 0) &
                            (by <= BY - 2) & (bz >= 0) & (bz <= BZ - 2);
      T xi[2], eta[2], zeta[2];
      if (in_block) {
        xi[0] = x - s_xn[bx]; xi[1] = s_xn[bx + 1] - x;
        eta[0] = y - s_yn[by]; eta[1] = s_yn[by + 1] - y;
        zeta[0] = z - s_zn[bz]; zeta[1] = s_zn[bz + 1] - z;
      } else {
        xi[0] = x - a.g.xn[ix - 1]; xi[1] = a.g.xn[ix] - x;
        eta[0] = y - a.g.yn[iy - 1]; eta[1] = a.g.yn[iy] - y;
        zeta[0] = z - a.g.zn[iz - 1]; zeta[1] = a.g.zn[iz] - z;
      }
      const T iv = a.g.invVOL;
      const T uu = u * u, uv = u * v, uw = u * w, vv = v * v, vw = v * w,
              ww = w * w;
      if (in_block) {
        ++n_ld;
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const T wgt = q * xi[aa] * eta[bb] * zeta[cc] * iv;
              const int nb = ((bx + 1 - aa) * BY + (by + 1 - bb)) * BZ + (bz + 1 - cc);
              atomicAdd(&s_mom[0 * NB + nb], (A)(wgt * iv));
              atomicAdd(&s_mom[1 * NB + nb], (A)((u * wgt) * iv));
              atomicAdd(&s_mom[2 * NB + nb], (A)((v * wgt) * iv));
              atomicAdd(&s_mom[3 * NB + nb], (A)((w * wgt) * iv));
              atomicAdd(&s_mom[4 * NB + nb], (A)((uu * wgt) * iv));
              atomicAdd(&s_mom[5 * NB + nb], (A)((uv * wgt) * iv));
              atomicAdd(&s_mom[6 * NB + nb], (A)((uw * wgt) * iv));
              atomicAdd(&s_mom[7 * NB + nb], (A)((vv * wgt) * iv));
              atomicAdd(&s_mom[8 * NB + nb], (A)((vw * wgt) * iv));
              atomicAdd(&s_mom[9 * NB + nb], (A)((ww * wgt) * iv));
            }
      } else {
        ++n_gd;
        const int base = (ix * nyn + iy) * nzn + iz;
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const T wgt = q * xi[aa] * eta[bb] * zeta[cc] * iv;
              const int n = base - aa * sxg - bb * nzn - cc;
              unsafeAtomicAdd(sp.m.m[0] + n, (double)(wgt * iv));
              unsafeAtomicAdd(sp.m.m[1] + n, (double)((u * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[2] + n, (double)((v * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[3] + n, (double)((w * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[4] + n, (double)((uu * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[5] + n, (double)((uv * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[6] + n, (double)((uw * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[7] + n, (double)((vv * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[8] + n, (double)((vw * wgt) * iv));
              unsafeAtomicAdd(sp.m.m[9] + n, (double)((ww * wgt) * iv));
            }
      }
    }
    __syncthreads();
    // 3. store the block's partial moments (coalesced, FP64); the reduction
    // kernel adds them into the global arrays per node in a fixed order
    {
      double *dst = a.partials + ((size_t)tile * a.ns + s) * (10 * NB);
      for (int i = threadIdx.x; i < 10 * NB; i += NT) dst[i] = (double)s_mom[i];
    }
    __syncthreads();
  }
  if (bad_move) atomicAdd(a.status, 1u);
  if (bad_dep) atomicAdd(a.status + 1, 1u);
  if (a.counters) {
    if (n_gg) atomicAdd(a.counters + 0, (unsigned long long)n_gg);
    if (n_gd) atomicAdd(a.counters + 1, (unsigned long long)n_gd);
    if (n_lg) atomicAdd(a.counters + 2, (unsigned long long)n_lg);
    if (n_ld) atomicAdd(a.counters + 3, (unsigned long long)n_ld);
  }
}

// Per node and species: the sum of the partials of every tile whose staged
// block holds the node (at most two tiles per axis), in tile order, with
// Neumaier's compensated summation, added to the FP64 moment arrays.
// Tiles without particles of the species wrote no partials and are skipped.
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_tile_reduce(TiledArgsBase a, int nn) {
  constexpr int BX = TX + 3, BY = TY + 3, BZ = TZ + 3, NB = BX * BY * BZ;
  const int s = blockIdx.y;
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= nn) return;
  const int nyn = a.nyn, nzn = a.nzn;
  const int gz = n % nzn, gy = (n / nzn) % nyn, gx = n / (nzn * nyn);
  const int xl = gx >= BX - 1 ? (gx - (BX - 1) + TX - 1) / TX : 0;
  const int xh = min(gx / TX, a.ntx - 1);
  const int yl = gy >= BY - 1 ? (gy - (BY - 1) + TY - 1) / TY : 0;
  const int yh = min(gy / TY, a.nty - 1);
  const int zl = gz >= BZ - 1 ? (gz - (BZ - 1) + TZ - 1) / TZ : 0;
  const int zh = min(gz / TZ, a.ntz - 1);
  const int64_t *ts = a.tile_start[s];
  double sum[10], cmp[10];
#pragma unroll
  for (int m = 0; m < 10; ++m) { sum[m] = 0.0; cmp[m] = 0.0; }
  bool any = false;
  for (int tx = xl; tx <= xh; ++tx)
    for (int ty = yl; ty <= yh; ++ty)
      for (int tz = zl; tz <= zh; ++tz) {
        const int tile = (tx * a.nty + ty) * a.ntz + tz;
        if (ts[tile] == ts[tile + 1]) continue;
        any = true;
        const int b = ((gx - tx * TX) * BY + (gy - ty * TY)) * BZ + (gz - tz * TZ);
        const double *src = a.partials + ((size_t)tile * a.ns + s) * (10 * NB) + b;
#pragma unroll
        for (int m = 0; m < 10; ++m) {
          const double v = src[m * NB];
          const double t = sum[m] + v;
          cmp[m] += fabs(sum[m]) >= fabs(v) ? (sum[m] - t) + v : (v - t) + sum[m];
          sum[m] = t;
        }
      }
  if (!any) return;
#pragma unroll
  for (int m = 0; m < 10; ++m) a.mom[s].m[m][n] += sum[m] + cmp[m];
}

template <typename T, typename A>
hipError_t launch_tiled_shape(const TiledArgs<T> &a, int tiles, hipStream_t s) {
  if (a.tx == 4 && a.ty == 4 && a.tz == 4)
    hipLaunchKernelGGL((k_tiled_phase<T, A, 4, 4, 4, 256>), dim3(tiles),
                       dim3(256), 0, s, a);
  else if (a.tx == 8 && a.ty == 4 && a.tz == 4)
    hipLaunchKernelGGL((k_tiled_phase<T, A, 8, 4, 4, 256>), dim3(tiles),
                       dim3(256), 0, s, a);
  else if (a.tx == 2 && a.ty == 2 && a.tz == 2)
    hipLaunchKernelGGL((k_tiled_phase<T, A, 2, 2, 2, 128>), dim3(tiles),
                       dim3(128), 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tiled_phase(const TiledArgs<T> &a, int tiles, hipStream_t s) {
  if (tiles <= 0) return hipSuccess;
  if (sizeof(T) == 4 && a.lds_fp64) return launch_tiled_shape<T, double>(a, tiles, s);
  return launch_tiled_shape<T, T>(a, tiles, s);
}

template <typename T>
hipError_t launch_tile_reduce(const TiledArgs<T> &a, int nn, hipStream_t s) {
  if (nn <= 0 || a.ns <= 0) return hipSuccess;
  TiledArgsBase b;
  b.nyn = a.g.nyn; b.nzn = a.g.nzn;
  b.ntx = a.ntx; b.nty = a.nty; b.ntz = a.ntz;
  b.ns = a.ns;
  b.partials = a.partials;
  for (int i = 0; i < a.ns; ++i) {
    b.tile_start[i] = a.sp[i].tile_start;
    b.mom[i] = a.sp[i].m;
  }
  dim3 grid((nn + 255) / 256, a.ns);
  if (a.tx == 4 && a.ty == 4 && a.tz == 4)
    hipLaunchKernelGGL((k_tile_reduce<4, 4, 4>), grid, dim3(256), 0, s, b, nn);
  else if (a.tx == 8 && a.ty == 4 && a.tz == 4)
    hipLaunchKernelGGL((k_tile_reduce<8, 4, 4>), grid, dim3(256), 0, s, b, nn);
  else if (a.tx == 2 && a.ty == 2 && a.tz == 2)
    hipLaunchKernelGGL((k_tile_reduce<2, 2, 2>), grid, dim3(256), 0, s, b, nn);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template hipError_t launch_tiled_phase<double>(const TiledArgs<double> &, int,
                                               hipStream_t);
template hipError_t launch_tile_reduce<double>(const TiledArgs<double> &, int,
                                               hipStream_t);
template hipError_t launch_tile_reduce<float>(const TiledArgs<float> &, int,
                                              hipStream_t);
template hipError_t launch_tiled_phase<float>(const TiledArgs<float> &, int,
                                              hipStream_t);

}  // namespace SPIC_NS
}  // namespace spic
