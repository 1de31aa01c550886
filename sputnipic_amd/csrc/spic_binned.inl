// spic_binned.inl -- the binned tile-resident phase kernels: every species'
// mover_PC followed by its interpP2G, with each particle written back into
// the segment of its destination tile (binned storage, spic_binned.hpp).
//
// k_binned_phase: one workgroup of four threads per cell per tile of
// TX x TY x TZ cells.  It
//   1. stages the six node fields and node coordinates of the tile plus a
//      one-cell halo into LDS;
//   2. for each species, batch by batch (one particle per thread):
//      - pushes the particle (src/Mover.cpp:33-157) with the eight-node
//        stencil gathered from LDS and kept in registers while the cell is
//        unchanged (from global memory if it lies outside the staged block);
//      - writes it into the output buffer: a particle that stays in the
//        tile takes the next slot of a block of its own segment that the
//        workgroup reserved for this species with one counter atomic (slots
//        from an LDS counter, one atomic per wave); a particle that left the
//        tile (it is in the list ring, below) is written when the ring is
//        flushed, at the next slot of its destination segment's counter;
//        the overflow region takes particles whose segment is full;
//      - deposits it (src/Mover.cpp:399-570): a particle in one of the tile's
//        cells goes into that cell's LDS bucket; one in a halo cell adds its
//        80 contributions to the LDS moment block; one outside the block adds
//        them to global memory;
//      - after a barrier the four threads of each cell take the cell's
//        bucket, each thread two of the cell's eight nodes, and accumulate
//        their ten moments in registers;
//   3. adds the register sums into the LDS moment block and stores the block
//      to the tile-partial scratch, summed per node by k_binned_reduce.
// k_binned_overflow: the particles of the input overflow regions, one per
// thread, with global-memory gathers and deposits, re-binned the same way.
//
// The arithmetic per particle is exactly the reference's (the operation
// order of spic_plain.inl); with SPIC_STRICT it matches it bit for bit.
// Only the order in which moment contributions are summed differs.
#include "spic_binned.hpp"

#ifndef SPIC_NS
#error "define SPIC_NS and SPIC_STRICT before including spic_binned.inl"
#endif

namespace spic {
namespace SPIC_NS {

constexpr bool kStrictB = (SPIC_STRICT != 0);

// Mover cell index 2 + int((p - start) * inv) (src/Mover.cpp:45-47) and
// deposit cell index 2 + floor(...) (src/Mover.cpp:411-413), in double for
// FP64 and strict FP32 (the reference promotes to double via xStart),
// clamped before the conversion so that no input overflows it.
template <typename T>
__device__ __forceinline__ int bcell_trunc(T p, double start, T inv, int n) {
  if (kStrictB || sizeof(T) == 8) {
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
__device__ __forceinline__ int bcell_floor(T p, double start, T inv, int n) {
  if (kStrictB || sizeof(T) == 8) {
    double c = ((double)p - start) * (double)inv;
    c = fmin(fmax(c, -8.0), (double)n + 8.0);
    return 2 + (int)floor(c);
  } else {
    float c = ((float)p - (float)start) * (float)inv;
    c = fminf(fmaxf(c, -8.0f), (float)n + 8.0f);
    return 2 + (int)floorf(c);
  }
}

// s_waitcnt vmcnt(0) (gfx9 encoding: vmcnt 0, expcnt 7, lgkmcnt 15).  Ends
// the rare global-memory stencil gather inside its own branch.  Without it
// the stencil registers, which the common branch fills from LDS, reach the
// interpolation still counted as vector-memory loads, and the compiler
// waits there for every vector memory operation in flight: the next
// batch's prefetched particles and the previous batch's stores.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Stencil indices must address nodes 0..n-1: ix-1 >= 0 and ix <= n-1.
__device__ __forceinline__ bool bclamp(int &ix, int n) {
  bool bad = (ix < 1) | (ix > n - 1);
  ix = min(max(ix, 1), n - 1);
  return bad;
}

// 1/(1 + omdtsq): the correctly rounded double division (src/Mover.cpp:71)
// in strict arithmetic; in production FP64 the hardware reciprocal refined
// by two Newton-Raphson steps r += r (1 - d r), within an ulp of the
// quotient at half the instructions (d >= 1, so no special cases); the
// hardware reciprocal (1 ulp) in the FP32 form.
template <typename T>
__device__ __forceinline__ T bdenom(T omdtsq) {
  if (kStrictB) return (T)(1.0 / (1.0 + (double)omdtsq));
  if constexpr (sizeof(T) == 8) {
    const double d = 1.0 + omdtsq;
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    return r;
  } else {
    return (T)__builtin_amdgcn_rcpf((float)(T(1) + omdtsq));
  }
}

// Fixed-point positions (FixedParams, FP32 engines in production
// arithmetic).  The mover iterate's cell truncates the cell coordinate
// toward zero as the reference's int() does (src/Mover.cpp:45-47); the
// deposit cell is its floor (src/Mover.cpp:411-413).
//
// A displacement d (fixed units) rounded to an integer, bounded to `lim`
// (64 cells): with k chosen so that (cells + 64) << k fits in int32, the sum
// with a coordinate in the box cannot overflow.  A particle whose final
// displacement exceeds the bound (or is not a number) is reported by the
// caller; one that moves that far also leaves any grid the bound allows.
__device__ __forceinline__ int fx_step(float d, float lim) {
  return __float2int_rn(__builtin_amdgcn_fmed3f(d, -lim, lim));
}

// Offset of X from the left node of cell c, in cells: exact while X lies
// in the cell (k <= 24 bits), extrapolated beyond it.
__device__ __forceinline__ float fx_frac(int X, int c, int k, float unit) {
  return (float)(X - (c << k)) * unit;
}

// The staged block of a tile as the mover's stencil gather sees it: node
// fields [NB][6] and node coordinates in LDS, and the global node index of
// block node 0 along each axis.
template <typename T, int TX, int TY, int TZ>
struct Block {
  static constexpr int BX = TX + 3, BY = TY + 3, BZ = TZ + 3, NB = BX * BY * BZ;
  const T *fld;
  const T *xn, *yn, *zn;
  // Ex Ey Ez Bx By Bz and the node coordinates xn yn zn in global memory,
  // read from LDS by the off-block gather when it runs (volatile: not hoisted
  // into registers held across the batch loop)
  const SPIC_GLOBAL T *const volatile __attribute__((address_space(3))) *gptr;
  int c0x, c0y, c0z;
};

// One particle through the predictor-corrector mover and the particle
// boundary conditions (src/Mover.cpp:33-157), in place.  With STAGED the
// stencil comes from the block when it lies inside it.
template <typename T, bool STAGED, int TX, int TY, int TZ>
__device__ __forceinline__ void push_particle_once(const KGrid<T> &g, const KField<T> &f,
                                              const KPush<T> &k,
                                              const Block<T, TX, TY, TZ> &blk, T &x,
                                              T &y, T &z, T &u, T &v, T &w,
                                              int &bad, unsigned int &n_lds,
                                              unsigned int &n_glob) {
  using B = Block<T, TX, TY, TZ>;
  const int nyn = g.nyn, nzn = g.nzn, sxg = nyn * nzn;
  const T xt = x, yt = y, zt = z;
  T upt = T(0), vpt = T(0), wpt = T(0);
  T xc = x, yc = y, zc = z;
  int cix = -100, ciy = -100, ciz = -100;  // cell of the cached stencil
  T F[2][2][2][6];
  T X0 = T(0), X1 = T(0), Y0 = T(0), Y1 = T(0), Z0 = T(0), Z1 = T(0);
  for (int it = 0; it < k.niter; ++it) {
    int ix = bcell_trunc<T>(xc, g.xStart, g.invdx, g.nxn);
    int iy = bcell_trunc<T>(yc, g.yStart, g.invdy, g.nyn);
    int iz = bcell_trunc<T>(zc, g.zStart, g.invdz, g.nzn);
    bad |= bclamp(ix, g.nxn);
    bad |= bclamp(iy, g.nyn);
    bad |= bclamp(iz, g.nzn);
    if (ix != cix || iy != ciy || iz != ciz) {
      cix = ix; ciy = iy; ciz = iz;
      const int bx = ix - 1 - blk.c0x, by = iy - 1 - blk.c0y, bz = iz - 1 - blk.c0z;
      const bool in_block = STAGED && (bx >= 0) & (bx <= B::BX - 2) & (by >= 0) &
                                          (by <= B::BY - 2) & (bz >= 0) & (bz <= B::BZ - 2);
      if (in_block) {
        ++n_lds;
        X0 = blk.xn[bx]; X1 = blk.xn[bx + 1];
        Y0 = blk.yn[by]; Y1 = blk.yn[by + 1];
        Z0 = blk.zn[bz]; Z1 = blk.zn[bz + 1];
        // node (bx + 1 - aa, by + 1 - bb, bz + 1 - cc) lies a compile-time
        // offset before node (bx + 1, by + 1, bz + 1): one address for all
        const T *fb = blk.fld + (((bx + 1) * B::BY + (by + 1)) * B::BZ + (bz + 1)) * 6;
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const int off = ((aa * B::BY + bb) * B::BZ + cc) * 6;
#pragma unroll
              for (int c = 0; c < 6; ++c) F[aa][bb][cc][c] = fb[c - off];
            }
      } else {
        ++n_glob;
        const SPIC_GLOBAL T *gp[9];
        if constexpr (STAGED) {
#pragma unroll
          for (int c = 0; c < 9; ++c) gp[c] = blk.gptr[c];
        } else {
          gp[0] = gmem(f.Ex); gp[1] = gmem(f.Ey); gp[2] = gmem(f.Ez);
          gp[3] = gmem(f.Bx); gp[4] = gmem(f.By); gp[5] = gmem(f.Bz);
          gp[6] = gmem(g.xn); gp[7] = gmem(g.yn); gp[8] = gmem(g.zn);
        }
        X0 = gp[6][ix - 1]; X1 = gp[6][ix];
        Y0 = gp[7][iy - 1]; Y1 = gp[7][iy];
        Z0 = gp[8][iz - 1]; Z1 = gp[8][iz];
        const int gb = (ix * nyn + iy) * nzn + iz;
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const int n = gb - aa * sxg - bb * nzn - cc;
#pragma unroll
              for (int c = 0; c < 6; ++c) F[aa][bb][cc][c] = gp[c][n];
            }
        if constexpr (STAGED) wait_vm_all();
      }
    }
    const T xi[2] = {xc - X0, X1 - xc};
    const T eta[2] = {yc - Y0, Y1 - yc};
    const T zeta[2] = {zc - Z0, Z1 - zc};
    // strict: the reference's ((xi eta) zeta) invVOL; production: the same
    // product as (xi invVOL) eta, then zeta (14 multiplies, not 20)
    const T xv[2] = {kStrictB ? xi[0] : xi[0] * g.invVOL, kStrictB ? xi[1] : xi[1] * g.invVOL};
    T Exl = T(0), Eyl = T(0), Ezl = T(0), Bxl = T(0), Byl = T(0), Bzl = T(0);
#pragma unroll
    for (int aa = 0; aa < 2; ++aa)
#pragma unroll
      for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          const T wgt = kStrictB ? xi[aa] * eta[bb] * zeta[cc] * g.invVOL
                                 : xv[aa] * eta[bb] * zeta[cc];
          Exl += wgt * F[aa][bb][cc][0];
          Eyl += wgt * F[aa][bb][cc][1];
          Ezl += wgt * F[aa][bb][cc][2];
          Bxl += wgt * F[aa][bb][cc][3];
          Byl += wgt * F[aa][bb][cc][4];
          Bzl += wgt * F[aa][bb][cc][5];
        }
    const T q2 = k.qomdt2;
    const T omdtsq = q2 * q2 * (Bxl * Bxl + Byl * Byl + Bzl * Bzl);
    const T denom = bdenom<T>(omdtsq);
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
  if (kStrictB || sizeof(T) == 8) {
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
  // particle boundary conditions: two independent tests per axis
  // (src/Mover.cpp:104-150)
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
}

// The push once, or k.nloop times with the particle BC after each (the
// reference GPU mover's sub-cycle loop, src/gpu/Particles_gpu.cu:336-437).
template <typename T, bool STAGED, int TX, int TY, int TZ>
__device__ __forceinline__ void push_particle(const KGrid<T> &g, const KField<T> &f,
                                              const KPush<T> &k,
                                              const Block<T, TX, TY, TZ> &blk, T &x, T &y,
                                              T &z, T &u, T &v, T &w, int &bad,
                                              unsigned int &n_lds, unsigned int &n_glob) {
#pragma unroll 1
  for (int isub = 0; isub < k.nloop; ++isub)
    push_particle_once<T, STAGED, TX, TY, TZ>(g, f, k, blk, x, y, z, u, v, w, bad, n_lds,
                                              n_glob);
}

// push_particle for fixed-point positions (FP32 form): the same
// predictor-corrector iteration, rotation and final update, with the
// interpolation weights as offsets within the cell in cell units (so that
// the weight is xi eta zeta invVOL of the reference without the multiply),
// the iterate and the new position advanced by displacements rounded to the
// fixed-point unit, and the boundary conditions applied exactly on the
// integers: x > Lx is X > lim, a reflection lim - (X - lim), a wrap
// X - lim.
//
// The iterate's cell truncates toward zero as the reference's int() does
// (src/Mover.cpp:45-47): an arithmetic shift (the floor), corrected for
// negative coordinates only in waves that hold one.  An iterate whose
// stencil lies in the staged block (lim_b: the last valid lower-node index
// per axis, the block's edge or the grid's) needs no clamp: every index is
// valid.  Any other takes the clamped, flagged indices of bclamp and
// gathers from global memory.  The cached stencil is keyed by its lower
// block node (staged) or by its packed global cell (not staged).
template <bool STAGED, int TX, int TY, int TZ>
__device__ __forceinline__ void push_fixed_once(const KGrid<float> &g, const KField<float> &f,
                                           const KPush<float> &k, const FixedParams &fp,
                                           const Block<float, TX, TY, TZ> &blk,
                                           const int lim_b[3], int &X, int &Y, int &Z,
                                           float &u, float &v, float &w, int &bad,
                                           unsigned int &n_lds, unsigned int &n_glob) {
  using B = Block<float, TX, TY, TZ>;
  const int nyn = g.nyn, nzn = g.nzn, sxg = nyn * nzn;
  const int Xt = X, Yt = Y, Zt = Z;
  int Xc = X, Yc = Y, Zc = Z;
  float upt = 0.f, vpt = 0.f, wpt = 0.f;
  int ckey = 0x7fffffff;  // key of the cached stencil (none yet)
  float F[2][2][2][6];
  const int kx = fp.k[0], ky = fp.k[1], kz = fp.k[2];
  const int mx = (1 << kx) - 1, my = (1 << ky) - 1, mz = (1 << kz) - 1;
  // fixed units per unit velocity over half a step and over the step; a
  // displacement is bounded to 64 cells, which keeps the sum with any
  // coordinate of the box in int32 (fixed_params leaves 64 cells of room)
  const float hx = k.dto2 * fp.s[0], hy = k.dto2 * fp.s[1], hz = k.dto2 * fp.s[2];
  const float bx_lim = (float)(64 << kx), by_lim = (float)(64 << ky), bz_lim = (float)(64 << kz);
  for (int it = 0; it < k.niter; ++it) {
    // physical cell (floor) and offset within it, in fixed units
    int cx = Xc >> kx, cy = Yc >> ky, cz = Zc >> kz;
    int ox_i = Xc & mx, oy_i = Yc & my, oz_i = Zc & mz;
    if (__builtin_expect(__any(min(Xc, min(Yc, Zc)) < 0), 0)) {
      // truncation toward zero: -((-X) >> k), offset -((-X) & mask)
      if (Xc < 0) { cx = -((-Xc) >> kx); ox_i = -((-Xc) & mx); }
      if (Yc < 0) { cy = -((-Yc) >> ky); oy_i = -((-Yc) & my); }
      if (Zc < 0) { cz = -((-Zc) >> kz); oz_i = -((-Zc) & mz); }
    }
    // block index of the stencil's lower node (global node cx + 1)
    const int bx = cx + 1 - blk.c0x, by = cy + 1 - blk.c0y, bz = cz + 1 - blk.c0z;
    const bool in_block = STAGED && ((unsigned)bx <= (unsigned)lim_b[0]) &
                                        ((unsigned)by <= (unsigned)lim_b[1]) &
                                        ((unsigned)bz <= (unsigned)lim_b[2]);
    int key, ix = 0, iy = 0, iz = 0;
    if (in_block) {
      key = (bx * B::BY + by) * B::BZ + bz;
    } else {
      ix = cx + 2; iy = cy + 2; iz = cz + 2;
      bad |= bclamp(ix, g.nxn);
      bad |= bclamp(iy, g.nyn);
      bad |= bclamp(iz, g.nzn);
      // offsets from the (clamped) cell's left node, extrapolated beyond it
      ox_i = Xc - ((ix - 2) << kx);
      oy_i = Yc - ((iy - 2) << ky);
      oz_i = Zc - ((iz - 2) << kz);
      key = ~((ix << 20) | (iy << 10) | iz);
    }
    if (key != ckey) {
      ckey = key;
      if (in_block) {
        ++n_lds;
        // node (bx + 1 - aa, by + 1 - bb, bz + 1 - cc) lies a compile-time
        // offset before the upper corner (bx + 1, by + 1, bz + 1)
        const float *fb = blk.fld + (key + (B::BY + 1) * B::BZ + 1) * 6;
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const int off = ((aa * B::BY + bb) * B::BZ + cc) * 6;
#pragma unroll
              for (int c = 0; c < 6; ++c) F[aa][bb][cc][c] = fb[c - off];
            }
      } else {
        ++n_glob;
        const SPIC_GLOBAL float *gp[6];
        if constexpr (STAGED) {
#pragma unroll
          for (int c = 0; c < 6; ++c) gp[c] = blk.gptr[c];
        } else {
          gp[0] = gmem(f.Ex); gp[1] = gmem(f.Ey); gp[2] = gmem(f.Ez);
          gp[3] = gmem(f.Bx); gp[4] = gmem(f.By); gp[5] = gmem(f.Bz);
        }
        const int gb = (ix * nyn + iy) * nzn + iz;
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              const int n = gb - aa * sxg - bb * nzn - cc;
#pragma unroll
              for (int c = 0; c < 6; ++c) F[aa][bb][cc][c] = gp[c][n];
            }
        if constexpr (STAGED) wait_vm_all();
      }
    }
    const float ox = (float)ox_i * fp.unit[0];
    const float oy = (float)oy_i * fp.unit[1];
    const float oz = (float)oz_i * fp.unit[2];
    const float xi[2] = {ox, 1.f - ox};
    const float eta[2] = {oy, 1.f - oy};
    const float zeta[2] = {oz, 1.f - oz};
    float Exl = 0.f, Eyl = 0.f, Ezl = 0.f, Bxl = 0.f, Byl = 0.f, Bzl = 0.f;
#pragma unroll
    for (int aa = 0; aa < 2; ++aa)
#pragma unroll
      for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          const float wgt = xi[aa] * eta[bb] * zeta[cc];
          Exl += wgt * F[aa][bb][cc][0];
          Eyl += wgt * F[aa][bb][cc][1];
          Ezl += wgt * F[aa][bb][cc][2];
          Bxl += wgt * F[aa][bb][cc][3];
          Byl += wgt * F[aa][bb][cc][4];
          Bzl += wgt * F[aa][bb][cc][5];
        }
    const float q2 = k.qomdt2;
    const float omdtsq = q2 * q2 * (Bxl * Bxl + Byl * Byl + Bzl * Bzl);
    const float denom = bdenom<float>(omdtsq);
    const float ut = u + q2 * Exl;
    const float vt = v + q2 * Eyl;
    const float wt = w + q2 * Ezl;
    const float udotb = ut * Bxl + vt * Byl + wt * Bzl;
    upt = (ut + q2 * (vt * Bzl - wt * Byl + q2 * udotb * Bxl)) * denom;
    vpt = (vt + q2 * (wt * Bxl - ut * Bzl + q2 * udotb * Byl)) * denom;
    wpt = (wt + q2 * (ut * Byl - vt * Bxl + q2 * udotb * Bzl)) * denom;
    Xc = Xt + fx_step(upt * hx, bx_lim);
    Yc = Yt + fx_step(vpt * hy, by_lim);
    Zc = Zt + fx_step(wpt * hz, bz_lim);
  }
  u = 2.f * upt - u;
  v = 2.f * vpt - v;
  w = 2.f * wpt - w;
  const float dX = upt * (k.dt_sub * fp.s[0]), dY = vpt * (k.dt_sub * fp.s[1]),
              dZ = wpt * (k.dt_sub * fp.s[2]);
  bad |= !((fabsf(dX) <= bx_lim) & (fabsf(dY) <= by_lim) & (fabsf(dZ) <= bz_lim));
  X = Xt + fx_step(dX, bx_lim);
  Y = Yt + fx_step(dY, by_lim);
  Z = Zt + fx_step(dZ, bz_lim);
  // particle boundary conditions, two independent tests per axis
  // (src/Mover.cpp:104-150), exact on the integers
  if (X > fp.lim[0]) {
    if (k.perx) X -= fp.lim[0];
    else { u = -u; X = fp.lim[0] - (X - fp.lim[0]); }
  }
  if (X < 0) {
    if (k.perx) X += fp.lim[0];
    else { u = -u; X = -X; }
  }
  if (Y > fp.lim[1]) {
    if (k.pery) Y -= fp.lim[1];
    else { v = -v; Y = fp.lim[1] - (Y - fp.lim[1]); }
  }
  if (Y < 0) {
    if (k.pery) Y += fp.lim[1];
    else { v = -v; Y = -Y; }
  }
  if (Z > fp.lim[2]) {
    if (k.perz) Z -= fp.lim[2];
    else { w = -w; Z = fp.lim[2] - (Z - fp.lim[2]); }
  }
  if (Z < 0) {
    if (k.perz) Z += fp.lim[2];
    else { w = -w; Z = -Z; }
  }
}

template <bool STAGED, int TX, int TY, int TZ>
__device__ __forceinline__ void push_fixed(const KGrid<float> &g, const KField<float> &f,
                                           const KPush<float> &k, const FixedParams &fp,
                                           const Block<float, TX, TY, TZ> &blk,
                                           const int lim_b[3], int &X, int &Y, int &Z,
                                           float &u, float &v, float &w, int &bad,
                                           unsigned int &n_lds, unsigned int &n_glob) {
#pragma unroll 1
  for (int isub = 0; isub < k.nloop; ++isub)
    push_fixed_once<STAGED, TX, TY, TZ>(g, f, k, fp, blk, lim_b, X, Y, Z, u, v, w, bad, n_lds,
                                        n_glob);
}
// Tile of the deposit cell (ix, iy, iz), clamped to the physical cells;
// the tile shape is a compile-time constant (divisions become shifts).
template <int TX, int TY, int TZ, typename T>
__device__ __forceinline__ int tile_of(const BinnedArgs<T> &a, int ix, int iy, int iz) {
  const int pcx = min(max(ix - 2, 0), a.ncx - 1);
  const int pcy = min(max(iy - 2, 0), a.ncy - 1);
  const int pcz = min(max(iz - 2, 0), a.ncz - 1);
  return ((pcx / TX) * a.nty + pcy / TY) * a.ntz + pcz / TZ;
}

// Slots in the phase kernels are 32-bit: the host keeps every component
// array of a species below 4 GiB (spic_engine_load), so a slot's byte
// offset fits in 32 bits and each load or store is one uniform component
// base plus one per-lane offset.  kNoSlot marks "no slot".
constexpr uint32_t kNoSlot = 0xffffffffu;

// The particles of range t (see RangeView) with 32-bit slots.
struct SlotRange {
  uint32_t first, count, gap_lo, gap_len;
  __device__ uint32_t slot(uint32_t r) const { return first + (r < gap_lo ? r : r + gap_len); }
};

template <typename T>
__device__ __forceinline__ SlotRange slot_range(const BinnedArgs<T> &a,
                                                const BinnedSpecies<T> &sp, int t) {
  const RangeView v = range_view(gmem(sp.seg), gmem(sp.cnt_in), a.ntiles, sp.ovf_cap, t);
  return SlotRange{(uint32_t)v.first, (uint32_t)v.count, v.gap_lo, v.gap_len};
}

// Component c (x y z u v w q) of the slot at byte offset `off`.
template <typename T>
__device__ __forceinline__ T load_comp(const BinnedSpecies<T> &sp, int c, uint32_t off) {
  const SPIC_GLOBAL char *base =
      (const SPIC_GLOBAL char *)gmem(sp.in) + (size_t)c * (size_t)sp.stride * sizeof(T);
  return *(const SPIC_GLOBAL T *)(base + off);
}

// A particle word written once per phase.  FP64 stores are nontemporal
// (streaming): 1.3% faster FP64 phases on the 262M-particle sheet, for
// 7-12% more bytes written (partial lines leave L2 unmerged).  The FP32
// form gains nothing from them and keeps cached stores; nontemporal loads
// changed neither form.
template <typename T, typename W>
__device__ __forceinline__ void store_word(SPIC_GLOBAL W *p, W v) {
  if constexpr (sizeof(T) == 8) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <typename T>
__device__ __forceinline__ void store_comp(const BinnedSpecies<T> &sp, int c, uint32_t off,
                                           T v) {
  SPIC_GLOBAL char *base =
      (SPIC_GLOBAL char *)gmem(sp.out) + (size_t)c * (size_t)sp.stride * sizeof(T);
  store_word<T>((SPIC_GLOBAL T *)(base + off), v);
}

// Slot of a particle to which segment `dest`'s counter gave `k`: in the
// segment while it has room, else the next slot of the overflow region;
// kNoSlot if that is full too.
template <typename T>
__device__ __forceinline__ uint32_t seg_slot(const BinnedArgs<T> &a,
                                             const BinnedSpecies<T> &sp, int dest,
                                             unsigned int k) {
  const uint32_t lo = (uint32_t)gmem(sp.seg)[dest];
  const uint32_t cap = (uint32_t)gmem(sp.seg)[dest + 1] - lo;
  if (k < cap) return lo + k;
  const unsigned int o = gatomic_add(sp.cnt_out + a.ntiles, 1u);
  return o < sp.ovf_cap ? (uint32_t)gmem(sp.seg)[a.ntiles] + o : kNoSlot;
}

template <typename T>
__device__ __forceinline__ void write_particle(const BinnedSpecies<T> &sp, uint32_t o,
                                               T x, T y, T z, T u, T v, T w, T q,
                                               uint32_t pid) {
  const uint32_t off = o * (uint32_t)sizeof(T);
  store_comp(sp, 0, off, x); store_comp(sp, 1, off, y); store_comp(sp, 2, off, z);
  store_comp(sp, 3, off, u); store_comp(sp, 4, off, v); store_comp(sp, 5, off, w);
  store_comp(sp, 6, off, q);
  // the particle's load index, in engines that keep the load order
  if (sp.id_out)
    store_word<T>((SPIC_GLOBAL uint32_t *)((SPIC_GLOBAL char *)gmem(sp.id_out) + o * 4u), pid);
}

// Write a pushed particle to the output buffer.  A particle that stays in
// its tile takes slot j of the block the workgroup reserved in its own
// segment (own_lo + stay_base + j while that is below own_cap); a leaver
// takes the next slot of its destination segment's counter; when the
// segment is full, the next slot of the overflow region.  False if that is
// full too (the particle is lost and the call reports it).
template <typename T>
__device__ __forceinline__ bool put_particle(const BinnedArgs<T> &a,
                                             const BinnedSpecies<T> &sp, uint32_t own_lo,
                                             uint32_t own_cap, unsigned int stay_base,
                                             bool stay, unsigned int j, int dest, T x, T y,
                                             T z, T u, T v, T w, T q, uint32_t pid) {
  uint32_t o;
  if (stay) {
    const unsigned int k = stay_base + j;
    if (k < own_cap) {
      o = own_lo + k;
    } else {
      const unsigned int e = gatomic_add(sp.cnt_out + a.ntiles, 1u);
      o = e < sp.ovf_cap ? (uint32_t)gmem(sp.seg)[a.ntiles] + e : kNoSlot;
    }
  } else {
    o = seg_slot(a, sp, dest, gatomic_add(sp.cnt_out + dest, 1u));
  }
  if (o == kNoSlot) return false;
  write_particle(sp, o, x, y, z, u, v, w, q, pid);
  return true;
}

// The ten moment terms of one node weight, in the reference's operation
// order (src/Mover.cpp:430-568): rho = wgt*iv, Jx = (u*wgt)*iv, ...,
// pzz = ((w*w)*wgt)*iv, each converted to A and added to acc.
template <typename T, typename A>
__device__ __forceinline__ void add_terms(A *acc, T wgt, T iv, T u, T v, T w, T uu,
                                          T uv, T uw, T vv, T vw, T ww) {
  acc[0] += (A)(wgt * iv);
  acc[1] += (A)((u * wgt) * iv);
  acc[2] += (A)((v * wgt) * iv);
  acc[3] += (A)((w * wgt) * iv);
  acc[4] += (A)((uu * wgt) * iv);
  acc[5] += (A)((uv * wgt) * iv);
  acc[6] += (A)((uw * wgt) * iv);
  acc[7] += (A)((vv * wgt) * iv);
  acc[8] += (A)((vw * wgt) * iv);
  acc[9] += (A)((ww * wgt) * iv);
}

// All 80 contributions of one particle with global FP64 atomics.
template <typename T>
__device__ __forceinline__ void deposit_global(const KGrid<T> &g, const KMom &m, int ix,
                                               int iy, int iz, T x, T y, T z, T u, T v,
                                               T w, T q) {
  const int nyn = g.nyn, nzn = g.nzn, sxg = nyn * nzn;
  const T xi[2] = {x - gmem(g.xn)[ix - 1], gmem(g.xn)[ix] - x};
  const T eta[2] = {y - gmem(g.yn)[iy - 1], gmem(g.yn)[iy] - y};
  const T zeta[2] = {z - gmem(g.zn)[iz - 1], gmem(g.zn)[iz] - z};
  const T iv = g.invVOL;
  const T uu = u * u, uv = u * v, uw = u * w, vv = v * v, vw = v * w, ww = w * w;
  const int gb = (ix * nyn + iy) * nzn + iz;
#pragma unroll
  for (int aa = 0; aa < 2; ++aa)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const T wgt = q * xi[aa] * eta[bb] * zeta[cc] * iv;
        double t[10];
#pragma unroll
        for (int mm = 0; mm < 10; ++mm) t[mm] = 0.0;
        add_terms<T, double>(t, wgt, iv, u, v, w, uu, uv, uw, vv, vw, ww);
        const int n = gb - aa * sxg - bb * nzn - cc;
#pragma unroll
        for (int mm = 0; mm < 10; ++mm) gatomic_add(m.m[mm] + n, t[mm]);
      }
}

// One node (corner aa, bb, cc) of one particle's deposit: its ten terms
// added to the LDS block when the particle's stencil lies inside the
// staged block, else to the global FP64 arrays.
template <typename T, typename A, int BX, int BY, int BZ>
__device__ __forceinline__ bool deposit_node(const KGrid<T> &g, const KMom &m, A *s_mom,
                                             const T *s_xn, const T *s_yn, const T *s_zn,
                                             int c0x, int c0y, int c0z, int ix, int iy,
                                             int iz, int aa, int bb, int cc, T x, T y,
                                             T z, T u, T v, T w, T q) {
  constexpr int NB = BX * BY * BZ;
  const int bx = ix - 1 - c0x, by = iy - 1 - c0y, bz = iz - 1 - c0z;
  const bool in_block = (bx >= 0) & (bx <= BX - 2) & (by >= 0) & (by <= BY - 2) &
                        (bz >= 0) & (bz <= BZ - 2);
  T xi, eta, zeta;
  if (in_block) {
    xi = aa == 0 ? x - s_xn[bx] : s_xn[bx + 1] - x;
    eta = bb == 0 ? y - s_yn[by] : s_yn[by + 1] - y;
    zeta = cc == 0 ? z - s_zn[bz] : s_zn[bz + 1] - z;
  } else {
    xi = aa == 0 ? x - gmem(g.xn)[ix - 1] : gmem(g.xn)[ix] - x;
    eta = bb == 0 ? y - gmem(g.yn)[iy - 1] : gmem(g.yn)[iy] - y;
    zeta = cc == 0 ? z - gmem(g.zn)[iz - 1] : gmem(g.zn)[iz] - z;
  }
  const T iv = g.invVOL;
  const T wgt = q * xi * eta * zeta * iv;
  A t[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) t[k] = A(0);
  add_terms<T, A>(t, wgt, iv, u, v, w, u * u, u * v, u * w, v * v, v * w, w * w);
  if (in_block) {
    const int nb = ((bx + 1 - aa) * BY + (by + 1 - bb)) * BZ + (bz + 1 - cc);
#pragma unroll
    for (int k = 0; k < 10; ++k) atomicAdd(&s_mom[k * NB + nb], t[k]);
  } else {
    const int n = ((ix - aa) * g.nyn + (iy - bb)) * g.nzn + (iz - cc);
#pragma unroll
    for (int k = 0; k < 10; ++k) gatomic_add(m.m[k] + n, (double)t[k]);
  }
  return in_block;
}

// deposit_node and deposit_global for fixed-point positions: the weight of
// corner (aa, bb, cc) is q times the offsets of the particle within its
// deposit cell (ix, iy, iz) in cell units, q xi eta zeta invVOL of the
// reference; each moment term is then that weight times p_m times invVOL,
// as the reference adds it (src/Mover.cpp:430-568).
template <typename A, int BX, int BY, int BZ>
__device__ __forceinline__ bool deposit_node_fx(const KGrid<float> &g, const FixedParams &fp,
                                                const KMom &m, A *s_mom, int c0x, int c0y,
                                                int c0z, int ix, int iy, int iz, int aa, int bb,
                                                int cc, int X, int Y, int Z, float u, float v,
                                                float w, float q) {
  constexpr int NB = BX * BY * BZ;
  const int bx = ix - 1 - c0x, by = iy - 1 - c0y, bz = iz - 1 - c0z;
  const bool in_block = (bx >= 0) & (bx <= BX - 2) & (by >= 0) & (by <= BY - 2) &
                        (bz >= 0) & (bz <= BZ - 2);
  const float ox = fx_frac(X, ix - 2, fp.k[0], fp.unit[0]);
  const float oy = fx_frac(Y, iy - 2, fp.k[1], fp.unit[1]);
  const float oz = fx_frac(Z, iz - 2, fp.k[2], fp.unit[2]);
  const float xi = aa == 0 ? ox : 1.f - ox;
  const float eta = bb == 0 ? oy : 1.f - oy;
  const float zeta = cc == 0 ? oz : 1.f - oz;
  const float iv = g.invVOL;
  const float wgt = q * xi * eta * zeta;
  A t[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) t[k] = A(0);
  add_terms<float, A>(t, wgt, iv, u, v, w, u * u, u * v, u * w, v * v, v * w, w * w);
  if (in_block) {
    const int nb = ((bx + 1 - aa) * BY + (by + 1 - bb)) * BZ + (bz + 1 - cc);
#pragma unroll
    for (int k = 0; k < 10; ++k) atomicAdd(&s_mom[k * NB + nb], t[k]);
  } else {
    const int n = ((ix - aa) * g.nyn + (iy - bb)) * g.nzn + (iz - cc);
#pragma unroll
    for (int k = 0; k < 10; ++k) gatomic_add(m.m[k] + n, (double)t[k]);
  }
  return in_block;
}

__device__ __forceinline__ void deposit_global_fx(const KGrid<float> &g, const FixedParams &fp,
                                                  const KMom &m, int ix, int iy, int iz, int X,
                                                  int Y, int Z, float u, float v, float w,
                                                  float q) {
  const int nyn = g.nyn, nzn = g.nzn, sxg = nyn * nzn;
  const float ox = fx_frac(X, ix - 2, fp.k[0], fp.unit[0]);
  const float oy = fx_frac(Y, iy - 2, fp.k[1], fp.unit[1]);
  const float oz = fx_frac(Z, iz - 2, fp.k[2], fp.unit[2]);
  const float xi[2] = {ox, 1.f - ox};
  const float eta[2] = {oy, 1.f - oy};
  const float zeta[2] = {oz, 1.f - oz};
  const float iv = g.invVOL;
  const float uu = u * u, uv = u * v, uw = u * w, vv = v * v, vw = v * w, ww = w * w;
  const int gb = (ix * nyn + iy) * nzn + iz;
#pragma unroll
  for (int aa = 0; aa < 2; ++aa)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const float wgt = q * xi[aa] * eta[bb] * zeta[cc];
        double t[10];
#pragma unroll
        for (int mm = 0; mm < 10; ++mm) t[mm] = 0.0;
        add_terms<float, double>(t, wgt, iv, u, v, w, uu, uv, uw, vv, vw, ww);
        const int n = gb - aa * sxg - bb * nzn - cc;
#pragma unroll
        for (int mm = 0; mm < 10; ++mm) gatomic_add(m.m[mm] + n, t[mm]);
      }
}

// Deposit cell of a fixed-point position, 2 + floor of the cell
// coordinate, clamped as bclamp does.
__device__ __forceinline__ int fx_deposit_cell(int X, int k, int n, bool &bad) {
  int c = 2 + (X >> k);
  bad |= bclamp(c, n);
  return c;
}

// A: type of the LDS moment block (T, or double for FP32 particles); the
// per-thread register sums are in the particle precision T.  FP32 kernels
// ask for three waves per SIMD (at most 168 VGPRs).  FX: fixed-point
// positions (FP32, production arithmetic): the x y z words of the particle
// storage, the output buffer and the list ring hold the int32 coordinates,
// the buckets the particle's offsets within its deposit cell.
template <typename T, typename A, int TX, int TY, int TZ, bool PUSH = true, bool FX = false,
          int PPT = 1>
__device__ __forceinline__ void binned_phase_tile(const BinnedArgs<T> &a, const int tile) {
  static_assert(!FX || (sizeof(T) == 4 && !kStrictB), "fixed point: FP32 production only");
  static_assert(PPT == 1 || PPT == 2, "one or two particles per thread and batch");
  using Blk = Block<T, TX, TY, TZ>;
  constexpr int NC = TX * TY * TZ;  // cells per tile
  constexpr int NT = 4 * NC;        // threads: four per cell
  constexpr int BX = Blk::BX, BY = Blk::BY, BZ = Blk::BZ, NB = Blk::NB;
  // PPT: particles each thread pushes and bins per batch before the bucket
  // step, which then sums PPT x NT particles; the buckets grow with it
  constexpr int CAP = kBucketCap * PPT;
  // bucket stride per cell: odd, so that the 16 cells a wave reads in the
  // bucket step fall on distinct LDS banks
  constexpr int CS = CAP | 1;
  constexpr int NL = kListCap;  // list ring entries
  __shared__ T s_fld[NB * 6];
  __shared__ A s_mom[10 * NB];
  __shared__ T s_xn[BX], s_yn[BY], s_zn[BZ];
  __shared__ T s_bkt[7][NC * CS];  // x y z q u v w of the bucketed particles
  __shared__ T s_lst[7][NL];       // x y z q u v w of the listed particles
  __shared__ unsigned int s_lcell[NL];  // their deposit cells, packed
  __shared__ uint32_t s_lid[NL];        // and their ids
  __shared__ unsigned int s_cnt[2][NC];
  __shared__ unsigned int s_nl;    // entries appended to the ring this species
  __shared__ unsigned int s_stay;  // particles written that stay in the tile
  __shared__ unsigned int s_base;  // start of the block reserved for them
  __shared__ const SPIC_GLOBAL T *s_gptr[9];  // Block::gptr

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int tz = tile % a.ntz, ty = (tile / a.ntz) % a.nty, tx = tile / (a.ntz * a.nty);
  // block node b along x is global node c0x + b; cell c of the tile uses
  // block nodes c+1 and c+2
  Blk blk;
  blk.fld = s_fld;
  blk.xn = s_xn; blk.yn = s_yn; blk.zn = s_zn;
  blk.c0x = tx * TX; blk.c0y = ty * TY; blk.c0z = tz * TZ;
  // last lower-node index (block coordinates) of a stencil inside both the
  // staged block and the grid: a stencil within these needs no clamp
  const int lim_b[3] = {min(BX - 2, a.g.nxn - 2 - blk.c0x), min(BY - 2, a.g.nyn - 2 - blk.c0y),
                        min(BZ - 2, a.g.nzn - 2 - blk.c0z)};
  blk.gptr = (const SPIC_GLOBAL T *const volatile __attribute__((address_space(3))) *)s_gptr;
  if (tid == 0) {
    s_gptr[0] = gmem(a.f.Ex); s_gptr[1] = gmem(a.f.Ey); s_gptr[2] = gmem(a.f.Ez);
    s_gptr[3] = gmem(a.f.Bx); s_gptr[4] = gmem(a.f.By); s_gptr[5] = gmem(a.f.Bz);
    s_gptr[6] = gmem(a.g.xn); s_gptr[7] = gmem(a.g.yn); s_gptr[8] = gmem(a.g.zn);
  }
  const int nyn = a.g.nyn, nzn = a.g.nzn;

  // 1. stage node coordinates and fields
  for (int b = tid; b < BX + BY + BZ; b += NT) {
    if (b < BX) s_xn[b] = (blk.c0x + b < a.g.nxn) ? gmem(a.g.xn)[blk.c0x + b] : T(0);
    else if (b < BX + BY)
      s_yn[b - BX] = (blk.c0y + b - BX < nyn) ? gmem(a.g.yn)[blk.c0y + b - BX] : T(0);
    else
      s_zn[b - BX - BY] =
          (blk.c0z + b - BX - BY < nzn) ? gmem(a.g.zn)[blk.c0z + b - BX - BY] : T(0);
  }
  for (int b = PUSH ? tid : NB; b < NB; b += NT) {  // fields: for the push only
    const int bz = b % BZ, by = (b / BZ) % BY, bx = b / (BZ * BY);
    const int gx = blk.c0x + bx, gy = blk.c0y + by, gz = blk.c0z + bz;
    T v[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (gx < a.g.nxn && gy < nyn && gz < nzn) {
      const int n = (gx * nyn + gy) * nzn + gz;
      v[0] = gmem(a.f.Ex)[n]; v[1] = gmem(a.f.Ey)[n]; v[2] = gmem(a.f.Ez)[n];
      v[3] = gmem(a.f.Bx)[n]; v[4] = gmem(a.f.By)[n]; v[5] = gmem(a.f.Bz)[n];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) s_fld[b * 6 + c] = v[c];
  }

  // this thread's cell and node pair in the bucket step: corners (a, b, 0)
  // and (a, b, 1) of cell (mlx, mly, mlz)
  const int my_cell = tid >> 2;
  const int my_a = (tid >> 1) & 1, my_b = tid & 1;
  const int mlx = my_cell / (TY * TZ), mly = (my_cell / TZ) % TY, mlz = my_cell % TZ;

  // per-lane error flags (bit 0: move, 1: deposit cell, 2: store), an int
  // so that the loops carry them in a VGPR: carried bools become 64-bit
  // lane masks in scalar registers, which the FP64 kernel spills
  int bad_bits = 0;
  unsigned int n_gg = 0, n_gd = 0, n_lg = 0, n_ld = 0;

  for (int s = 0; s < a.ns; ++s) {
    const BinnedSpecies<T> &sp = a.sp[s];
    const SlotRange in = slot_range(a, sp, tile);
    const uint32_t nin = in.count;
    if (tid == 0) gmem(a.pflag)[(size_t)tile * a.ns + s] = nin > 0 ? 1 : 0;
    if (nin == 0) continue;  // uniform across the workgroup
    for (int i = tid; i < 10 * NB; i += NT) s_mom[i] = A(0);
    for (int i = tid; i < 2 * NC; i += NT) (&s_cnt[0][0])[i] = 0u;
    if (tid == 0) {
      s_nl = 0u;
      s_stay = 0u;
      // at most nin particles stay: reserve that many slots of the output
      // segment in one block
      s_base = gatomic_add(sp.cnt_out + tile, (unsigned int)nin);
    }
    // entries [0, lst_done) of the ring have been deposited; the same value
    // in every thread
    unsigned int lst_done = 0u;
    __syncthreads();
    const unsigned int stay_base = s_base;
    // this tile's output segment
    const uint32_t own_lo = (uint32_t)gmem(sp.seg)[tile];
    const uint32_t own_cap = (uint32_t)gmem(sp.seg)[tile + 1] - own_lo;
    T acc[20];  // register sums, particle precision
#pragma unroll
    for (int k = 0; k < 20; ++k) acc[k] = T(0);

    const uint32_t nbatches = (nin + NT * PPT - 1) / (NT * PPT);
    // this thread's particle of batch j (x y z u v w q and its id), loaded
    // one batch ahead.  PREFETCH: batch j + 1 is loaded while batch j is
    // processed -- FP32, which has the registers for it at three waves per
    // SIMD, and the deposit alone (8.2 -> 7.3 ms per FP64 deposit of the
    // 262M-particle sheet).  LATE (the FP64 phase, whose push needs all 256
    // VGPRs; loading during the push spills): batch j + 1 is loaded once
    // batch j is pushed and binned, before its particles are stored, so that
    // the loads cannot take the registers those stores still read (a load
    // into a register a pending store reads waits for the store) and are in
    // flight over the stores, the bucket and the list steps: 2% faster than
    // loading at the start of each batch, and than loading after the stores.
    constexpr bool PREFETCH = sizeof(T) == 4 || !PUSH;
    constexpr bool LATE = !PREFETCH;
    T nxt[7];
    uint32_t nxt_id = 0;
    auto load_particle = [&](uint32_t r) {
#pragma unroll
      for (int c = 0; c < 7; ++c) nxt[c] = T(0);
      nxt_id = 0;
      if (r < nin) {
        const uint32_t i = in.slot(r);
        const uint32_t off = i * (uint32_t)sizeof(T);
#pragma unroll
        for (int c = 0; c < 7; ++c) nxt[c] = load_comp(sp, c, off);
        if (sp.id_in)
          nxt_id = *(const SPIC_GLOBAL uint32_t *)((const SPIC_GLOBAL char *)gmem(sp.id_in) +
                                                   i * 4u);
      }
    };
    if (PREFETCH || LATE) load_particle(tid);
    for (uint32_t j = 0; j < nbatches; ++j) {
      const int buf = (int)(j & 1);
      int bad_move = 0, bad_dep = 0, bad_store = 0;  // this batch's
#pragma unroll 1
      for (int pp = 0; pp < PPT; ++pp) {
      // this thread's particle pp of batch j, and whether one follows
      const uint32_t step = j * (uint32_t)PPT + (uint32_t)pp;
      const uint32_t r = step * NT + tid;
      const bool active = r < nin;
      if (!PREFETCH && !LATE) load_particle(r);
      T x = nxt[0], y = nxt[1], z = nxt[2], u = nxt[3], v = nxt[4], w = nxt[5], q = nxt[6];
      uint32_t pid = nxt_id;
      if (PREFETCH && step + 1 < nbatches * (uint32_t)PPT) load_particle(r + NT);
      int ix = 2, iy = 2, iz = 2, dest = -1;
      bool in_tile = false;  // the deposit cell is one of the tile's cells
      if (active) {
        if constexpr (FX) {
          int X = __float_as_int(x), Y = __float_as_int(y), Z = __float_as_int(z);
          if constexpr (PUSH)
            push_fixed<true, TX, TY, TZ>(a.g, a.f, sp.k, a.fx, blk, lim_b, X, Y, Z, u, v, w,
                                         bad_move, n_lg, n_gg);
          // deposit cell 2 + floor (src/Mover.cpp:411-413) by a shift; a
          // cell of the tile is valid as it is, any other is clamped and
          // flagged as fx_deposit_cell does
          ix = 2 + (X >> a.fx.k[0]);
          iy = 2 + (Y >> a.fx.k[1]);
          iz = 2 + (Z >> a.fx.k[2]);
          in_tile = ((unsigned)(ix - 2 - blk.c0x) < (unsigned)TX) &
                    ((unsigned)(iy - 2 - blk.c0y) < (unsigned)TY) &
                    ((unsigned)(iz - 2 - blk.c0z) < (unsigned)TZ);
          if (!in_tile) {
            bad_dep |= bclamp(ix, a.g.nxn);
            bad_dep |= bclamp(iy, a.g.nyn);
            bad_dep |= bclamp(iz, a.g.nzn);
          }
          x = __int_as_float(X);
          y = __int_as_float(Y);
          z = __int_as_float(Z);
        } else {
          if constexpr (PUSH)
            push_particle<T, true, TX, TY, TZ>(a.g, a.f, sp.k, blk, x, y, z, u, v, w,
                                               bad_move, n_lg, n_gg);
          ix = bcell_floor<T>(x, a.g.xStart, a.g.invdx, a.g.nxn);
          iy = bcell_floor<T>(y, a.g.yStart, a.g.invdy, a.g.nyn);
          iz = bcell_floor<T>(z, a.g.zStart, a.g.invdz, a.g.nzn);
          bad_dep |= bclamp(ix, a.g.nxn);
          bad_dep |= bclamp(iy, a.g.nyn);
          bad_dep |= bclamp(iz, a.g.nzn);
          in_tile = ((unsigned)(ix - 2 - blk.c0x) < (unsigned)TX) &
                    ((unsigned)(iy - 2 - blk.c0y) < (unsigned)TY) &
                    ((unsigned)(iz - 2 - blk.c0z) < (unsigned)TZ);
        }
        dest = in_tile ? tile : tile_of<TX, TY, TZ>(a, ix, iy, iz);
      }

      // deposit: bucket (tile cell), else the list ring (halo cell,
      // outside the block, full bucket), else directly
      bool bucketed = false;
      if (active) {
        bool done = false;
        if (in_tile) {
          const int cell = ((ix - 2 - blk.c0x) * TY + (iy - 2 - blk.c0y)) * TZ + (iz - 2 - blk.c0z);
          const unsigned int sl = atomicAdd(&s_cnt[buf][cell], 1u);
          if (sl < (unsigned int)CAP) {
            const int e = cell * CS + (int)sl;
            if constexpr (FX) {
              // offsets within the (unclamped) cell: the fixed-point
              // coordinate's low bits
              s_bkt[0][e] = (float)(__float_as_int(x) & ((1 << a.fx.k[0]) - 1)) * a.fx.unit[0];
              s_bkt[1][e] = (float)(__float_as_int(y) & ((1 << a.fx.k[1]) - 1)) * a.fx.unit[1];
              s_bkt[2][e] = (float)(__float_as_int(z) & ((1 << a.fx.k[2]) - 1)) * a.fx.unit[2];
            } else {
              s_bkt[0][e] = x; s_bkt[1][e] = y; s_bkt[2][e] = z;
            }
            s_bkt[3][e] = q;
            s_bkt[4][e] = u; s_bkt[5][e] = v; s_bkt[6][e] = w;
            done = true;
            bucketed = true;
            ++n_ld;
          }
        }
        if (!done) {
          // halo cell, outside the block or full bucket: to the list
          const unsigned int k = atomicAdd(&s_nl, 1u);
          if (k - lst_done < (unsigned int)NL) {
            const unsigned int e = k % NL;
            s_lst[0][e] = x; s_lst[1][e] = y; s_lst[2][e] = z; s_lst[3][e] = q;
            s_lst[4][e] = u; s_lst[5][e] = v; s_lst[6][e] = w;
            s_lcell[e] = ((unsigned int)ix << 20) | ((unsigned int)iy << 10) | (unsigned int)iz;
            s_lid[e] = pid;
            done = true;
          }
        }
        if (!done) {
          // list full: this thread deposits all eight nodes itself
          bool lds = false;
#pragma unroll 1
          for (int nd = 0; nd < 8; ++nd) {
            if constexpr (FX)
              lds = deposit_node_fx<A, BX, BY, BZ>(
                  a.g, a.fx, sp.m, s_mom, blk.c0x, blk.c0y, blk.c0z, ix, iy, iz, nd >> 2,
                  (nd >> 1) & 1, nd & 1, __float_as_int(x), __float_as_int(y),
                  __float_as_int(z), u, v, w, q);
            else
              lds = deposit_node<T, A, BX, BY, BZ>(a.g, sp.m, s_mom, s_xn, s_yn, s_zn,
                                                   blk.c0x, blk.c0y, blk.c0z, ix, iy, iz,
                                                   nd >> 2, (nd >> 1) & 1, nd & 1, x, y, z,
                                                   u, v, w, q);
          }
          if (lds) ++n_ld;
          else ++n_gd;
          const bool stay = dest == tile;
          const unsigned int jj = stay ? atomicAdd(&s_stay, 1u) : 0u;
          bad_store |= !put_particle(a, sp, own_lo, own_cap, stay_base, stay, jj, dest, x, y,
                                     z, u, v, w, q, pid);
        }
      }

      if (LATE && step + 1 < nbatches * (uint32_t)PPT) load_particle(r + NT);
      // the bucketed particles stay in the tile: their slots come from the
      // workgroup's reserved block, one LDS counter atomic per wave
      {
        const unsigned long long m = __ballot(bucketed);
        if (m) {
          const int leader = __ffsll((long long)m) - 1;
          unsigned int jb = 0;
          if (lane == leader) jb = atomicAdd(&s_stay, (unsigned int)__popcll(m));
          jb = __shfl(jb, leader);
          if (bucketed) {
            const unsigned int jj = jb + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
            bad_store |= !put_particle(a, sp, own_lo, own_cap, stay_base, true, jj, tile, x,
                                       y, z, u, v, w, q, pid);
          }
        }
      }
      }  // particle pp of the batch

      // bucket step: each cell's four threads sum the cell's bucket
      __syncthreads();
      {
        const unsigned int nb_c = min(s_cnt[buf][my_cell], (unsigned int)CAP);
        if ((tid & 3) == 0) s_cnt[buf ^ 1][my_cell] = 0u;  // for the next batch
        const T X0 = s_xn[mlx + 1], X1 = s_xn[mlx + 2];
        const T Y0 = s_yn[mly + 1], Y1 = s_yn[mly + 2];
        const T Z0 = s_zn[mlz + 1], Z1 = s_zn[mlz + 2];
        const T iv = a.g.invVOL;
        // xa = px - X0 (a = 0) or X1 - px (a = 1), as one fused multiply-add
        // with this thread's constants (exact: a product by +-1 then one
        // rounding); with FX px is already the offset: px or 1 - px
        const T sa = my_a ? T(-1) : T(1), sb = my_b ? T(-1) : T(1);
        const T ca = FX ? (my_a ? T(1) : T(0)) : (my_a ? X1 : -X0);
        const T cb = FX ? (my_b ? T(1) : T(0)) : (my_b ? Y1 : -Y0);
        for (unsigned int k = 0; k < nb_c; ++k) {
          const int e = my_cell * CS + (int)k;
          const T px = s_bkt[0][e], py = s_bkt[1][e], pz = s_bkt[2][e];
          const T pq = s_bkt[3][e], pu = s_bkt[4][e], pv = s_bkt[5][e], pw = s_bkt[6][e];
          // px py pz: the particle's position, or with FX its offsets
          // within the cell (the distances from nodes X0 Y0 Z0 in cells)
          const T xa = fma(sa, px, ca);
          const T yb = fma(sb, py, cb);
          const T uu = pu * pu, uv = pu * pv, uw = pu * pw, vv = pv * pv, vw = pv * pw,
                  ww = pw * pw;
          if (kStrictB) {
            // the reference's terms: (p_m * (q xi eta zeta iv)) * iv
            const T w0 = pq * xa * yb * (pz - Z0) * iv;  // corner (a, b, 0)
            const T w1 = pq * xa * yb * (Z1 - pz) * iv;  // corner (a, b, 1)
            add_terms<T, T>(acc, w0, iv, pu, pv, pw, uu, uv, uw, vv, vw, ww);
            add_terms<T, T>(acc + 10, w1, iv, pu, pv, pw, uu, uv, uw, vv, vw, ww);
          } else {
            // sum of p_m * (q xi eta zeta); the factor iv^2 (with FX, whose
            // offsets are in cells, iv) is applied once when the sums are
            // added to the LDS block
            const T qxy = pq * xa * yb;
            const T w0 = qxy * (FX ? pz : pz - Z0);
            const T w1 = FX ? qxy - w0 : qxy * (Z1 - pz);
            const T pm[10] = {T(1), pu, pv, pw, uu, uv, uw, vv, vw, ww};
#pragma unroll
            for (int m = 0; m < 10; ++m) {
              acc[m] += pm[m] * w0;
              acc[10 + m] += pm[m] * w1;
            }
          }
        }
      }
      // list step: once the ring is half full (and after the species' last
      // batch), all threads deposit its pending entries, eight threads per
      // entry, one node each.  s_nl is not written again before the next
      // barrier, so every thread reads the same value and takes the same
      // decision.
      {
        const unsigned int nl = s_nl;
        const unsigned int pend = min(nl - lst_done, (unsigned int)NL);
        if (pend >= (unsigned int)(NL / 2) || (j == nbatches - 1 && pend > 0u)) {
          const int nd = tid & 7;
          for (unsigned int k = (unsigned int)(tid >> 3); k < pend; k += NT / 8) {
            const unsigned int e = (lst_done + k) % NL;
            const T px = s_lst[0][e], py = s_lst[1][e], pz = s_lst[2][e];
            const T pq = s_lst[3][e], pu = s_lst[4][e], pv = s_lst[5][e], pw = s_lst[6][e];
            const unsigned int c = s_lcell[e];
            const int jx = (int)(c >> 20), jy = (int)((c >> 10) & 1023u), jz = (int)(c & 1023u);
            bool lds;
            if constexpr (FX)
              lds = deposit_node_fx<A, BX, BY, BZ>(
                  a.g, a.fx, sp.m, s_mom, blk.c0x, blk.c0y, blk.c0z, jx, jy, jz, nd >> 2,
                  (nd >> 1) & 1, nd & 1, __float_as_int(px), __float_as_int(py),
                  __float_as_int(pz), pu, pv, pw, pq);
            else
              lds = deposit_node<T, A, BX, BY, BZ>(
                  a.g, sp.m, s_mom, s_xn, s_yn, s_zn, blk.c0x, blk.c0y, blk.c0z, jx, jy, jz,
                  nd >> 2, (nd >> 1) & 1, nd & 1, px, py, pz, pu, pv, pw, pq);
            if (nd == 0) {
              if (lds) ++n_ld;
              else ++n_gd;
              // and the entry's particle to the output buffer
              const int pdest = tile_of<TX, TY, TZ>(a, jx, jy, jz);
              const bool stay = pdest == tile;
              const unsigned int jj = stay ? atomicAdd(&s_stay, 1u) : 0u;
              bad_store |= !put_particle(a, sp, own_lo, own_cap, stay_base, stay, jj, pdest,
                                         px, py, pz, pu, pv, pw, pq, s_lid[e]);
            }
          }
          lst_done = nl;
        }
      }
      bad_bits |= (bad_move ? 1 : 0) | (bad_dep ? 2 : 0) | (bad_store ? 4 : 0);
      __syncthreads();
    }

    // 3. register sums into the LDS block: corner (a, b, c) of cell
    // (lx, ly, lz) is block node (lx + 2 - a, ly + 2 - b, lz + 2 - c)
    const A scale = kStrictB ? A(1) : FX ? (A)a.g.invVOL : (A)a.g.invVOL * (A)a.g.invVOL;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      const int nb = ((mlx + 2 - my_a) * BY + (mly + 2 - my_b)) * BZ + (mlz + 2 - cc);
#pragma unroll
      for (int m = 0; m < 10; ++m)
        if (acc[cc * 10 + m] != T(0))
          atomicAdd(&s_mom[m * NB + nb], (A)acc[cc * 10 + m] * scale);
    }
    __syncthreads();
    using P = PartialT<T>;
    // stored node-major (the ten moments of a block node together), the
    // order in which the reduce reads them; the store stays contiguous
    SPIC_GLOBAL P *dst = gmem((P *)a.partials) + ((size_t)tile * a.ns + s) * (10 * NB);
    for (int i = tid; i < 10 * NB; i += NT) {
      const int b = i / 10, m = i - 10 * b;
      dst[i] = (P)s_mom[m * NB + b];
    }
    if (tid == 0) {
      // the reserved block [stay_base, stay_base + nin) holds s_stay
      // particles; its remainder inside the segment stays unused
      const int64_t cap = gmem(sp.seg)[tile + 1] - gmem(sp.seg)[tile];
      const int64_t room = cap > (int64_t)stay_base ? cap - (int64_t)stay_base : 0;
      const int64_t lo = (int64_t)stay_base + min((int64_t)s_stay, room);
      const int64_t hi = min((int64_t)stay_base + nin, cap);
      if (hi > lo) {
        gmem(sp.cnt_out)[(a.ntiles + 1) + tile] = (unsigned int)lo;
        gmem(sp.cnt_out)[2 * (a.ntiles + 1) + tile] = (unsigned int)(hi - lo);
      }
    }
    __syncthreads();
  }
  if (bad_bits & 1) gatomic_add(a.status, 1u);
  if (bad_bits & 2) gatomic_add(a.status + 1, 1u);
  if (bad_bits & 4) gatomic_add(a.status + 2, 1u);
  if (a.counters) {
    if (n_gg) gatomic_add(a.counters + 0, (unsigned long long)n_gg);
    if (n_gd) gatomic_add(a.counters + 1, (unsigned long long)n_gd);
    if (n_lg) gatomic_add(a.counters + 2, (unsigned long long)n_lg);
    if (n_ld) gatomic_add(a.counters + 3, (unsigned long long)n_ld);
  }
}


// PUSH false: the deposit alone (spic_engine_deposit), in its own
// instances so that the phase kernels carry no test for it.  PPT 2 (two
// particles per thread and batch, FP32 fixed point only): its buckets take
// the LDS of a third workgroup, so it asks for two waves per SIMD.
template <typename T, typename A, int TX, int TY, int TZ, bool PUSH = true, bool FX = false,
          int PPT = 1>
__global__ __launch_bounds__(4 * TX * TY * TZ)
__attribute__((amdgpu_waves_per_eu(PPT == 2 ? 2 : sizeof(T) == 4 ? 3 : 2)))
void k_binned_phase(BinnedArgs<T> a) {
  binned_phase_tile<T, A, TX, TY, TZ, PUSH, FX, PPT>(a, blockIdx.x);
}

// The particles of every species' input overflow region: global gathers
// and deposits, re-binned like the tile kernel's.
// Part `part` of `nparts` (workgroups of blockDim.x threads) takes every
// (nparts * blockDim.x)-th particle.
template <typename T, int TX, int TY, int TZ, bool PUSH = true, bool FX = false>
__device__ __forceinline__ void binned_overflow_part(const BinnedArgs<T> &a, int part,
                                                     int nparts) {
  Block<T, 1, 1, 1> none;
  const int no_lim[3] = {-1, -1, -1};  // nothing is staged
  none.fld = nullptr;
  none.xn = none.yn = none.zn = nullptr;
  none.c0x = none.c0y = none.c0z = 0;
  int bad_move = 0;
  bool bad_dep = false, bad_store = false;
  unsigned int n_gg = 0, n_gd = 0, n_lg = 0;
  for (int s = 0; s < a.ns; ++s) {
    const BinnedSpecies<T> &sp = a.sp[s];
    const SlotRange ov = slot_range(a, sp, a.ntiles);
    const uint32_t nin = ov.count;
    for (uint32_t r = part * blockDim.x + threadIdx.x; r < nin;
         r += (uint32_t)nparts * blockDim.x) {
      const uint32_t i = ov.slot(r);
      const uint32_t off = i * (uint32_t)sizeof(T);
      T x = load_comp(sp, 0, off), y = load_comp(sp, 1, off), z = load_comp(sp, 2, off);
      T u = load_comp(sp, 3, off), v = load_comp(sp, 4, off), w = load_comp(sp, 5, off);
      const T q = load_comp(sp, 6, off);
      const uint32_t pid =
          sp.id_in
              ? *(const SPIC_GLOBAL uint32_t *)((const SPIC_GLOBAL char *)gmem(sp.id_in) + i * 4u)
              : 0u;
      int ix, iy, iz;
      if constexpr (FX) {
        int X = __float_as_int(x), Y = __float_as_int(y), Z = __float_as_int(z);
        if constexpr (PUSH)
          push_fixed<false, 1, 1, 1>(a.g, a.f, sp.k, a.fx, none, no_lim, X, Y, Z, u, v, w,
                                     bad_move, n_lg, n_gg);
        ix = fx_deposit_cell(X, a.fx.k[0], a.g.nxn, bad_dep);
        iy = fx_deposit_cell(Y, a.fx.k[1], a.g.nyn, bad_dep);
        iz = fx_deposit_cell(Z, a.fx.k[2], a.g.nzn, bad_dep);
        x = __int_as_float(X);
        y = __int_as_float(Y);
        z = __int_as_float(Z);
      } else {
        if constexpr (PUSH)
          push_particle<T, false, 1, 1, 1>(a.g, a.f, sp.k, none, x, y, z, u, v, w,
                                           bad_move, n_lg, n_gg);
        ix = bcell_floor<T>(x, a.g.xStart, a.g.invdx, a.g.nxn);
        iy = bcell_floor<T>(y, a.g.yStart, a.g.invdy, a.g.nyn);
        iz = bcell_floor<T>(z, a.g.zStart, a.g.invdz, a.g.nzn);
        bad_dep |= bclamp(ix, a.g.nxn);
        bad_dep |= bclamp(iy, a.g.nyn);
        bad_dep |= bclamp(iz, a.g.nzn);
      }
      const int dest = tile_of<TX, TY, TZ>(a, ix, iy, iz);
      const uint32_t o = seg_slot(a, sp, dest, gatomic_add(sp.cnt_out + dest, 1u));
      if (o != kNoSlot) write_particle(sp, o, x, y, z, u, v, w, q, pid);
      else bad_store = true;
      ++n_gd;
      if constexpr (FX)
        deposit_global_fx(a.g, a.fx, sp.m, ix, iy, iz, __float_as_int(x), __float_as_int(y),
                          __float_as_int(z), u, v, w, q);
      else
        deposit_global<T>(a.g, sp.m, ix, iy, iz, x, y, z, u, v, w, q);
    }
  }
  if (bad_move) gatomic_add(a.status, 1u);
  if (bad_dep) gatomic_add(a.status + 1, 1u);
  if (bad_store) gatomic_add(a.status + 2, 1u);
  if (a.counters) {
    if (n_gg) gatomic_add(a.counters + 0, (unsigned long long)n_gg);
    if (n_gd) gatomic_add(a.counters + 1, (unsigned long long)n_gd);
  }
}

template <typename T, int TX, int TY, int TZ, bool PUSH = true, bool FX = false>
__global__ __launch_bounds__(256) void k_binned_overflow(BinnedArgs<T> a) {
  binned_overflow_part<T, TX, TY, TZ, PUSH, FX>(a, blockIdx.x, gridDim.x);
}

// Every instance's particle work of a batch phase in one persistent
// launch: each workgroup takes the next item of the work list `items`
// (spic_binned.hpp, QueueItem) from the device counter ctr[0] until the
// list is exhausted, so the grid is sized to the workgroups the device holds
// at once and the longest items, listed first, start first.  Items are
// independent (every tile and overflow part reads the input buffers and
// writes the outputs through counters), so no order among them is needed and
// any grid size completes.  The last workgroup to leave (ctr[1]) resets both
// counters for the next launch: a replayed graph needs no reset node.
template <typename T, typename A, int TX, int TY, int TZ, bool FX = false>
__global__ __launch_bounds__(4 * TX * TY * TZ)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 2)))
void k_binned_phase_queue(const BinnedArgs<T> *__restrict__ tab,
                          const uint32_t *__restrict__ items, uint32_t nitems,
                          unsigned int *ctr) {
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = gatomic_add(ctr, 1u);
    __syncthreads();
    const uint32_t it = s_item;
    __syncthreads();  // every thread has read s_item before it is written again
    if (it >= nitems) break;
    const uint32_t w = gmem(items)[it];
    const int inst = (int)(w >> kQueueInstShift);
    const uint32_t idx = w & kQueueIndexMask;
    if (w & kQueueOverflowBit)
      binned_overflow_part<T, TX, TY, TZ, true, FX>(tab[inst], (int)idx, kQueueOverflowParts);
    else
      binned_phase_tile<T, A, TX, TY, TZ, true, FX>(tab[inst], (int)idx);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (gatomic_add(ctr + 1, 1u) == gridDim.x - 1) {
      // every workgroup has taken its last item: no counter use remains
      gmem(ctr)[0] = 0u;
      gmem(ctr)[1] = 0u;
      __threadfence();
    }
  }
}

template <typename T, typename A, int TX, int TY, int TZ, bool FX>
hipError_t launch_queue_shape(const BinnedArgs<T> *tab, const uint32_t *items,
                              uint32_t nitems, unsigned int *ctr, hipStream_t s) {
  constexpr int NT = 4 * TX * TY * TZ;
  // workgroups the device holds at once (occupancy x compute units), once
  // per kernel instance
  static int grid = 0;
  if (grid == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err == hipSuccess)
      err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (err == hipSuccess)
      err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(&k_binned_phase_queue<T, A, TX, TY, TZ, FX>),
          NT, 0);
    if (err != hipSuccess) return err;
    grid = max(1, cus * max(1, per_cu));
  }
  const int blocks = (int)min((uint32_t)grid, nitems);
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL((k_binned_phase_queue<T, A, TX, TY, TZ, FX>), dim3(blocks), dim3(NT), 0, s,
                     tab, items, nitems, ctr);
  return hipGetLastError();
}

template <typename T, typename A, int TX, int TY, int TZ>
hipError_t launch_queue_pos(const BinnedArgs<T> *tab, const uint32_t *items, uint32_t nitems,
                            unsigned int *ctr, int fixed, hipStream_t s) {
  if constexpr (sizeof(T) == 4 && !kStrictB) {
    if (fixed) return launch_queue_shape<T, A, TX, TY, TZ, true>(tab, items, nitems, ctr, s);
  } else {
    if (fixed) return hipErrorInvalidValue;
  }
  return launch_queue_shape<T, A, TX, TY, TZ, false>(tab, items, nitems, ctr, s);
}

template <typename T>
hipError_t launch_binned_phase_queue(const BinnedArgs<T> *tab, const uint32_t *items,
                                     uint32_t nitems, unsigned int *ctr, int tx, int ty,
                                     int tz, int lds_fp64, int fixed, hipStream_t s) {
  if (nitems == 0) return hipSuccess;
  const bool wide = sizeof(T) == 4 && lds_fp64;
  if (tx == 4 && ty == 4 && tz == 4)
    return wide ? launch_queue_pos<T, double, 4, 4, 4>(tab, items, nitems, ctr, fixed, s)
                : launch_queue_pos<T, T, 4, 4, 4>(tab, items, nitems, ctr, fixed, s);
  if (tx == 8 && ty == 4 && tz == 4)
    return wide ? launch_queue_pos<T, double, 8, 4, 4>(tab, items, nitems, ctr, fixed, s)
                : launch_queue_pos<T, T, 8, 4, 4>(tab, items, nitems, ctr, fixed, s);
  return hipErrorInvalidValue;
}

template hipError_t launch_binned_phase_queue<double>(const BinnedArgs<double> *,
                                                      const uint32_t *, uint32_t,
                                                      unsigned int *, int, int, int, int, int,
                                                      hipStream_t);
template hipError_t launch_binned_phase_queue<float>(const BinnedArgs<float> *,
                                                     const uint32_t *, uint32_t,
                                                     unsigned int *, int, int, int, int, int,
                                                     hipStream_t);

template <typename T, typename A, bool PUSH, bool FX>
hipError_t launch_binned_shape(const BinnedArgs<T> &a, hipStream_t s) {
  // two particles per thread and batch: the FP32 fixed-point phase only
  constexpr bool kPpt2 = sizeof(T) == 4 && FX && PUSH;
  const bool ppt2 = PUSH && a.ppt == 2;  // the deposit alone keeps one
  if (ppt2 && !kPpt2) return hipErrorInvalidValue;
  if (a.tx == 4 && a.ty == 4 && a.tz == 4) {
    if constexpr (kPpt2) {
      if (ppt2)
        hipLaunchKernelGGL((k_binned_phase<T, A, 4, 4, 4, PUSH, FX, 2>), dim3(a.ntiles),
                           dim3(256), 0, s, a);
    }
    if (!ppt2)
      hipLaunchKernelGGL((k_binned_phase<T, A, 4, 4, 4, PUSH, FX>), dim3(a.ntiles), dim3(256),
                         0, s, a);
    hipLaunchKernelGGL((k_binned_overflow<T, 4, 4, 4, PUSH, FX>), dim3(a.n_ovf_blocks),
                       dim3(256), 0, s, a);
  } else if (a.tx == 8 && a.ty == 4 && a.tz == 4) {
    if constexpr (kPpt2) {
      if (ppt2)
        hipLaunchKernelGGL((k_binned_phase<T, A, 8, 4, 4, PUSH, FX, 2>), dim3(a.ntiles),
                           dim3(512), 0, s, a);
    }
    if (!ppt2)
      hipLaunchKernelGGL((k_binned_phase<T, A, 8, 4, 4, PUSH, FX>), dim3(a.ntiles), dim3(512),
                         0, s, a);
    hipLaunchKernelGGL((k_binned_overflow<T, 8, 4, 4, PUSH, FX>), dim3(a.n_ovf_blocks),
                       dim3(256), 0, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, typename A, bool PUSH>
hipError_t launch_binned_pos(const BinnedArgs<T> &a, hipStream_t s) {
  if constexpr (sizeof(T) == 4 && !kStrictB) {
    if (a.fixed) return launch_binned_shape<T, A, PUSH, true>(a, s);
  } else {
    // fixed-point positions exist only in FP32 production arithmetic; the
    // engine refuses the other combinations before launching
    if (a.fixed) return hipErrorInvalidValue;
  }
  return launch_binned_shape<T, A, PUSH, false>(a, s);
}

template <typename T>
hipError_t launch_binned_phase(const BinnedArgs<T> &a, hipStream_t s) {
  if (a.ntiles <= 0 || a.ns <= 0) return hipSuccess;
  const bool lds64 = sizeof(T) == 4 && a.lds_fp64;
  if (a.push_off)
    return lds64 ? launch_binned_pos<T, double, false>(a, s)
                 : launch_binned_pos<T, T, false>(a, s);
  return lds64 ? launch_binned_pos<T, double, true>(a, s)
               : launch_binned_pos<T, T, true>(a, s);
}

template hipError_t launch_binned_phase<double>(const BinnedArgs<double> &, hipStream_t);
template hipError_t launch_binned_phase<float>(const BinnedArgs<float> &, hipStream_t);

}  // namespace SPIC_NS
}  // namespace spic
