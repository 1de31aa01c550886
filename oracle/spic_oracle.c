/*
 * spic_oracle.c -- CPU oracle for the sputniPIC particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle, never the product:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load the library built from it.  The product path (sputnipic_amd/)
 * never links or calls it.
 *
 * It restates, operation for operation and type for type, the reference's
 * CPU algorithms (paths relative to the reference repository root):
 *   oracle_mover_pc      src/Mover.cpp:4-162      (mover_PC)
 *   oracle_interp_p2g    src/Mover.cpp:385-572    (interpP2G)
 *   oracle_mover_interp  src/Mover.cpp:575-862    (mover_interp, fused)
 *   oracle_particle_sort src/Particles.cpp:67-234 (particle_sort, binned branch)
 *   oracle_set_grid      src/Grid.cpp:5-62        (setGrid)
 *   oracle_init_gem      src/IC.cpp:7-177 + src/InterpDensSpecies.cpp:64-77
 *                        (initGEM fields/densities/particles, interpN2Crho)
 *   oracle_apply_bc_ids  src/BC.cpp:362-619       (applyBCids)
 *   oracle_sum_species   src/InterpDensNet.cpp:120-144 (sumOverSpecies)
 *   oracle_bc_scalar     src/BC.cpp:11-123, 242-352, 623-735, 737-850
 *                        (applyBCscalarDensN / FieldNzero / DensC / FieldC)
 *   oracle_rhoc_from_rhon  src/sputniPIC_CPU.cpp:276-277 (applyBCscalarDensN
 *                        on the net rhon, interpN2Cinterp src/Grid.cpp:164-185)
 *   oracle_hat_densities src/MaxwellSolver.cpp:465-604 (calculateHatDensities)
 *   oracle_maxwell_image src/MaxwellSolver.cpp:95-257  (MaxwellImage)
 *
 * The file is compiled twice.  With -DSPIC_ORACLE_FP64 the three precision
 * typedefs are double (the reference built with include/PrecisionTypes.h set
 * to double); without it they are float, the reference's default
 * (include/PrecisionTypes.h:6-10).  Every expression keeps the reference's
 * operand types, so the usual arithmetic conversions (float*float in float,
 * float op double in double) round exactly as the reference does when both
 * are built without FMA contraction (-ffp-contract=off, no -ffast-math).
 *
 * Parity pinning of each function (what checks it) is listed in the
 * module docstring of oracle/__init__.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef SPIC_ORACLE_FP64
typedef double FPpart;
typedef double FPfield;
typedef double FPinterp;
#define OSFX(name) name##_f64
#else
typedef float FPpart;
typedef float FPfield;
typedef float FPinterp;
#define OSFX(name) name##_f32
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Scalar members of the reference's struct grid (include/Grid.h:11-63).
 * invd* and invVOL hold FPfield values; they are re-narrowed to FPfield
 * before use so arithmetic matches the reference's operand types. */
typedef struct oracle_grid {
  int nxc, nyc, nzc; /* guard-inclusive cell counts (Grid.cpp:9-11) */
  int nxn, nyn, nzn; /* nxc + 1 (Grid.cpp:13-15) */
  double dx, dy, dz;
  double invdx, invdy, invdz, invVOL;
  double xStart, yStart, zStart;
  double Lx, Ly, Lz;
} oracle_grid;

#define IDX3(i, j, k, ny, nz) (((long)(i) * (ny) + (j)) * (long)(nz) + (k))

/* ------------------------------------------------------------------ */
/* setGrid (src/Grid.cpp:5-62): constants and the three node arrays. */
/* XN, YN, ZN are flat [nxn][nyn][nzn] arrays, k fastest.              */
/* ------------------------------------------------------------------ */
void OSFX(oracle_set_grid)(int nxc_phys, int nyc_phys, int nzc_phys, double Lx,
                           double Ly, double Lz, oracle_grid *g, FPfield *XN,
                           FPfield *YN, FPfield *ZN) {
  g->nxc = nxc_phys + 2;
  g->nyc = nyc_phys + 2;
  g->nzc = nzc_phys + 2;
  g->nxn = g->nxc + 1;
  g->nyn = g->nyc + 1;
  g->nzn = g->nzc + 1;
  g->dx = Lx / nxc_phys;
  g->dy = Ly / nyc_phys;
  g->dz = Lz / nzc_phys;
  /* (FPfield)1.0 / (double) -> double quotient narrowed to FPfield */
  g->invVOL = (double)(FPfield)((FPfield)1.0 / (g->dx * g->dy * g->dz));
  g->invdx = (double)(FPfield)((FPfield)1.0 / g->dx);
  g->invdy = (double)(FPfield)((FPfield)1.0 / g->dy);
  g->invdz = (double)(FPfield)((FPfield)1.0 / g->dz);
  g->xStart = 0.0;
  g->yStart = 0.0;
  g->zStart = 0.0;
  g->Lx = Lx;
  g->Ly = Ly;
  g->Lz = Lz;
  if (XN && YN && ZN) {
    for (int i = 0; i < g->nxn; i++)
      for (int j = 0; j < g->nyn; j++)
        for (int k = 0; k < g->nzn; k++) {
          long n = IDX3(i, j, k, g->nyn, g->nzn);
          XN[n] = (FPfield)(g->xStart + (i - 1) * g->dx);
          YN[n] = (FPfield)(g->yStart + (j - 1) * g->dy);
          ZN[n] = (FPfield)(g->zStart + (k - 1) * g->dz);
        }
  }
}

/* ------------------------------------------------------------------ */
/* mover_PC (src/Mover.cpp:4-162).                                     */
/* Returns 0, as the reference does (Mover.cpp:161).  The oracle adds  */
/* no bounds checks: inputs must keep every iterate within one cell of */
/* the domain, as the reference silently requires.                    */
/* ------------------------------------------------------------------ */
typedef struct oracle_fields {
  const FPfield *Ex, *Ey, *Ez, *Bxn, *Byn, *Bzn;
} oracle_fields;

typedef struct oracle_nodes {
  const FPfield *XN, *YN, *ZN;
} oracle_nodes;

typedef struct oracle_particles {
  FPpart *x, *y, *z, *u, *v, *w;
  FPinterp *q;
  long nop;
  long npmax;
  double qom; /* narrowed to FPpart as particle_allocate does (Particles.cpp:25) */
  int NiterMover;
  int n_sub_cycles;
} oracle_particles;

typedef struct oracle_moments {
  FPinterp *rhon, *Jx, *Jy, *Jz, *pxx, *pxy, *pxz, *pyy, *pyz, *pzz;
} oracle_moments;

typedef struct oracle_push {
  double dt, c;
  int PERIODICX, PERIODICY, PERIODICZ;
} oracle_push;

/* One particle's predictor-corrector push + particle BC; the body of the
 * reference loop at Mover.cpp:33-157 (also :602-726 in mover_interp). */
static inline void push_one(const oracle_grid *grd, const oracle_nodes *nd,
                            const oracle_fields *field, oracle_particles *part,
                            const oracle_push *param, long i,
                            FPpart dt_sub_cycling, FPpart dto2, FPpart qomdt2) {
  const FPfield invdx = (FPfield)grd->invdx, invdy = (FPfield)grd->invdy,
                invdz = (FPfield)grd->invdz, invVOL = (FPfield)grd->invVOL;
  const int nyn = grd->nyn, nzn = grd->nzn;
  FPpart omdtsq, denom, ut, vt, wt, udotb;
  FPfield Exl, Eyl, Ezl, Bxl, Byl, Bzl;
  int ix, iy, iz;
  FPfield weight[2][2][2];
  FPfield xi[2], eta[2], zeta[2];
  FPpart xptilde, yptilde, zptilde, uptilde = 0, vptilde = 0, wptilde = 0;

  xptilde = part->x[i];
  yptilde = part->y[i];
  zptilde = part->z[i];
  for (int innter = 0; innter < part->NiterMover; innter++) {
    ix = 2 + (int)((part->x[i] - grd->xStart) * invdx);
    iy = 2 + (int)((part->y[i] - grd->yStart) * invdy);
    iz = 2 + (int)((part->z[i] - grd->zStart) * invdz);

    xi[0] = part->x[i] - nd->XN[IDX3(ix - 1, iy, iz, nyn, nzn)];
    eta[0] = part->y[i] - nd->YN[IDX3(ix, iy - 1, iz, nyn, nzn)];
    zeta[0] = part->z[i] - nd->ZN[IDX3(ix, iy, iz - 1, nyn, nzn)];
    xi[1] = nd->XN[IDX3(ix, iy, iz, nyn, nzn)] - part->x[i];
    eta[1] = nd->YN[IDX3(ix, iy, iz, nyn, nzn)] - part->y[i];
    zeta[1] = nd->ZN[IDX3(ix, iy, iz, nyn, nzn)] - part->z[i];
    for (int ii = 0; ii < 2; ii++)
      for (int jj = 0; jj < 2; jj++)
        for (int kk = 0; kk < 2; kk++)
          weight[ii][jj][kk] = xi[ii] * eta[jj] * zeta[kk] * invVOL;

    Exl = 0.0, Eyl = 0.0, Ezl = 0.0, Bxl = 0.0, Byl = 0.0, Bzl = 0.0;
    for (int ii = 0; ii < 2; ii++)
      for (int jj = 0; jj < 2; jj++)
        for (int kk = 0; kk < 2; kk++) {
          long n = IDX3(ix - ii, iy - jj, iz - kk, nyn, nzn);
          Exl += weight[ii][jj][kk] * field->Ex[n];
          Eyl += weight[ii][jj][kk] * field->Ey[n];
          Ezl += weight[ii][jj][kk] * field->Ez[n];
          Bxl += weight[ii][jj][kk] * field->Bxn[n];
          Byl += weight[ii][jj][kk] * field->Byn[n];
          Bzl += weight[ii][jj][kk] * field->Bzn[n];
        }

    omdtsq = qomdt2 * qomdt2 * (Bxl * Bxl + Byl * Byl + Bzl * Bzl);
    denom = 1.0 / (1.0 + omdtsq);
    ut = part->u[i] + qomdt2 * Exl;
    vt = part->v[i] + qomdt2 * Eyl;
    wt = part->w[i] + qomdt2 * Ezl;
    udotb = ut * Bxl + vt * Byl + wt * Bzl;
    uptilde = (ut + qomdt2 * (vt * Bzl - wt * Byl + qomdt2 * udotb * Bxl)) * denom;
    vptilde = (vt + qomdt2 * (wt * Bxl - ut * Bzl + qomdt2 * udotb * Byl)) * denom;
    wptilde = (wt + qomdt2 * (ut * Byl - vt * Bxl + qomdt2 * udotb * Bzl)) * denom;
    part->x[i] = xptilde + uptilde * dto2;
    part->y[i] = yptilde + vptilde * dto2;
    part->z[i] = zptilde + wptilde * dto2;
  }
  part->u[i] = 2.0 * uptilde - part->u[i];
  part->v[i] = 2.0 * vptilde - part->v[i];
  part->w[i] = 2.0 * wptilde - part->w[i];
  part->x[i] = xptilde + uptilde * dt_sub_cycling;
  part->y[i] = yptilde + vptilde * dt_sub_cycling;
  part->z[i] = zptilde + wptilde * dt_sub_cycling;

  /* particle BC: two independent tests per axis (Mover.cpp:103-157) */
  if (part->x[i] > grd->Lx) {
    if (param->PERIODICX) {
      part->x[i] = part->x[i] - grd->Lx;
    } else {
      part->u[i] = -part->u[i];
      part->x[i] = 2 * grd->Lx - part->x[i];
    }
  }
  if (part->x[i] < 0) {
    if (param->PERIODICX) {
      part->x[i] = part->x[i] + grd->Lx;
    } else {
      part->u[i] = -part->u[i];
      part->x[i] = -part->x[i];
    }
  }
  if (part->y[i] > grd->Ly) {
    if (param->PERIODICY) {
      part->y[i] = part->y[i] - grd->Ly;
    } else {
      part->v[i] = -part->v[i];
      part->y[i] = 2 * grd->Ly - part->y[i];
    }
  }
  if (part->y[i] < 0) {
    if (param->PERIODICY) {
      part->y[i] = part->y[i] + grd->Ly;
    } else {
      part->v[i] = -part->v[i];
      part->y[i] = -part->y[i];
    }
  }
  if (part->z[i] > grd->Lz) {
    if (param->PERIODICZ) {
      part->z[i] = part->z[i] - grd->Lz;
    } else {
      part->w[i] = -part->w[i];
      part->z[i] = 2 * grd->Lz - part->z[i];
    }
  }
  if (part->z[i] < 0) {
    if (param->PERIODICZ) {
      part->z[i] = part->z[i] + grd->Lz;
    } else {
      part->w[i] = -part->w[i];
      part->z[i] = -part->z[i];
    }
  }
}

/* Deposit of one particle: the body of the reference loop at
 * Mover.cpp:397-571 (also :729-855 in mover_interp). */
static inline void deposit_one(const oracle_grid *grd, const oracle_nodes *nd,
                               const oracle_particles *part,
                               oracle_moments *ids, long i) {
  const FPfield invdx = (FPfield)grd->invdx, invdy = (FPfield)grd->invdy,
                invdz = (FPfield)grd->invdz, invVOL = (FPfield)grd->invVOL;
  const int nyn = grd->nyn, nzn = grd->nzn;
  FPpart weight[2][2][2];
  FPpart temp[2][2][2];
  FPpart xi[2], eta[2], zeta[2];
  int ix, iy, iz;

  ix = 2 + (int)(floor((part->x[i] - grd->xStart) * invdx));
  iy = 2 + (int)(floor((part->y[i] - grd->yStart) * invdy));
  iz = 2 + (int)(floor((part->z[i] - grd->zStart) * invdz));

  xi[0] = part->x[i] - nd->XN[IDX3(ix - 1, iy, iz, nyn, nzn)];
  eta[0] = part->y[i] - nd->YN[IDX3(ix, iy - 1, iz, nyn, nzn)];
  zeta[0] = part->z[i] - nd->ZN[IDX3(ix, iy, iz - 1, nyn, nzn)];
  xi[1] = nd->XN[IDX3(ix, iy, iz, nyn, nzn)] - part->x[i];
  eta[1] = nd->YN[IDX3(ix, iy, iz, nyn, nzn)] - part->y[i];
  zeta[1] = nd->ZN[IDX3(ix, iy, iz, nyn, nzn)] - part->z[i];

  for (int ii = 0; ii < 2; ii++)
    for (int jj = 0; jj < 2; jj++)
      for (int kk = 0; kk < 2; kk++)
        weight[ii][jj][kk] = part->q[i] * xi[ii] * eta[jj] * zeta[kk] * invVOL;

#define DEPOSIT_LOOP(ARR, EXPR)                                            \
  for (int ii = 0; ii < 2; ii++)                                           \
    for (int jj = 0; jj < 2; jj++)                                         \
      for (int kk = 0; kk < 2; kk++) temp[ii][jj][kk] = (EXPR);            \
  for (int ii = 0; ii < 2; ii++)                                           \
    for (int jj = 0; jj < 2; jj++)                                         \
      for (int kk = 0; kk < 2; kk++)                                       \
        ids->ARR[IDX3(ix - ii, iy - jj, iz - kk, nyn, nzn)] +=             \
            temp[ii][jj][kk] * invVOL;

  for (int ii = 0; ii < 2; ii++)
    for (int jj = 0; jj < 2; jj++)
      for (int kk = 0; kk < 2; kk++)
        ids->rhon[IDX3(ix - ii, iy - jj, iz - kk, nyn, nzn)] +=
            weight[ii][jj][kk] * invVOL;
  DEPOSIT_LOOP(Jx, part->u[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(Jy, part->v[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(Jz, part->w[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(pxx, part->u[i] * part->u[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(pxy, part->u[i] * part->v[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(pxz, part->u[i] * part->w[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(pyy, part->v[i] * part->v[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(pyz, part->v[i] * part->w[i] * weight[ii][jj][kk])
  DEPOSIT_LOOP(pzz, part->w[i] * part->w[i] * weight[ii][jj][kk])
#undef DEPOSIT_LOOP
}

/* mover_PC: omp parallel for over particles with an int index, as at
 * Mover.cpp:27-31. */
int OSFX(oracle_mover_pc)(const oracle_grid *grd, const oracle_nodes *nd,
                          const oracle_fields *field, oracle_particles *part,
                          const oracle_push *param) {
  const FPpart qom = (FPpart)part->qom;
  FPpart dt_sub_cycling = (FPpart)param->dt / ((double)part->n_sub_cycles);
  FPpart dto2 = .5 * dt_sub_cycling, qomdt2 = qom * dto2 / param->c;
  const int nop = (int)part->nop;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nop; i++)
    push_one(grd, nd, field, part, param, i, dt_sub_cycling, dto2, qomdt2);
  return 0;
}

/* interpP2G: serial over particles, in particle order (Mover.cpp:395-397:
 * the omp pragma is commented out in the reference). */
void OSFX(oracle_interp_p2g)(const oracle_grid *grd, const oracle_nodes *nd,
                             const oracle_particles *part, oracle_moments *ids) {
  for (long i = 0; i < part->nop; i++) deposit_one(grd, nd, part, ids, i);
}

/* mover_interp: serial fused push + deposit (Mover.cpp:575-862). */
int OSFX(oracle_mover_interp)(const oracle_grid *grd, const oracle_nodes *nd,
                              const oracle_fields *field,
                              oracle_particles *part, oracle_moments *ids,
                              const oracle_push *param) {
  const FPpart qom = (FPpart)part->qom;
  FPpart dt_sub_cycling = (FPpart)param->dt / ((double)part->n_sub_cycles);
  FPpart dto2 = .5 * dt_sub_cycling, qomdt2 = qom * dto2 / param->c;
  for (int i = 0; i < part->nop; i++) {
    push_one(grd, nd, field, part, param, i, dt_sub_cycling, dto2, qomdt2);
    deposit_one(grd, nd, part, ids, i);
  }
  return 0;
}

/* Zero the ten node moments (InterpDensNet.cpp:40-75, rhon..pzz part). */
void OSFX(oracle_zero_moments)(const oracle_grid *grd, oracle_moments *ids) {
  long nn = (long)grd->nxn * grd->nyn * grd->nzn;
  FPinterp *arr[10] = {ids->rhon, ids->Jx,  ids->Jy,  ids->Jz,  ids->pxx,
                       ids->pxy,  ids->pxz, ids->pyy, ids->pyz, ids->pzz};
  for (int m = 0; m < 10; m++)
    for (long n = 0; n < nn; n++) arr[m][n] = 0.0;
}

/* ------------------------------------------------------------------ */
/* particle_sort, binned branch (Particles.cpp:155-233): stable bin    */
/* sort over npmax with a trailing "lost" bin.  Returns 0, or -1 when  */
/* scratch allocation fails.                                           */
/* ------------------------------------------------------------------ */
int OSFX(oracle_particle_sort)(const oracle_grid *grd, long sort_cps,
                               oracle_particles *part) {
  long n_binx = 1 + ((grd->nxc - 1) / sort_cps);
  long n_biny = 1 + ((grd->nyc - 1) / sort_cps);
  long n_binz = 1 + ((grd->nzc - 1) / sort_cps);
  FPpart dist_binx = grd->Lx / (FPpart)n_binx;
  FPpart dist_biny = grd->Ly / (FPpart)n_biny;
  FPpart dist_binz = grd->Lz / (FPpart)n_binz;
  long bins_tot = n_binx * n_biny * n_binz;
  long np = part->npmax;

  long *bin = (long *)malloc(sizeof(long) * (size_t)(np > 0 ? np : 1));
  long *start = (long *)calloc((size_t)(bins_tot + 2), sizeof(long));
  FPpart *tmp = (FPpart *)malloc(sizeof(FPpart) * (size_t)(np > 0 ? np : 1));
  FPinterp *tmpq = (FPinterp *)malloc(sizeof(FPinterp) * (size_t)(np > 0 ? np : 1));
  long *perm = (long *)malloc(sizeof(long) * (size_t)(np > 0 ? np : 1));
  if (!bin || !start || !tmp || !tmpq || !perm) {
    free(bin); free(start); free(tmp); free(tmpq); free(perm);
    return -1;
  }
  for (long i = 0; i < np; i++) {
    /* FPpart / FPpart, then floor: the reference calls floor on the FPpart
     * quotient (Particles.cpp:180-182); floor is exact in either width. */
    FPpart qx = part->x[i] / dist_binx;
    FPpart qy = part->y[i] / dist_biny;
    FPpart qz = part->z[i] / dist_binz;
    long pos_x = (long)floor((double)qx);
    long pos_y = (long)floor((double)qy);
    long pos_z = (long)floor((double)qz);
    if (pos_x < 0 || pos_x >= n_binx || pos_y < 0 || pos_y >= n_biny ||
        pos_z < 0 || pos_z >= n_binz)
      bin[i] = bins_tot; /* lost bin, placed last */
    else
      bin[i] = pos_z + pos_y * n_binz + pos_x * n_binz * n_biny;
    start[bin[i] + 1]++;
  }
  for (long b = 0; b <= bins_tot; b++) start[b + 1] += start[b];
  /* stable placement: particles of one bin keep their index order, which
   * is what push_back in index order followed by bin-major concatenation
   * produces (Particles.cpp:178-216). */
  for (long i = 0; i < np; i++) perm[start[bin[i]]++] = i;

  FPpart *arrs[6] = {part->x, part->y, part->z, part->u, part->v, part->w};
  for (int a = 0; a < 6; a++) {
    for (long p = 0; p < np; p++) tmp[p] = arrs[a][perm[p]];
    memcpy(arrs[a], tmp, sizeof(FPpart) * (size_t)np);
  }
  for (long p = 0; p < np; p++) tmpq[p] = part->q[perm[p]];
  memcpy(part->q, tmpq, sizeof(FPinterp) * (size_t)np);
  free(bin); free(start); free(tmp); free(tmpq); free(perm);
  return 0;
}

/* ------------------------------------------------------------------ */
/* initGEM (src/IC.cpp:7-177), non-restart branch, plus interpN2Crho   */
/* (src/InterpDensSpecies.cpp:64-77).                                  */
/* Inputs mirror struct parameters members; arrays are flat, k fastest.*/
/* rhon/rhoc: ns arrays each ([nxn][nyn][nzn] and [nxc][nyc][nzc]).    */
/* Particles are written up to max_per_species each (<= 0: all); the   */
/* lattice and rand() order are those of the reference, so a truncated */
/* run produces a prefix of the full one.  Returns 0, or 1 when a      */
/* velocity exceeds c (the reference exits at IC.cpp:151-172).         */
/* ------------------------------------------------------------------ */
typedef struct oracle_gem {
  int ns;
  double B0x, B0y, B0z, delta, fourpi, c;
  double rhoINIT[16];
  double qom[16], u0[16], v0[16], w0[16], uth[16], vth[16], wth[16];
  int npcelx[16], npcely[16], npcelz[16];
  unsigned int seed_base; /* srand(seed_base + is); the reference uses 0 */
  double pertX;           /* 0.4 in the reference (IC.cpp:11) */
} oracle_gem;

void OSFX(oracle_gem_fields)(const oracle_grid *grd, const FPfield *XN,
                             const FPfield *YN, const oracle_gem *p,
                             FPfield *Ex, FPfield *Ey, FPfield *Ez,
                             FPfield *Bxn, FPfield *Byn, FPfield *Bzn,
                             FPinterp **rhon, FPinterp **rhoc) {
  const double pertX = p->pertX;
  const int nxn = grd->nxn, nyn = grd->nyn, nzn = grd->nzn;
  for (int i = 0; i < nxn; i++)
    for (int j = 0; j < nyn; j++)
      for (int k = 0; k < nzn; k++) {
        long n = IDX3(i, j, k, nyn, nzn);
        for (int is = 0; is < p->ns; is++) {
          if (is < 2)
            rhon[is][n] = (FPinterp)((FPinterp)((p->rhoINIT[is] /
                                      (cosh((YN[n] - grd->Ly / 2) / p->delta) *
                                       cosh((YN[n] - grd->Ly / 2) / p->delta)))) /
                          p->fourpi);
          else
            rhon[is][n] = (FPinterp)((FPinterp)p->rhoINIT[is] / p->fourpi);
        }
        Ex[n] = 0.0;
        Ey[n] = 0.0;
        Ez[n] = 0.0;
        Bxn[n] = (FPfield)(p->B0x * tanh((YN[n] - grd->Ly / 2) / p->delta));
        Byn[n] = (FPfield)p->B0y;
        double xpert = XN[n] - grd->Lx / 2;
        double ypert = YN[n] - grd->Ly / 2;
        double exp_pert = exp(-(xpert / p->delta) * (xpert / p->delta) -
                              (ypert / p->delta) * (ypert / p->delta));
        Bxn[n] = (FPfield)(Bxn[n] +
                 (p->B0x * pertX) * exp_pert *
                     (-cos(M_PI * xpert / 10.0 / p->delta) *
                          cos(M_PI * ypert / 10.0 / p->delta) * 2.0 * ypert /
                          p->delta -
                      cos(M_PI * xpert / 10.0 / p->delta) *
                          sin(M_PI * ypert / 10.0 / p->delta) * M_PI / 10.0));
        Byn[n] = (FPfield)(Byn[n] +
                 (p->B0x * pertX) * exp_pert *
                     (cos(M_PI * xpert / 10.0 / p->delta) *
                          cos(M_PI * ypert / 10.0 / p->delta) * 2.0 * xpert /
                          p->delta +
                      sin(M_PI * xpert / 10.0 / p->delta) *
                          cos(M_PI * ypert / 10.0 / p->delta) * M_PI / 10.0));
        Bzn[n] = (FPfield)p->B0z;
      }
  /* interpN2Crho, interior cells only (InterpDensSpecies.cpp:64-77);
   * guard cells of rhoc are left as the caller initialised them. */
  for (int is = 0; is < p->ns; is++)
    for (int i = 1; i < grd->nxc - 1; i++)
      for (int j = 1; j < grd->nyc - 1; j++)
        for (int k = 1; k < grd->nzc - 1; k++) {
          const FPinterp *r = rhon[is];
          FPinterp s = r[IDX3(i, j, k, nyn, nzn)] + r[IDX3(i + 1, j, k, nyn, nzn)];
          s = s + r[IDX3(i, j + 1, k, nyn, nzn)];
          s = s + r[IDX3(i, j, k + 1, nyn, nzn)];
          s = s + r[IDX3(i + 1, j + 1, k, nyn, nzn)];
          s = s + r[IDX3(i + 1, j, k + 1, nyn, nzn)];
          s = s + r[IDX3(i, j + 1, k + 1, nyn, nzn)];
          s = s + r[IDX3(i + 1, j + 1, k + 1, nyn, nzn)];
          rhoc[is][IDX3(i, j, k, grd->nyc, grd->nzc)] = (FPinterp)(.125 * s);
        }
}

/* Particles of one species (IC.cpp:117-177).  Returns the number of
 * particles written, or -1 when a velocity exceeds c. */
long OSFX(oracle_gem_particles)(const oracle_grid *grd, const FPfield *XN,
                                const FPfield *YN, const FPfield *ZN,
                                const oracle_gem *p, int is,
                                const FPinterp *rhoc_is, long max_particles,
                                FPpart *x, FPpart *y, FPpart *z, FPpart *u,
                                FPpart *v, FPpart *w, FPinterp *q) {
  const int nyn = grd->nyn, nzn = grd->nzn;
  const int npcelx = p->npcelx[is], npcely = p->npcely[is], npcelz = p->npcelz[is];
  const int npcel = npcelx * npcely * npcelz;
  const FPpart qom = (FPpart)p->qom[is];
  const FPpart u0 = (FPpart)p->u0[is], v0 = (FPpart)p->v0[is], w0 = (FPpart)p->w0[is];
  const FPpart uth = (FPpart)p->uth[is], vth = (FPpart)p->vth[is], wth = (FPpart)p->wth[is];
  const FPfield invVOL = (FPfield)grd->invVOL;
  double harvest, prob, theta;
  long counter = 0;
  srand(p->seed_base + (unsigned int)is);
  for (int i = 1; i < grd->nxc - 1; i++)
    for (int j = 1; j < grd->nyc - 1; j++)
      for (int k = 1; k < grd->nzc - 1; k++)
        for (int ii = 0; ii < npcelx; ii++)
          for (int jj = 0; jj < npcely; jj++)
            for (int kk = 0; kk < npcely; kk++) { /* npcely bound: IC.cpp:128 */
              if (max_particles > 0 && counter >= max_particles) return counter;
              long n = IDX3(i, j, k, nyn, nzn);
              x[counter] = (FPpart)((ii + .5) * (grd->dx / npcelx) + XN[n]);
              y[counter] = (FPpart)((jj + .5) * (grd->dy / npcely) + YN[n]);
              z[counter] = (FPpart)((kk + .5) * (grd->dz / npcelz) + ZN[n]);
              {
                /* (qom/|qom|) in FPpart, times rhoc/npcel in FPinterp, then a
                 * double factor (IC.cpp:138-140) */
                FPpart sgn = qom / (FPpart)fabs((double)qom);
                FPinterp rq = rhoc_is[IDX3(i, j, k, grd->nyc, grd->nzc)] / npcel;
                q[counter] = (FPinterp)((sgn * rq) * (1.0 / invVOL));
              }
              harvest = rand() / (double)RAND_MAX;
              prob = sqrt(-2.0 * log(1.0 - .999999 * harvest));
              harvest = rand() / (double)RAND_MAX;
              theta = 2.0 * M_PI * harvest;
              u[counter] = (FPpart)(u0 + uth * prob * cos(theta));
              if (u[counter] > p->c) return -1;
              harvest = rand() / (double)RAND_MAX;
              prob = sqrt(-2.0 * log(1.0 - .999999 * harvest));
              harvest = rand() / (double)RAND_MAX;
              theta = 2.0 * M_PI * harvest;
              w[counter] = (FPpart)(w0 + wth * prob * cos(theta));
              v[counter] = (FPpart)(v0 + vth * prob * sin(theta));
              if (v[counter] > p->c) return -1;
              if (w[counter] > p->c) return -1;
              counter++;
            }
  return counter;
}

/* ------------------------------------------------------------------ */
/* applyBCids (src/BC.cpp:362-619): fold periodic duplicate planes in  */
/* x, then y, then z; then double the boundary planes of every         */
/* non-periodic axis, in x, y, z order.                                */
/* ------------------------------------------------------------------ */
void OSFX(oracle_apply_bc_ids)(const oracle_grid *grd, oracle_moments *ids,
                               int PERIODICX, int PERIODICY, int PERIODICZ) {
  const int nxn = grd->nxn, nyn = grd->nyn, nzn = grd->nzn;
  FPinterp *arr[10] = {ids->rhon, ids->Jx,  ids->Jy,  ids->Jz,  ids->pxx,
                       ids->pxy,  ids->pxz, ids->pyy, ids->pyz, ids->pzz};
  if (PERIODICX)
    for (int j = 1; j < nyn - 1; j++)
      for (int k = 1; k < nzn - 1; k++)
        for (int m = 0; m < 10; m++) {
          FPinterp *a = arr[m];
          a[IDX3(1, j, k, nyn, nzn)] += a[IDX3(nxn - 2, j, k, nyn, nzn)];
          a[IDX3(nxn - 2, j, k, nyn, nzn)] = a[IDX3(1, j, k, nyn, nzn)];
        }
  if (PERIODICY)
    for (int i = 1; i < nxn - 1; i++)
      for (int k = 1; k < nzn - 1; k++)
        for (int m = 0; m < 10; m++) {
          FPinterp *a = arr[m];
          a[IDX3(i, 1, k, nyn, nzn)] += a[IDX3(i, nyn - 2, k, nyn, nzn)];
          a[IDX3(i, nyn - 2, k, nyn, nzn)] = a[IDX3(i, 1, k, nyn, nzn)];
        }
  if (PERIODICZ)
    for (int i = 1; i < nxn - 1; i++)
      for (int j = 1; j < nyn - 1; j++)
        for (int m = 0; m < 10; m++) {
          FPinterp *a = arr[m];
          a[IDX3(i, j, 1, nyn, nzn)] += a[IDX3(i, j, nzn - 2, nyn, nzn)];
          a[IDX3(i, j, nzn - 2, nyn, nzn)] = a[IDX3(i, j, 1, nyn, nzn)];
        }
  if (!PERIODICX)
    for (int j = 1; j < nyn - 1; j++)
      for (int k = 1; k < nzn - 1; k++)
        for (int m = 0; m < 10; m++) {
          FPinterp *a = arr[m];
          a[IDX3(1, j, k, nyn, nzn)] *= 2;
          a[IDX3(nxn - 2, j, k, nyn, nzn)] *= 2;
        }
  if (!PERIODICY)
    for (int i = 1; i < nxn - 1; i++)
      for (int k = 1; k < nzn - 1; k++)
        for (int m = 0; m < 10; m++) {
          FPinterp *a = arr[m];
          a[IDX3(i, 1, k, nyn, nzn)] *= 2;
          a[IDX3(i, nyn - 2, k, nyn, nzn)] *= 2;
        }
  if (!PERIODICZ)
    for (int i = 1; i < nxn - 1; i++)
      for (int j = 1; j < nyn - 1; j++)
        for (int m = 0; m < 10; m++) {
          FPinterp *a = arr[m];
          a[IDX3(i, j, 1, nyn, nzn)] *= 2;
          a[IDX3(i, j, nzn - 2, nyn, nzn)] *= 2;
        }
}

/* sumOverSpecies (src/InterpDensNet.cpp:120-144): net += species, species
 * in order. */
void OSFX(oracle_sum_species)(const oracle_grid *grd, oracle_moments *net,
                              const oracle_moments *ids, int ns) {
  long nn = (long)grd->nxn * grd->nyn * grd->nzn;
  for (int is = 0; is < ns; is++) {
    FPinterp *dst[10] = {net->rhon, net->Jx,  net->Jy,  net->Jz,  net->pxx,
                         net->pxy,  net->pxz, net->pyy, net->pyz, net->pzz};
    FPinterp *src[10] = {ids[is].rhon, ids[is].Jx,  ids[is].Jy,  ids[is].Jz,
                         ids[is].pxx,  ids[is].pxy, ids[is].pxz, ids[is].pyy,
                         ids[is].pyz,  ids[is].pzz};
    for (long n = 0; n < nn; n++)
      for (int m = 0; m < 10; m++) dst[m][n] += src[m][n];
  }
}

/* ------------------------------------------------------------------ */
/* Ghost updates of one scalar array: the four routines of src/BC.cpp  */
/*   applyBCscalarDensN      (BC.cpp:11-123)   nodes, copied faces     */
/*   applyBCscalarFieldNzero (BC.cpp:242-352)  nodes, zeroed faces     */
/*   applyBCscalarDensC      (BC.cpp:623-735)  cells, copied faces     */
/*   applyBCscalarFieldC     (BC.cpp:737-850)  cells, copied faces     */
/* share one sequence: periodic faces x, y, z; edges along x, y, z;    */
/* the eight corners; non-periodic faces x, y, z.  A node array takes  */
/* its ghosts from planes n-3 and 2, a cell array from n-2 and 1.  All */
/* four run the y-edge loop to the node count (j < nyn - 1), which for */
/* a cell array reaches the ghost layer j = nyc - 1.                   */
/* ------------------------------------------------------------------ */
static void bc_scalar(FPinterp *a, int n0, int n1, int n2, int nyn, int nodes,
                      int zero_faces, int PX, int PY, int PZ) {
  const int lo = nodes ? 3 : 2, hi = nodes ? 2 : 1;
#define A(i, j, k) a[IDX3(i, j, k, n1, n2)]
  if (PX)
    for (int j = 1; j < n1 - 1; j++)
      for (int k = 1; k < n2 - 1; k++) {
        A(0, j, k) = A(n0 - lo, j, k);
        A(n0 - 1, j, k) = A(hi, j, k);
      }
  if (PY)
    for (int i = 1; i < n0 - 1; i++)
      for (int k = 1; k < n2 - 1; k++) {
        A(i, 0, k) = A(i, n1 - lo, k);
        A(i, n1 - 1, k) = A(i, hi, k);
      }
  if (PZ)
    for (int i = 1; i < n0 - 1; i++)
      for (int j = 1; j < n1 - 1; j++) {
        A(i, j, 0) = A(i, j, n2 - lo);
        A(i, j, n2 - 1) = A(i, j, hi);
      }
  if (PY || PZ)
    for (int i = 1; i < n0 - 1; i++) {
      A(i, n1 - 1, n2 - 1) = A(i, hi, hi);
      A(i, 0, 0) = A(i, n1 - lo, n2 - lo);
      A(i, 0, n2 - 1) = A(i, n1 - lo, hi);
      A(i, n1 - 1, 0) = A(i, hi, n2 - lo);
    }
  if (PX || PZ)
    for (int j = 1; j < nyn - 1; j++) {
      A(n0 - 1, j, n2 - 1) = A(hi, j, hi);
      A(0, j, 0) = A(n0 - lo, j, n2 - lo);
      A(0, j, n2 - 1) = A(n0 - lo, j, hi);
      A(n0 - 1, j, 0) = A(hi, j, n2 - lo);
    }
  if (PX || PY)
    for (int k = 1; k < n2 - 1; k++) {
      A(n0 - 1, n1 - 1, k) = A(hi, hi, k);
      A(0, 0, k) = A(n0 - lo, n1 - lo, k);
      A(n0 - 1, 0, k) = A(hi, n1 - lo, k);
      A(0, n1 - 1, k) = A(n0 - lo, hi, k);
    }
  if (PX || PY || PZ) {
    A(n0 - 1, n1 - 1, n2 - 1) = A(hi, hi, hi);
    A(0, n1 - 1, n2 - 1) = A(n0 - lo, hi, hi);
    A(n0 - 1, 0, n2 - 1) = A(hi, n1 - lo, hi);
    A(0, 0, n2 - 1) = A(n0 - lo, n1 - lo, hi);
    A(n0 - 1, n1 - 1, 0) = A(hi, hi, n2 - lo);
    A(0, n1 - 1, 0) = A(n0 - lo, hi, n2 - lo);
    A(n0 - 1, 0, 0) = A(hi, n1 - lo, n2 - lo);
    A(0, 0, 0) = A(n0 - lo, n1 - lo, n2 - lo);
  }
  if (!PX)
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n2; k++) {
        A(0, j, k) = zero_faces ? 0.0 : A(1, j, k);
        A(n0 - 1, j, k) = zero_faces ? 0.0 : A(n0 - 2, j, k);
      }
  if (!PY)
    for (int i = 0; i < n0; i++)
      for (int k = 0; k < n2; k++) {
        A(i, 0, k) = zero_faces ? 0.0 : A(i, 1, k);
        A(i, n1 - 1, k) = zero_faces ? 0.0 : A(i, n1 - 2, k);
      }
  if (!PZ)
    for (int i = 0; i < n0; i++)
      for (int j = 0; j < n1; j++) {
        A(i, j, 0) = zero_faces ? 0.0 : A(i, j, 1);
        A(i, j, n2 - 1) = zero_faces ? 0.0 : A(i, j, n2 - 2);
      }
#undef A
}

/* kind: 0 applyBCscalarDensN, 1 applyBCscalarFieldNzero,
 *       2 applyBCscalarDensC, 3 applyBCscalarFieldC */
void OSFX(oracle_bc_scalar)(const oracle_grid *grd, FPinterp *a, int kind, int PX,
                            int PY, int PZ) {
  if (kind <= 1)
    bc_scalar(a, grd->nxn, grd->nyn, grd->nzn, grd->nyn, 1, kind == 1, PX, PY, PZ);
  else
    bc_scalar(a, grd->nxc, grd->nyc, grd->nzc, grd->nyn, 0, 0, PX, PY, PZ);
}

/* interpN2Cinterp (src/Grid.cpp:164-185): interior cells from their eight
 * nodes, summed in the reference's order. */
static void interp_n2c(const oracle_grid *grd, FPinterp *c, const FPinterp *n) {
  const int nyn = grd->nyn, nzn = grd->nzn, nyc = grd->nyc, nzc = grd->nzc;
#define N(i, j, k) n[IDX3(i, j, k, nyn, nzn)]
  for (int i = 1; i < grd->nxc - 1; i++)
    for (int j = 1; j < nyc - 1; j++)
      for (int k = 1; k < nzc - 1; k++)
        c[IDX3(i, j, k, nyc, nzc)] =
            .125 * (N(i, j, k) + N(i + 1, j, k) + N(i, j + 1, k) + N(i, j, k + 1) +
                    N(i + 1, j + 1, k) + N(i + 1, j, k + 1) + N(i, j + 1, k + 1) +
                    N(i + 1, j + 1, k + 1));
#undef N
}

/* interpC2Ninterp (src/Grid.cpp:122-141): interior nodes from their eight
 * cells. */
static void interp_c2n(const oracle_grid *grd, FPinterp *n, const FPinterp *c) {
  const int nyn = grd->nyn, nzn = grd->nzn, nyc = grd->nyc, nzc = grd->nzc;
#define C(i, j, k) c[IDX3(i, j, k, nyc, nzc)]
  for (int i = 1; i < grd->nxn - 1; i++)
    for (int j = 1; j < nyn - 1; j++)
      for (int k = 1; k < nzn - 1; k++)
        n[IDX3(i, j, k, nyn, nzn)] =
            .125 * (C(i, j, k) + C(i - 1, j, k) + C(i, j - 1, k) + C(i, j, k - 1) +
                    C(i - 1, j - 1, k) + C(i - 1, j, k - 1) + C(i, j - 1, k - 1) +
                    C(i - 1, j - 1, k - 1));
#undef C
}

/* divN2C (src/Grid.cpp:267-319): interior cells. */
static void div_n2c(const oracle_grid *grd, FPfield *d, const FPfield *X, const FPfield *Y,
                    const FPfield *Z) {
  const int nyn = grd->nyn, nzn = grd->nzn, nyc = grd->nyc, nzc = grd->nzc;
  const FPfield invdx = (FPfield)grd->invdx, invdy = (FPfield)grd->invdy,
                invdz = (FPfield)grd->invdz;
#define V(a, i, j, k) a[IDX3(i, j, k, nyn, nzn)]
  for (int i = 1; i < grd->nxc - 1; i++)
    for (int j = 1; j < nyc - 1; j++)
      for (int k = 1; k < nzc - 1; k++) {
        FPfield cx = .25 * (V(X, i + 1, j, k) - V(X, i, j, k)) * invdx +
                     .25 * (V(X, i + 1, j, k + 1) - V(X, i, j, k + 1)) * invdx +
                     .25 * (V(X, i + 1, j + 1, k) - V(X, i, j + 1, k)) * invdx +
                     .25 * (V(X, i + 1, j + 1, k + 1) - V(X, i, j + 1, k + 1)) * invdx;
        FPfield cy = .25 * (V(Y, i, j + 1, k) - V(Y, i, j, k)) * invdy +
                     .25 * (V(Y, i, j + 1, k + 1) - V(Y, i, j, k + 1)) * invdy +
                     .25 * (V(Y, i + 1, j + 1, k) - V(Y, i + 1, j, k)) * invdy +
                     .25 * (V(Y, i + 1, j + 1, k + 1) - V(Y, i + 1, j, k + 1)) * invdy;
        FPfield cz = .25 * (V(Z, i, j, k + 1) - V(Z, i, j, k)) * invdz +
                     .25 * (V(Z, i + 1, j, k + 1) - V(Z, i + 1, j, k)) * invdz +
                     .25 * (V(Z, i, j + 1, k + 1) - V(Z, i, j + 1, k)) * invdz +
                     .25 * (V(Z, i + 1, j + 1, k + 1) - V(Z, i + 1, j + 1, k)) * invdz;
        d[IDX3(i, j, k, nyc, nzc)] = cx + cy + cz;
      }
#undef V
}

/* The driver's step between sumOverSpecies and calculateHatDensities
 * (src/sputniPIC_CPU.cpp:276-277): ghost nodes of the net rhon
 * (applyBCscalarDensN, in place), then rhoc = interpN2Cinterp(rhon) on the
 * interior cells. */
void OSFX(oracle_rhoc_from_rhon)(const oracle_grid *grd, FPinterp *rhoc, FPinterp *rhon,
                                 int PX, int PY, int PZ) {
  OSFX(oracle_bc_scalar)(grd, rhon, 0, PX, PY, PZ);
  interp_n2c(grd, rhoc, rhon);
}

/* calculateHatDensities (src/MaxwellSolver.cpp:465-604).  ids: ns species'
 * node moments (Jx Jy Jz and the six pressure components are read);
 * B: Bxn Byn Bzn; rhoc: net cell density.  Writes rhoh (cells) and
 * Jxh Jyh Jzh (nodes). */
void OSFX(oracle_hat_densities)(const oracle_grid *grd, const oracle_moments *ids, int ns,
                                const double *qom, const FPfield *Bxn, const FPfield *Byn,
                                const FPfield *Bzn, const FPinterp *rhoc, double dt,
                                double c, double th, int PX, int PY, int PZ,
                                FPinterp *rhoh, FPinterp *Jxh, FPinterp *Jyh,
                                FPinterp *Jzh) {
  const int nxn = grd->nxn, nyn = grd->nyn, nzn = grd->nzn;
  const int nxc = grd->nxc, nyc = grd->nyc, nzc = grd->nzc;
  const long nn = (long)nxn * nyn * nzn, nc = (long)nxc * nyc * nzc;
  const FPinterp invdx = (FPinterp)grd->invdx, invdy = (FPinterp)grd->invdy,
                 invdz = (FPinterp)grd->invdz;
  FPinterp *tXC = calloc(nc, sizeof(FPinterp)), *tYC = calloc(nc, sizeof(FPinterp)),
           *tZC = calloc(nc, sizeof(FPinterp));
  FPinterp *tXN = calloc(nn, sizeof(FPinterp)), *tYN = calloc(nn, sizeof(FPinterp)),
           *tZN = calloc(nn, sizeof(FPinterp));
  for (long n = 0; n < nn; n++) Jxh[n] = Jyh[n] = Jzh[n] = 0.0;
  for (long n = 0; n < nc; n++) rhoh[n] = 0.0;
  OSFX(oracle_bc_scalar)(grd, rhoh, 2, PX, PY, PZ);
  for (int is = 0; is < ns; is++) {
    const oracle_moments *m = &ids[is];
    /* divSymmTensorN2C (src/Grid.cpp:322-398) */
#define P(a, i, j, k) m->a[IDX3(i, j, k, nyn, nzn)]
#define D4(a, i0, j0, k0, i1, j1, k1, inv)                                     \
  (.25 * (P(a, i0 + i1, j0 + j1, k0 + k1) - P(a, i0, j0, k0)) * inv)
    for (int i = 1; i < nxc - 1; i++)
      for (int j = 1; j < nyc - 1; j++)
        for (int k = 1; k < nzc - 1; k++) {
          FPinterp c1X, c2X, c3X, c1Y, c2Y, c3Y, c1Z, c2Z, c3Z;
          c1X = .25 * (P(pxx, i + 1, j, k) - P(pxx, i, j, k)) * invdx +
                .25 * (P(pxx, i + 1, j, k + 1) - P(pxx, i, j, k + 1)) * invdx +
                .25 * (P(pxx, i + 1, j + 1, k) - P(pxx, i, j + 1, k)) * invdx +
                .25 * (P(pxx, i + 1, j + 1, k + 1) - P(pxx, i, j + 1, k + 1)) * invdx;
          c2X = .25 * (P(pxy, i + 1, j, k) - P(pxy, i, j, k)) * invdx +
                .25 * (P(pxy, i + 1, j, k + 1) - P(pxy, i, j, k + 1)) * invdx +
                .25 * (P(pxy, i + 1, j + 1, k) - P(pxy, i, j + 1, k)) * invdx +
                .25 * (P(pxy, i + 1, j + 1, k + 1) - P(pxy, i, j + 1, k + 1)) * invdx;
          c3X = .25 * (P(pxz, i + 1, j, k) - P(pxz, i, j, k)) * invdx +
                .25 * (P(pxz, i + 1, j, k + 1) - P(pxz, i, j, k + 1)) * invdx +
                .25 * (P(pxz, i + 1, j + 1, k) - P(pxz, i, j + 1, k)) * invdx +
                .25 * (P(pxz, i + 1, j + 1, k + 1) - P(pxz, i, j + 1, k + 1)) * invdx;
          c1Y = .25 * (P(pxy, i, j + 1, k) - P(pxy, i, j, k)) * invdy +
                .25 * (P(pxy, i, j + 1, k + 1) - P(pxy, i, j, k + 1)) * invdy +
                .25 * (P(pxy, i + 1, j + 1, k) - P(pxy, i + 1, j, k)) * invdy +
                .25 * (P(pxy, i + 1, j + 1, k + 1) - P(pxy, i + 1, j, k + 1)) * invdy;
          c2Y = .25 * (P(pyy, i, j + 1, k) - P(pyy, i, j, k)) * invdy +
                .25 * (P(pyy, i, j + 1, k + 1) - P(pyy, i, j, k + 1)) * invdy +
                .25 * (P(pyy, i + 1, j + 1, k) - P(pyy, i + 1, j, k)) * invdy +
                .25 * (P(pyy, i + 1, j + 1, k + 1) - P(pyy, i + 1, j, k + 1)) * invdy;
          c3Y = .25 * (P(pyz, i, j + 1, k) - P(pyz, i, j, k)) * invdy +
                .25 * (P(pyz, i, j + 1, k + 1) - P(pyz, i, j, k + 1)) * invdy +
                .25 * (P(pyz, i + 1, j + 1, k) - P(pyz, i + 1, j, k)) * invdy +
                .25 * (P(pyz, i + 1, j + 1, k + 1) - P(pyz, i + 1, j, k + 1)) * invdy;
          c1Z = .25 * (P(pxz, i, j, k + 1) - P(pxz, i, j, k)) * invdz +
                .25 * (P(pxz, i + 1, j, k + 1) - P(pxz, i + 1, j, k)) * invdz +
                .25 * (P(pxz, i, j + 1, k + 1) - P(pxz, i, j + 1, k)) * invdz +
                .25 * (P(pxz, i + 1, j + 1, k + 1) - P(pxz, i + 1, j + 1, k)) * invdz;
          c2Z = .25 * (P(pyz, i, j, k + 1) - P(pyz, i, j, k)) * invdz +
                .25 * (P(pyz, i + 1, j, k + 1) - P(pyz, i + 1, j, k)) * invdz +
                .25 * (P(pyz, i, j + 1, k + 1) - P(pyz, i, j + 1, k)) * invdz +
                .25 * (P(pyz, i + 1, j + 1, k + 1) - P(pyz, i + 1, j + 1, k)) * invdz;
          c3Z = .25 * (P(pzz, i, j, k + 1) - P(pzz, i, j, k)) * invdz +
                .25 * (P(pzz, i + 1, j, k + 1) - P(pzz, i + 1, j, k)) * invdz +
                .25 * (P(pzz, i, j + 1, k + 1) - P(pzz, i, j + 1, k)) * invdz +
                .25 * (P(pzz, i + 1, j + 1, k + 1) - P(pzz, i + 1, j + 1, k)) * invdz;
          tXC[IDX3(i, j, k, nyc, nzc)] = c1X + c1Y + c1Z;
          tYC[IDX3(i, j, k, nyc, nzc)] = c2X + c2Y + c2Z;
          tZC[IDX3(i, j, k, nyc, nzc)] = c3X + c3Y + c3Z;
        }
#undef D4
#undef P
    /* scale (src/Basic.cpp:135-141) by -dt/2, FPfield factor */
    const FPfield alfa = (FPfield)(-dt / 2.0);
    for (long n = 0; n < nc; n++) {
      tXC[n] *= alfa;
      tYC[n] *= alfa;
      tZC[n] *= alfa;
    }
    OSFX(oracle_bc_scalar)(grd, tXC, 2, PX, PY, PZ);
    OSFX(oracle_bc_scalar)(grd, tYC, 2, PX, PY, PZ);
    OSFX(oracle_bc_scalar)(grd, tZC, 2, PX, PY, PZ);
    interp_c2n(grd, tXN, tXC);
    interp_c2n(grd, tYN, tYC);
    interp_c2n(grd, tZN, tZC);
    for (long n = 0; n < nn; n++) {
      tXN[n] += m->Jx[n];
      tYN[n] += m->Jy[n];
      tZN[n] += m->Jz[n];
    }
    /* PIdot */
    const FPfield beta = (FPfield)(.5 * qom[is] * dt / c);
    for (long n = 0; n < nn; n++) {
      const FPfield omcx = beta * Bxn[n], omcy = beta * Byn[n], omcz = beta * Bzn[n];
      const FPfield edotb = tXN[n] * omcx + tYN[n] * omcy + tZN[n] * omcz;
      const FPfield denom = (FPfield)(1 / (1.0 + omcx * omcx + omcy * omcy + omcz * omcz));
      Jxh[n] += (tXN[n] + (tYN[n] * omcz - tZN[n] * omcy + edotb * omcx)) * denom;
      Jyh[n] += (tYN[n] + (tZN[n] * omcx - tXN[n] * omcz + edotb * omcy)) * denom;
      Jzh[n] += (tZN[n] + (tXN[n] * omcy - tYN[n] * omcx + edotb * omcz)) * denom;
    }
  }
  for (long n = 0; n < nc; n++) tXC[n] = 0.0;
  OSFX(oracle_bc_scalar)(grd, Jxh, 0, PX, PY, PZ);
  OSFX(oracle_bc_scalar)(grd, Jyh, 0, PX, PY, PZ);
  OSFX(oracle_bc_scalar)(grd, Jzh, 0, PX, PY, PZ);
  div_n2c(grd, tXC, Jxh, Jyh, Jzh);
  for (long n = 0; n < nc; n++) rhoh[n] = rhoc[n] - dt * th * tXC[n];
  OSFX(oracle_bc_scalar)(grd, rhoh, 2, PX, PY, PZ);
  free(tXC); free(tYC); free(tZC);
  free(tXN); free(tYN); free(tZN);
}

/* MaxwellImage (src/MaxwellSolver.cpp:95-257): the implicit field
 * operator applied to a Krylov vector, vec and im holding the interior
 * nodes' (x, y, z) components interleaved, i slowest (solver2phys and
 * phys2solver, src/Basic.cpp:365-392). */
void OSFX(oracle_maxwell_image)(const oracle_grid *grd, const FPfield *Bxn,
                                const FPfield *Byn, const FPfield *Bzn,
                                const oracle_moments *ids, int ns, const double *qom,
                                double dt, double c, double th, double fourpi, int PX,
                                int PY, int PZ, const FPfield *vec, FPfield *im) {
  const int nxn = grd->nxn, nyn = grd->nyn, nzn = grd->nzn;
  const int nxc = grd->nxc, nyc = grd->nyc, nzc = grd->nzc;
  const long nn = (long)nxn * nyn * nzn, nc = (long)nxc * nyc * nzc;
  const FPfield invdx = (FPfield)grd->invdx, invdy = (FPfield)grd->invdy,
                invdz = (FPfield)grd->invdz;
  const FPfield delt = (FPfield)(c * th * dt);
  FPfield *v[3], *img[3], *D[3], *tmp[3];
  for (int a = 0; a < 3; a++) {
    v[a] = calloc(nn, sizeof(FPfield));
    img[a] = calloc(nn, sizeof(FPfield));
    D[a] = calloc(nn, sizeof(FPfield));
    tmp[a] = calloc(nn, sizeof(FPfield));
  }
  FPfield *divC = calloc(nc, sizeof(FPfield));
  const FPfield *p = vec;
  for (int i = 1; i < nxn - 1; i++)
    for (int j = 1; j < nyn - 1; j++)
      for (int k = 1; k < nzn - 1; k++) {
        const long n = IDX3(i, j, k, nyn, nzn);
        v[0][n] = *p++;
        v[1][n] = *p++;
        v[2][n] = *p++;
      }
  for (int a = 0; a < 3; a++) OSFX(oracle_bc_scalar)(grd, (FPinterp *)v[a], 1, PX, PY, PZ);
  /* lapN2N_V (src/Grid.cpp:668-717), then neg (src/Basic.cpp:281-286) */
  for (int a = 0; a < 3; a++) {
    const FPfield *s = v[a];
    for (int i = 1; i < nxn - 1; i++)
      for (int j = 1; j < nyn - 1; j++)
        for (int k = 1; k < nzn - 1; k++) {
          const long n = IDX3(i, j, k, nyn, nzn);
          img[a][n] = (s[IDX3(i - 1, j, k, nyn, nzn)] - 2 * s[n] +
                       s[IDX3(i + 1, j, k, nyn, nzn)]) * invdx * invdx +
                      (s[IDX3(i, j - 1, k, nyn, nzn)] - 2 * s[n] +
                       s[IDX3(i, j + 1, k, nyn, nzn)]) * invdy * invdy +
                      (s[IDX3(i, j, k - 1, nyn, nzn)] - 2 * s[n] +
                       s[IDX3(i, j, k + 1, nyn, nzn)]) * invdz * invdz;
        }
    for (long n = 0; n < nn; n++) img[a][n] = -img[a][n];
  }
  for (int is = 0; is < ns; is++) {
    const FPfield beta = (FPfield)(.5 * qom[is] * dt / c);
    const FPinterp *rhon = ids[is].rhon;
    for (long n = 0; n < nn; n++) {
      const FPfield omcx = beta * Bxn[n], omcy = beta * Byn[n], omcz = beta * Bzn[n];
      const FPfield edotb = v[0][n] * omcx + v[1][n] * omcy + v[2][n] * omcz;
      const FPfield denom = (FPfield)(fourpi / 2 * delt * dt / c * qom[is] * rhon[n] /
                                      (1.0 + omcx * omcx + omcy * omcy + omcz * omcz));
      D[0][n] += (v[0][n] + (v[1][n] * omcz - v[2][n] * omcy + edotb * omcx)) * denom;
      D[1][n] += (v[1][n] + (v[2][n] * omcx - v[0][n] * omcz + edotb * omcy)) * denom;
      D[2][n] += (v[2][n] + (v[0][n] * omcy - v[1][n] * omcx + edotb * omcz)) * denom;
    }
  }
  for (int a = 0; a < 3; a++) OSFX(oracle_bc_scalar)(grd, (FPinterp *)D[a], 1, PX, PY, PZ);
  div_n2c(grd, divC, D[0], D[1], D[2]);
  OSFX(oracle_bc_scalar)(grd, (FPinterp *)divC, 3, PX, PY, PZ);
  /* gradC2N (src/Grid.cpp:187-225) */
#define S(i, j, k) divC[IDX3(i, j, k, nyc, nzc)]
  for (int i = 1; i < nxn - 1; i++)
    for (int j = 1; j < nyn - 1; j++)
      for (int k = 1; k < nzn - 1; k++) {
        const long n = IDX3(i, j, k, nyn, nzn);
        tmp[0][n] = .25 * (S(i, j, k) - S(i - 1, j, k)) * invdx +
                    .25 * (S(i, j, k - 1) - S(i - 1, j, k - 1)) * invdx +
                    .25 * (S(i, j - 1, k) - S(i - 1, j - 1, k)) * invdx +
                    .25 * (S(i, j - 1, k - 1) - S(i - 1, j - 1, k - 1)) * invdx;
        tmp[1][n] = .25 * (S(i, j, k) - S(i, j - 1, k)) * invdy +
                    .25 * (S(i, j, k - 1) - S(i, j - 1, k - 1)) * invdy +
                    .25 * (S(i - 1, j, k) - S(i - 1, j - 1, k)) * invdy +
                    .25 * (S(i - 1, j, k - 1) - S(i - 1, j - 1, k - 1)) * invdy;
        tmp[2][n] = .25 * (S(i, j, k) - S(i, j, k - 1)) * invdz +
                    .25 * (S(i - 1, j, k) - S(i - 1, j, k - 1)) * invdz +
                    .25 * (S(i, j - 1, k) - S(i, j - 1, k - 1)) * invdz +
                    .25 * (S(i - 1, j - 1, k) - S(i - 1, j - 1, k - 1)) * invdz;
      }
#undef S
  /* sub, scale by delt*delt, sum D, sum vect (src/Basic.cpp) */
  const FPfield dd = delt * delt;
  for (int a = 0; a < 3; a++)
    for (long n = 0; n < nn; n++) {
      img[a][n] -= tmp[a][n];
      img[a][n] *= dd;
      img[a][n] += D[a][n];
      img[a][n] += v[a][n];
    }
  FPfield *q = im;
  for (int i = 1; i < nxn - 1; i++)
    for (int j = 1; j < nyn - 1; j++)
      for (int k = 1; k < nzn - 1; k++) {
        const long n = IDX3(i, j, k, nyn, nzn);
        *q++ = img[0][n];
        *q++ = img[1][n];
        *q++ = img[2][n];
      }
  for (int a = 0; a < 3; a++) {
    free(v[a]); free(img[a]); free(D[a]); free(tmp[a]);
  }
  free(divC);
}

/* Wall-clock helper for the CPU baseline (omp_get_wtime when OpenMP is
 * on). */
double OSFX(oracle_wtime)(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

int OSFX(oracle_max_threads)(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* The oracle's form of the CPU baseline's timed region: ncycles of mover
 * for every species (parallel over particles), zero, deposit with an
 * OpenMP loop over species (src/sputniPIC_CPU.cpp:232-256). */
double OSFX(oracle_time_cycles)(const oracle_grid *grd, const oracle_nodes *nd,
                                const oracle_fields *field,
                                oracle_particles *parts, oracle_moments *moms,
                                int ns, const oracle_push *push, int ncycles,
                                double *t_mover, double *t_deposit) {
  double tm = 0, td = 0;
  for (int c = 0; c < ncycles; c++) {
    double t0 = OSFX(oracle_wtime)();
    for (int is = 0; is < ns; is++)
      OSFX(oracle_mover_pc)(grd, nd, field, &parts[is], push);
    double t1 = OSFX(oracle_wtime)();
#pragma omp parallel for
    for (int is = 0; is < ns; is++) {
      OSFX(oracle_zero_moments)(grd, &moms[is]);
      OSFX(oracle_interp_p2g)(grd, nd, &parts[is], &moms[is]);
    }
    double t2 = OSFX(oracle_wtime)();
    tm += t1 - t0;
    td += t2 - t1;
  }
  if (t_mover) *t_mover = tm;
  if (t_deposit) *t_deposit = td;
  return tm + td;
}
