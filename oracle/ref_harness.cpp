// ref_harness.cpp -- test-only C entry points around the reference's own
// hot-path translation units, compiled from the reference checkout by
// oracle/Makefile into oracle/_ref/libref_{f32,f64}.so.
//
// TEST INFRASTRUCTURE ONLY: used to pin the C restatement in spic_oracle.c
// and to write the fixtures under tests/golden/.  Nothing here is product
// code.  It includes the reference headers and calls, unmodified:
//   setGrid (src/Grid.cpp:5), field_allocate / field_aux_allocate
//   (src/EMfield.cpp), interp_dens_species_allocate
//   (src/InterpDensSpecies.cpp:4), particle_allocate (src/Particles.cpp:5),
//   initGEM (src/IC.cpp:7), mover_PC / interpP2G / mover_interp
//   (src/Mover.cpp), particle_sort (src/Particles.cpp:72), applyBCids
//   (src/BC.cpp:362), setZeroSpeciesDensities (src/InterpDensNet.cpp:40),
//   setZeroNetDensities / sumOverSpecies (src/InterpDensNet.cpp:78-144),
//   applyBCscalarDensN (src/BC.cpp:11), interpN2Cinterp (src/Grid.cpp:164),
//   calculateHatDensities / MaxwellImage (src/MaxwellSolver.cpp:465, 95).
// It fills struct parameters directly, with the defaults of
// src/RW_IO.cpp:37-295 for keys it does not set, because readInputFile
// lives in RW_IO.cpp, which needs HDF5 (absent here).
#include <cstdio>
#include <cstring>
#include <string>

#include "BC.h"
#include "EMfield.h"
#include "Grid.h"
#include "IC.h"
#include "InterpDensNet.h"
#include "InterpDensSpecies.h"
#include "InterpDens_aux.h"
#include "MaxwellSolver.h"
#include "Mover.h"
#include "Parameters.h"
#include "Particles.h"

template <class T>
static void to_double(const T *src, double *dst, long n) {
  for (long i = 0; i < n; i++) dst[i] = (double)src[i];
}
template <class T>
static void from_double(const double *src, T *dst, long n) {
  for (long i = 0; i < n; i++) dst[i] = (T)src[i];
}

extern "C" {

struct ref_cfg {
  int nxc, nyc, nzc;
  double Lx, Ly, Lz;
  double dt, c;
  int NiterMover, n_sub_cycles;
  int ns;
  int npcelx[6], npcely[6], npcelz[6];
  double qom[6], rhoINIT[6], uth[6], vth[6], wth[6], u0[6], v0[6], w0[6];
  double B0x, B0y, B0z, delta;
  int PERIODICX, PERIODICY, PERIODICZ;
  long sort_cps;
};

struct ref_state {
  parameters param;
  grid grd;
  EMfield field;
  EMfield_aux field_aux;
  interpDensSpecies *ids;
  particles *part;
};

static void set_defaults(parameters *p) {
  // src/RW_IO.cpp:37-295 defaults for the members the hot path touches
  p->c = 1;
  p->fourpi = 12.5663706144;
  p->th = 1;
  p->NiterMover = 3;
  p->n_sub_cycles = 1;
  p->sort = false;
  p->sort_every_n = 50;
  p->sort_cps = 4;
  p->number_of_batches = 16;
  p->threads_per_block = 256;
  p->NpMaxNpRatio = 1.0;
  p->PERIODICX = true;
  p->PERIODICY = false;
  p->PERIODICZ = true;
  p->PERIODICX_P = true;
  p->PERIODICY_P = false;
  p->PERIODICZ_P = true;
  p->delta = 0.5;
  p->RESTART = false;
  p->first_cycle_n = 1;
  // a directory that does not exist keeps initGEM from writing IC dumps
  // (src/IC.cpp:178-183)
  p->RestartDirName = "/nonexistent-spic-ref-harness";
  p->SaveDirName = p->RestartDirName;
}

// The reference's state for cfg; with `ic`, initGEM fills fields and
// particles, else they stay as the allocators leave them, for the caller to
// set (ref_set_field / ref_set_particles).  With io_dir, initGEM's own
// dump directory: it writes its save_ic_data files there (restart 0) or
// reads its read_ic_data files from there instead of computing the IC
// (restart 1).
static void *create_state(const ref_cfg *cfg, bool ic, const char *io_dir = nullptr,
                          int restart = 0) {
  ref_state *s = new ref_state();
  parameters *p = &s->param;
  set_defaults(p);
  if (io_dir) {
    p->RestartDirName = io_dir;
    p->SaveDirName = io_dir;
    p->RESTART = restart != 0;
  }
  p->nxc = cfg->nxc;
  p->nyc = cfg->nyc;
  p->nzc = cfg->nzc;
  p->Lx = cfg->Lx;
  p->Ly = cfg->Ly;
  p->Lz = cfg->Lz;
  p->dt = cfg->dt;
  p->c = cfg->c;
  p->NiterMover = cfg->NiterMover;
  p->n_sub_cycles = cfg->n_sub_cycles;
  p->ns = cfg->ns;
  p->B0x = cfg->B0x;
  p->B0y = cfg->B0y;
  p->B0z = cfg->B0z;
  p->delta = cfg->delta;
  p->PERIODICX = cfg->PERIODICX != 0;
  p->PERIODICY = cfg->PERIODICY != 0;
  p->PERIODICZ = cfg->PERIODICZ != 0;
  p->sort_cps = cfg->sort_cps;
  for (int i = 0; i < p->ns; i++) {
    p->npcelx[i] = cfg->npcelx[i];
    p->npcely[i] = cfg->npcely[i];
    p->npcelz[i] = cfg->npcelz[i];
    p->qom[i] = cfg->qom[i];
    p->rhoINIT[i] = cfg->rhoINIT[i];
    p->uth[i] = cfg->uth[i];
    p->vth[i] = cfg->vth[i];
    p->wth[i] = cfg->wth[i];
    p->u0[i] = cfg->u0[i];
    p->v0[i] = cfg->v0[i];
    p->w0[i] = cfg->w0[i];
    // particle count as src/RW_IO.cpp:207-211
    int npcel = p->npcelx[i] * p->npcely[i] * p->npcelz[i];
    p->np[i] = (long)npcel * p->nxc * p->nyc * p->nzc;
    p->npMax[i] = (long)(p->NpMaxNpRatio * p->np[i]);
  }
  setGrid(p, &s->grd);
  field_allocate(&s->grd, &s->field);
  field_aux_allocate(&s->grd, &s->field_aux);
  s->ids = new interpDensSpecies[p->ns];
  for (int is = 0; is < p->ns; is++)
    interp_dens_species_allocate(&s->grd, &s->ids[is], is);
  s->part = new particles[p->ns];
  for (int is = 0; is < p->ns; is++) particle_allocate(p, &s->part[is], is);
  // initGEM fills rhoc only on interior cells; start from zero so the
  // guard cells hold a defined value
  long nc = (long)s->grd.nxc * s->grd.nyc * s->grd.nzc;
  for (int is = 0; is < p->ns; is++)
    std::memset(s->ids[is].rhoc_flat, 0, sizeof(FPinterp) * nc);
  // the cell-centred B is computed on interior cells only: defined guards
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  std::memset(s->field_aux.Bxc_flat, 0, sizeof(FPfield) * nc);
  std::memset(s->field_aux.Byc_flat, 0, sizeof(FPfield) * nc);
  std::memset(s->field_aux.Bzc_flat, 0, sizeof(FPfield) * nc);
  std::memset(s->field_aux.Exth_flat, 0, sizeof(FPfield) * nn);
  std::memset(s->field_aux.Eyth_flat, 0, sizeof(FPfield) * nn);
  std::memset(s->field_aux.Ezth_flat, 0, sizeof(FPfield) * nn);
  if (ic) initGEM(p, &s->grd, &s->field, &s->field_aux, s->part, s->ids);
  return s;
}

void *ref_create(const ref_cfg *cfg) { return create_state(cfg, true); }

// initGEM with its IC dump directory (save_ic_data or read_ic_data).
void *ref_create_io(const ref_cfg *cfg, const char *dir, int restart) {
  return create_state(cfg, true, dir, restart);
}

// which: 0..2 = Bxc Byc Bzc (cells), 3..5 = Exth Eyth Ezth (nodes)
void ref_get_aux(void *h, int which, double *out) {
  ref_state *s = (ref_state *)h;
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  long nc = (long)s->grd.nxc * s->grd.nyc * s->grd.nzc;
  FPfield *a[6] = {s->field_aux.Bxc_flat,  s->field_aux.Byc_flat,  s->field_aux.Bzc_flat,
                   s->field_aux.Exth_flat, s->field_aux.Eyth_flat, s->field_aux.Ezth_flat};
  to_double(a[which], out, which < 3 ? nc : nn);
}

// Without initGEM: for inputs made elsewhere (the device generator's
// Harris sheet at the benchmark's size), copied in afterwards.
void *ref_create_empty(const ref_cfg *cfg) { return create_state(cfg, false); }

void ref_destroy(void *h) {
  ref_state *s = (ref_state *)h;
  for (int is = 0; is < s->param.ns; is++) {
    particle_deallocate(&s->part[is]);
    interp_dens_species_deallocate(&s->grd, &s->ids[is]);
  }
  delete[] s->part;
  delete[] s->ids;
  field_deallocate(&s->grd, &s->field);
  field_aux_deallocate(&s->grd, &s->field_aux);
  grid_deallocate(&s->grd);
  delete s;
}

int ref_fp_bytes(void) { return (int)sizeof(FPpart); }

// grid scalars: nxc nyc nzc nxn nyn nzn | dx dy dz invdx invdy invdz invVOL
// xStart yStart zStart Lx Ly Lz
void ref_grid(void *h, int *ints, double *dbl) {
  grid *g = &((ref_state *)h)->grd;
  ints[0] = g->nxc; ints[1] = g->nyc; ints[2] = g->nzc;
  ints[3] = g->nxn; ints[4] = g->nyn; ints[5] = g->nzn;
  dbl[0] = g->dx; dbl[1] = g->dy; dbl[2] = g->dz;
  dbl[3] = g->invdx; dbl[4] = g->invdy; dbl[5] = g->invdz;
  dbl[6] = g->invVOL;
  dbl[7] = g->xStart; dbl[8] = g->yStart; dbl[9] = g->zStart;
  dbl[10] = g->Lx; dbl[11] = g->Ly; dbl[12] = g->Lz;
}

long ref_nop(void *h, int is) { return ((ref_state *)h)->part[is].nop; }


// which: 0..6 = x y z u v w q
void ref_get_particles(void *h, int is, int which, double *out) {
  particles *p = &((ref_state *)h)->part[is];
  FPpart *a[6] = {p->x, p->y, p->z, p->u, p->v, p->w};
  if (which < 6) to_double(a[which], out, p->npmax);
  else to_double(p->q, out, p->npmax);
}
void ref_set_particles(void *h, int is, int which, const double *in) {
  particles *p = &((ref_state *)h)->part[is];
  FPpart *a[6] = {p->x, p->y, p->z, p->u, p->v, p->w};
  if (which < 6) from_double(in, a[which], p->npmax);
  else from_double(in, p->q, p->npmax);
}

// which: 0..5 = Ex Ey Ez Bxn Byn Bzn ; 6..8 = XN YN ZN
void ref_get_field(void *h, int which, double *out) {
  ref_state *s = (ref_state *)h;
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  FPfield *a[9] = {s->field.Ex_flat,  s->field.Ey_flat,  s->field.Ez_flat,
                   s->field.Bxn_flat, s->field.Byn_flat, s->field.Bzn_flat,
                   s->grd.XN_flat,    s->grd.YN_flat,    s->grd.ZN_flat};
  to_double(a[which], out, nn);
}
void ref_set_field(void *h, int which, const double *in) {
  ref_state *s = (ref_state *)h;
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  FPfield *a[6] = {s->field.Ex_flat,  s->field.Ey_flat,  s->field.Ez_flat,
                   s->field.Bxn_flat, s->field.Byn_flat, s->field.Bzn_flat};
  from_double(in, a[which], nn);
}

// which: 0..9 = rhon Jx Jy Jz pxx pxy pxz pyy pyz pzz ; 10 = rhoc
void ref_get_moment(void *h, int is, int which, double *out) {
  ref_state *s = (ref_state *)h;
  interpDensSpecies *d = &s->ids[is];
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  long nc = (long)s->grd.nxc * s->grd.nyc * s->grd.nzc;
  FPinterp *a[10] = {d->rhon_flat, d->Jx_flat,  d->Jy_flat,  d->Jz_flat,
                     d->pxx_flat,  d->pxy_flat, d->pxz_flat, d->pyy_flat,
                     d->pyz_flat,  d->pzz_flat};
  if (which < 10) to_double(a[which], out, nn);
  else to_double(d->rhoc_flat, out, nc);
}
void ref_set_moment(void *h, int is, int which, const double *in) {
  ref_state *s = (ref_state *)h;
  interpDensSpecies *d = &s->ids[is];
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  FPinterp *a[10] = {d->rhon_flat, d->Jx_flat,  d->Jy_flat,  d->Jz_flat,
                     d->pxx_flat,  d->pxy_flat, d->pxz_flat, d->pyy_flat,
                     d->pyz_flat,  d->pzz_flat};
  from_double(in, a[which], nn);
}

void ref_zero_species_moments(void *h) {
  ref_state *s = (ref_state *)h;
  setZeroSpeciesDensities(s->ids, &s->grd, s->param.ns);
}

int ref_mover_pc(void *h, int is) {
  ref_state *s = (ref_state *)h;
  return mover_PC(&s->part[is], &s->field, &s->grd, &s->param);
}
void ref_interp_p2g(void *h, int is) {
  ref_state *s = (ref_state *)h;
  interpP2G(&s->part[is], &s->ids[is], &s->grd);
}
int ref_mover_interp(void *h, int is) {
  ref_state *s = (ref_state *)h;
  return mover_interp(&s->part[is], &s->field, &s->ids[is], &s->grd, &s->param);
}
void ref_particle_sort(void *h, int is) {
  ref_state *s = (ref_state *)h;
  particle_sort(&s->param, &s->part[is], &s->grd);
}
void ref_apply_bc_ids(void *h, int is) {
  ref_state *s = (ref_state *)h;
  applyBCids(&s->ids[is], &s->grd, &s->param);
}

// sumOverSpecies (src/InterpDensNet.cpp:120-144) into a zeroed net
// (setZeroNetDensities, src/InterpDensNet.cpp:78-110); out receives the ten
// node moments, each nn long, in the order rhon Jx Jy Jz pxx .. pzz.
void ref_sum_species(void *h, double *out) {
  ref_state *s = (ref_state *)h;
  interpDensNet idn;
  interp_dens_net_allocate(&s->grd, &idn);
  setZeroNetDensities(&idn, &s->grd);
  sumOverSpecies(&idn, s->ids, &s->grd, s->param.ns);
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  FPinterp *a[10] = {idn.rhon_flat, idn.Jx_flat,  idn.Jy_flat,  idn.Jz_flat,
                     idn.pxx_flat,  idn.pxy_flat, idn.pxz_flat, idn.pyy_flat,
                     idn.pyz_flat,  idn.pzz_flat};
  for (int m = 0; m < 10; ++m) to_double(a[m], out + m * nn, nn);
  interp_dens_net_deallocate(&s->grd, &idn);
}

// The field-solve side's first steps (src/sputniPIC_CPU.cpp:258-281):
// sumOverSpecies into a zeroed net, applyBCscalarDensN on the net rhon,
// interpN2Cinterp into the net rhoc, then calculateHatDensities.  out
// receives rhon (nn, after the BC), rhoc (nc), rhoh (nc), Jxh Jyh Jzh (nn
// each), in that order.
void ref_hat_densities(void *h, double *out) {
  ref_state *s = (ref_state *)h;
  interpDensNet idn;
  interp_dens_net_allocate(&s->grd, &idn);
  setZeroNetDensities(&idn, &s->grd);
  sumOverSpecies(&idn, s->ids, &s->grd, s->param.ns);
  applyBCscalarDensN(idn.rhon, &s->grd, &s->param);
  interpN2Cinterp(idn.rhoc, idn.rhon, &s->grd);
  interpDens_aux id_aux;
  interp_dens_aux_allocate(&s->grd, &id_aux);
  calculateHatDensities(&id_aux, &idn, s->ids, &s->field, &s->grd, &s->param);
  long nn = (long)s->grd.nxn * s->grd.nyn * s->grd.nzn;
  long nc = (long)s->grd.nxc * s->grd.nyc * s->grd.nzc;
  to_double(idn.rhon_flat, out, nn);
  to_double(&idn.rhoc[0][0][0], out + nn, nc);
  to_double(&id_aux.rhoh[0][0][0], out + nn + nc, nc);
  to_double(&id_aux.Jxh[0][0][0], out + nn + 2 * nc, nn);
  to_double(&id_aux.Jyh[0][0][0], out + 2 * nn + 2 * nc, nn);
  to_double(&id_aux.Jzh[0][0][0], out + 3 * nn + 2 * nc, nn);
  interp_dens_aux_deallocate(&s->grd, &id_aux);
  interp_dens_net_deallocate(&s->grd, &idn);
}

// MaxwellImage of vec (3 (nxn-2)(nyn-2)(nzn-2) values) into im, with the
// species' rhon as they are.
void ref_maxwell_image(void *h, const double *vec, double *im) {
  ref_state *s = (ref_state *)h;
  long n = 3L * (s->grd.nxn - 2) * (s->grd.nyn - 2) * (s->grd.nzn - 2);
  FPfield *v = new FPfield[n];
  FPfield *r = new FPfield[n];
  from_double(vec, v, n);
  MaxwellImage(r, v, &s->field, s->ids, &s->grd, &s->param);
  to_double(r, im, n);
  delete[] v;
  delete[] r;
}

}  // extern "C"

#include <omp.h>

extern "C" {
// The timed region of the CPU baseline: `ncycles` repetitions of the
// reference driver's particle phase (src/sputniPIC_CPU.cpp:232-256):
// mover_PC for each species in turn (each parallel over particles), zero
// the species moments, then interpP2G for all species in an OpenMP loop
// over species.  The BC fold that shares that loop in the driver is not on
// this path and is left out.  Returns total seconds; t_mover / t_deposit
// receive the two parts.
double ref_time_cycles(void *h, int ncycles, double *t_mover, double *t_deposit) {
  ref_state *s = (ref_state *)h;
  double tm = 0, td = 0;
  for (int c = 0; c < ncycles; c++) {
    double t0 = omp_get_wtime();
    for (int is = 0; is < s->param.ns; is++)
      mover_PC(&s->part[is], &s->field, &s->grd, &s->param);
    double t1 = omp_get_wtime();
    setZeroSpeciesDensities(s->ids, &s->grd, s->param.ns);
#pragma omp parallel for
    for (int is = 0; is < s->param.ns; is++)
      interpP2G(&s->part[is], &s->ids[is], &s->grd);
    double t2 = omp_get_wtime();
    tm += t1 - t0;
    td += t2 - t1;
  }
  if (t_mover) *t_mover = tm;
  if (t_deposit) *t_deposit = td;
  return tm + td;
}

// CPU timing of the field-solve side (src/sputniPIC_CPU.cpp:258-281):
// `reps` times sumOverSpecies + applyBCscalarDensN + interpN2Cinterp +
// calculateHatDensities, and `reps` calls of MaxwellImage on one vector.
void ref_time_field(void *h, int reps, double *t_hat, double *t_image) {
  ref_state *s = (ref_state *)h;
  interpDensNet idn;
  interp_dens_net_allocate(&s->grd, &idn);
  interpDens_aux id_aux;
  interp_dens_aux_allocate(&s->grd, &id_aux);
  double t0 = omp_get_wtime();
  for (int r = 0; r < reps; r++) {
    setZeroNetDensities(&idn, &s->grd);
    sumOverSpecies(&idn, s->ids, &s->grd, s->param.ns);
    applyBCscalarDensN(idn.rhon, &s->grd, &s->param);
    interpN2Cinterp(idn.rhoc, idn.rhon, &s->grd);
    calculateHatDensities(&id_aux, &idn, s->ids, &s->field, &s->grd, &s->param);
  }
  double t1 = omp_get_wtime();
  long n = 3L * (s->grd.nxn - 2) * (s->grd.nyn - 2) * (s->grd.nzn - 2);
  FPfield *v = new FPfield[n];
  FPfield *r = new FPfield[n];
  for (long i = 0; i < n; i++) v[i] = (FPfield)(1e-3 * (double)(i % 97));
  double t2 = omp_get_wtime();
  for (int q = 0; q < reps; q++) MaxwellImage(r, v, &s->field, s->ids, &s->grd, &s->param);
  double t3 = omp_get_wtime();
  delete[] v;
  delete[] r;
  interp_dens_aux_deallocate(&s->grd, &id_aux);
  interp_dens_net_deallocate(&s->grd, &idn);
  if (t_hat) *t_hat = t1 - t0;
  if (t_image) *t_image = t3 - t2;
}

int ref_max_threads(void) { return omp_get_max_threads(); }
}  // extern "C"
