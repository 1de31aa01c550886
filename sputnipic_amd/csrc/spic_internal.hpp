// spic_internal.hpp -- types shared by the gfx950 kernels and the C-ABI
// layer of libspic.so.  Not installed; the public ABI is include/spic.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "spic.h"

namespace spic {

// Kernel-side view of the grid in precision T (the reference's FPfield).
// Node coordinates are 1-D tables: XN[i][j][k] == xn[i] (src/Grid.cpp:53-61).
template <typename T>
struct KGrid {
  int nxn, nyn, nzn;
  double xStart, yStart, zStart;
  double Lx, Ly, Lz;
  T invdx, invdy, invdz, invVOL;
  const T *xn, *yn, *zn;
};

template <typename T>
struct KField {
  const T *Ex, *Ey, *Ez, *Bx, *By, *Bz;
};

template <typename T>
struct KPart {
  T *x, *y, *z, *u, *v, *w;
  const T *q;
  int n;
};

struct KMom {
  double *m[10];  // rhon Jx Jy Jz pxx pxy pxz pyy pyz pzz
};

// Per-species push constants, formed on the host exactly as mover_PC forms
// them (src/Mover.cpp:11-12) in the reference's types.
template <typename T>
struct KPush {
  T dt_sub;  // (FPpart)dt / (double)n_sub_cycles
  T dto2;    // .5 * dt_sub
  T qomdt2;  // qom * dto2 / c
  int niter;
  int nloop;  // pushes per call: n_sub_cycles in SPIC_SUBCYCLE_LOOP, else 1
  int perx, pery, perz;
};

// Launchers of the plain (caller-order) kernels, one set per arithmetic
// mode: spic_fast (FMA contraction allowed) and spic_strict (compiled with
// -ffp-contract=off, reference rounding order).
#define SPIC_DECLARE_PLAIN_LAUNCHERS(NS)                                      \
  namespace NS {                                                              \
  template <typename T>                                                       \
  hipError_t launch_mover(const KGrid<T> &g, const KField<T> &f,              \
                          const KPart<T> &p, const KPush<T> &k,               \
                          unsigned int *status, hipStream_t s);               \
  template <typename T>                                                       \
  hipError_t launch_deposit(const KGrid<T> &g, const KPart<T> &p,             \
                            const KMom &m, unsigned int *status,              \
                            hipStream_t s);                                   \
  template <typename T>                                                       \
  hipError_t launch_mover_deposit(const KGrid<T> &g, const KField<T> &f,      \
                                  const KPart<T> &p, const KPush<T> &k,       \
                                  const KMom &m, unsigned int *status,        \
                                  hipStream_t s);                             \
  }

SPIC_DECLARE_PLAIN_LAUNCHERS(spic_fast)
SPIC_DECLARE_PLAIN_LAUNCHERS(spic_strict)

// Reference-order particle sort (spic_sort.hip).
template <typename T>
hipError_t sort_reference_order(const spic_grid &g, int64_t sort_cps,
                                spic_species &sp, void *workspace,
                                size_t workspace_bytes, size_t *needed,
                                hipStream_t s);

hipError_t zero_f64(double *p, size_t n, hipStream_t s);
hipError_t zero_u32(unsigned int *p, size_t n, hipStream_t s);

// Synthetic Harris-sheet input (spic_gem.hip).
hipError_t gem_fields(const spic_grid &g, const spic_gem_params &p,
                      const spic_field &f, hipStream_t s);
hipError_t gem_particles(const spic_grid &g, const spic_gem_species &sp,
                         const double *q_row_dev, spic_species &out,
                         hipStream_t s);

template <typename T>
inline KGrid<T> kgrid(const spic_grid *g) {
  KGrid<T> k;
  k.nxn = g->nxn; k.nyn = g->nyn; k.nzn = g->nzn;
  k.xStart = g->xStart; k.yStart = g->yStart; k.zStart = g->zStart;
  k.Lx = g->Lx; k.Ly = g->Ly; k.Lz = g->Lz;
  k.invdx = (T)g->invdx; k.invdy = (T)g->invdy; k.invdz = (T)g->invdz;
  k.invVOL = (T)g->invVOL;
  k.xn = (const T *)g->xn; k.yn = (const T *)g->yn; k.zn = (const T *)g->zn;
  return k;
}

template <typename T>
inline KField<T> kfield(const spic_field *f) {
  return KField<T>{(const T *)f->Ex, (const T *)f->Ey, (const T *)f->Ez,
                         (const T *)f->Bxn, (const T *)f->Byn,
                         (const T *)f->Bzn};
}

template <typename T>
inline KPart<T> kpart(const spic_species *sp) {
  return KPart<T>{(T *)sp->x, (T *)sp->y, (T *)sp->z, (T *)sp->u,
                        (T *)sp->v, (T *)sp->w, (const T *)sp->q,
                        (int)sp->nop};
}

inline KMom kmom(const spic_moments *m) {
  KMom k;
  k.m[0] = m->rhon; k.m[1] = m->Jx; k.m[2] = m->Jy; k.m[3] = m->Jz;
  k.m[4] = m->pxx; k.m[5] = m->pxy; k.m[6] = m->pxz;
  k.m[7] = m->pyy; k.m[8] = m->pyz; k.m[9] = m->pzz;
  return k;
}

int subcycle_mode();

// Push constants in the reference's types (src/Mover.cpp:11-12):
//   FPpart dt_sub = (FPpart)dt / ((double)n_sub_cycles);
//   FPpart dto2 = .5 * dt_sub, qomdt2 = qom * dto2 / c;   (qom is FPpart)
// With the GPU mover's sub-cycle loop the step is formed as that mover forms
// it, (FPpart)dt / ((float)n_sub_cycles) (src/gpu/Particles_gpu.cu:319), and
// the push runs n_sub_cycles times.
template <typename T>
inline KPush<T> kpush(const spic_species *sp, const spic_push *p) {
  KPush<T> k;
  const T qom = (T)sp->qom;
  const bool loop = subcycle_mode() == SPIC_SUBCYCLE_LOOP;
  k.nloop = loop ? sp->n_sub_cycles : 1;
  if (loop)
    k.dt_sub = (T)((T)p->dt / (T)(float)sp->n_sub_cycles);
  else
    k.dt_sub = (T)((double)(T)p->dt / ((double)sp->n_sub_cycles));
  k.dto2 = (T)(.5 * (double)k.dt_sub);
  const T qd = qom * k.dto2;  // FPpart * FPpart
  k.qomdt2 = (T)((double)qd / p->c);
  k.niter = sp->NiterMover;
  k.perx = p->PERIODICX != 0;
  k.pery = p->PERIODICY != 0;
  k.perz = p->PERIODICZ != 0;
  return k;
}

// Call protocol shared by the entry points (spic_capi.cpp).
bool async_mode();
bool strict_mode();
unsigned int *status_buffer(hipError_t *err);
int begin_call(hipStream_t s, unsigned int **st);
int read_status(hipStream_t s, unsigned int *st, bool clear);
int end_call(hipStream_t s, unsigned int *st, hipError_t launch);
int check_grid(const spic_grid *g);
int check_species(const spic_species *sp, bool need_q, bool need_move);
int check_field(const spic_field *f);
int check_moments(const spic_moments *m);
int check_push(const spic_push *p);

// Error reporting shared by every entry point: sets the thread-local
// message returned by spic_last_error() and returns `code`.
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

}  // namespace spic
