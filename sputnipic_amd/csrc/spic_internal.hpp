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

// Error reporting shared by every entry point: sets the thread-local
// message returned by spic_last_error() and returns `code`.
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

}  // namespace spic
