/*
 * spic.h -- C ABI of libspic.so, the MI355X (gfx950) particle hot path of
 * sputniPIC: the implicit predictor-corrector mover and the ten-moment
 * cloud-in-cell deposit.
 *
 * Every array argument is a DEVICE pointer (HBM-resident); sizes are plain
 * integers; no C++ or framework types cross this boundary.  Every entry
 * point returns an int status (SPIC_OK == 0) and never throws or exits;
 * spic_last_error() gives a thread-local message for the last failure.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   spic_mover_pc       int  mover_PC(particles*, EMfield*, grid*, parameters*)
 *                       include/Mover.h:11-12, src/Mover.cpp:4-162
 *   spic_interp_p2g     void interpP2G(particles*, interpDensSpecies*, grid*)
 *                       include/Mover.h:19-21, src/Mover.cpp:385-572
 *   spic_mover_interp   int  mover_interp(particles*, EMfield*, interpDensSpecies*,
 *                                         grid*, parameters*)
 *                       include/Mover.h:24-26, src/Mover.cpp:575-862
 *   spic_particle_sort  void particle_sort(parameters*, particles*, grid*)
 *                       include/Particles.h:95-97, src/Particles.cpp:72-234
 *   spic_zero_moments   void setZeroSpeciesDensities(interpDensSpecies*, grid*, int)
 *                       include/InterpDensNet.h, src/InterpDensNet.cpp:40-75
 *                       (the ten node moments; rhoc is not on this path)
 * The reference's GPU form of the same path is move_and_interpolate /
 * batch_update_particles (include/gpu/Particles_gpu.cuh:130-138,
 * src/gpu/Particles_gpu.cu:221-307); this ABI supersedes it with
 * HBM-resident particles (no host batching).
 *
 * The C++ layer in include/sputnipic/sputnipic.h keeps the reference's own
 * C++ signatures (host structs) on top of these calls.
 */
#ifndef SPIC_H
#define SPIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPIC_ABI_VERSION 1

/* Status codes. */
enum {
  SPIC_OK = 0,
  SPIC_EINVAL = 1,      /* bad argument (null pointer, size, layout) */
  SPIC_EHIP = 2,        /* HIP runtime error */
  SPIC_EDOMAIN = 3,     /* a particle left the grid by more than one cell:
                           the reference would index outside its arrays */
  SPIC_EUNSUPPORTED = 4,/* a configuration outside what is built (e.g.
                           n_sub_cycles != 1 while sub-cycling is refused,
                           see spic_set_subcycling) */
  SPIC_ENOMEM = 5       /* device allocation failed */
};

/* Precision of particle, field and per-particle arithmetic (the
 * reference's FPpart/FPfield/FPinterp, include/PrecisionTypes.h:6-10).
 * Node moments on the device are always FP64. */
enum { SPIC_F64 = 0, SPIC_F32 = 1 };

/* Grid constants (struct grid, include/Grid.h:11-63).  Cell counts include
 * the two guard cells (nxc = physical + 2) and node counts are nxc + 1, as
 * setGrid makes them (src/Grid.cpp:9-15).  invd* / invVOL are the
 * FPfield values of the reference (exactly representable in double).
 * xn/yn/zn are device arrays of the node coordinates along each axis
 * (length nxn, nyn, nzn; element type per prec): the reference's XN[i][j][k]
 * depends on i only (src/Grid.cpp:53-61), so XN[i][*][*] == xn[i]. */
typedef struct spic_grid {
  int nxc, nyc, nzc;
  int nxn, nyn, nzn;
  double dx, dy, dz;
  double invdx, invdy, invdz, invVOL;
  double xStart, yStart, zStart;
  double Lx, Ly, Lz;
  int prec; /* SPIC_F64 or SPIC_F32 */
  const void *xn, *yn, *zn;
} spic_grid;

/* Node-centred fields (struct EMfield, include/EMfield.h:8-25): device
 * arrays [nxn][nyn][nzn], k fastest, element type per prec. */
typedef struct spic_field {
  const void *Ex, *Ey, *Ez, *Bxn, *Byn, *Bzn;
} spic_field;

/* One species (struct particles, include/Particles.h:11-56): device SoA
 * arrays of npmax elements, element type per prec.  nop particles are
 * moved and deposited; particle_sort covers npmax as the reference does
 * (src/Particles.cpp:178). */
typedef struct spic_species {
  void *x, *y, *z, *u, *v, *w;
  void *q;
  int64_t nop;
  int64_t npmax;
  double qom;
  int NiterMover;
  int n_sub_cycles;
} spic_species;

/* Ten node moments of one species (struct interpDensSpecies,
 * include/InterpDensSpecies.h:9-40): device FP64 arrays [nxn][nyn][nzn]. */
typedef struct spic_moments {
  double *rhon, *Jx, *Jy, *Jz, *pxx, *pxy, *pxz, *pyy, *pyz, *pzz;
} spic_moments;

/* The members of struct parameters the mover reads (include/Parameters.h):
 * dt, c and the field periodicity flags PERIODICX/Y/Z, which the mover uses
 * for the particle boundary conditions (src/Mover.cpp:103-157). */
typedef struct spic_push {
  double dt;
  double c;
  int PERIODICX, PERIODICY, PERIODICZ;
} spic_push;

/* ---------------------------------------------------------------------- */
/* Reference-path entry points (caller's particle order is preserved).    */
/* `stream` is a hipStream_t (NULL: the null stream).  In synchronous mode */
/* (the default, spic_set_async(0)) each call waits for its stream and     */
/* reports out-of-domain particles as SPIC_EDOMAIN; in asynchronous mode   */
/* calls only enqueue work and spic_check_domain() reports later.          */
/* ---------------------------------------------------------------------- */

/* mover_PC: NiterMover predictor-corrector iterations per particle, final
 * update, per-axis periodic / reflecting particle BC.  Updates x..w in
 * place; q is read by nothing here. */
int spic_mover_pc(const spic_grid *grd, const spic_field *field,
                  spic_species *sp, const spic_push *push, void *stream);

/* interpP2G: ADDS (+=) the ten CIC moments of sp's nop particles to mom. */
int spic_interp_p2g(const spic_grid *grd, const spic_species *sp,
                    spic_moments *mom, void *stream);

/* mover_interp: mover_PC then interpP2G, fused per particle (13 words of
 * particle traffic instead of 19). */
int spic_mover_interp(const spic_grid *grd, const spic_field *field,
                      spic_species *sp, spic_moments *mom,
                      const spic_push *push, void *stream);

/* particle_sort: stable bin sort of all 7 arrays over npmax, bins of
 * sort_cps guard-inclusive cells per side with a trailing "lost" bin, in
 * exactly the reference's order (src/Particles.cpp:155-233). */
int spic_particle_sort(const spic_grid *grd, int64_t sort_cps,
                       spic_species *sp, void *stream);

/* particle_sort without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, result and errors as spic_particle_sort; the sort's scratch
 * is a fixed device array of SPIC_SORT_NOALLOC_SCRATCH_MIB MiB per GPU,
 * held from the first call on that GPU until the process ends, which
 * covers about 2.2M particles of FP64 (2.7M FP32); a larger npmax is
 * refused with SPIC_EINVAL.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other threads, other streams) race. */
#define SPIC_SORT_NOALLOC_SCRATCH_MIB 64
int spic_particle_sort_noalloc(const spic_grid *grd, int64_t sort_cps,
                               spic_species *sp, void *stream);

/* The same snapshot words for a species in the caller's order (its nop
 * particles, element type per grd->prec). */
int spic_snapshot_be32(const spic_grid *grd, const spic_species *sp,
                       void *const dst[6], void *stream);

/* dst[(a i + b) mod nop] = src[i] for x..w and q: a bijection when a is
 * coprime to nop.  A utility for measuring on input in random order (the
 * engine's sort on load removes any input order). */
int spic_permute_species(const spic_grid *grd, const spic_species *src,
                         const spic_species *dst, uint64_t a, uint64_t b,
                         void *stream);

/* setZeroSpeciesDensities, node-moment part: zero the ten arrays. */
int spic_zero_moments(const spic_grid *grd, spic_moments *mom, void *stream);

/* applyBCids (src/BC.cpp:362-619) on `ns` consecutive moment sets (1..8):
 * fold the ghost-node moments in place, the periodic axes (push->PERIODICX,
 * Y, Z) first in the order x, y, z (a[1] += a[n-2]; a[n-2] = a[1]), then the
 * other axes in the same order (a[1] *= 2; a[n-2] *= 2), each over the
 * interior of the two other axes.  Bit-identical to the FP64 reference. */
int spic_apply_bc_ids(const spic_grid *grd, const spic_moments *mom, int ns,
                      const spic_push *push, void *stream);

/* sumOverSpecies (src/InterpDensNet.cpp:120-144): net += ids[0] + ... +
 * ids[ns-1] (1..8 sets), added in species order per element, ten node
 * moments.  Bit-identical to the FP64 reference. */
int spic_sum_species(const spic_grid *grd, const spic_moments *net,
                     const spic_moments *ids, int ns, void *stream);

/* ---------------------------------------------------------------------- */
/* Field-solve side of the cycle, on the device, FP64 (grid prec F64):     */
/* the consumers of the deposited moments before and inside the implicit   */
/* field solve.  Bit-identical to the reference's FP64 build.  Arrays are  */
/* FP64 device arrays: nodes [nxn][nyn][nzn], cells [nxc][nyc][nzc]       */
/* (guard-inclusive), k fastest.                                           */
/* ---------------------------------------------------------------------- */

/* The members of struct parameters these operators read
 * (include/Parameters.h): dt, c, th (theta), fourpi, the field periodicity
 * flags and the species' qom. */
typedef struct spic_solve_params {
  double dt, c, th, fourpi;
  int PERIODICX, PERIODICY, PERIODICZ;
  double qom[8];
} spic_solve_params;

/* struct interpDens_aux (include/InterpDens_aux.h:6-13): rho hat on the
 * cells, J hat on the nodes. */
typedef struct spic_hat {
  double *rhoh;
  double *Jxh, *Jyh, *Jzh;
} spic_hat;

/* Workspace bytes that spic_hat_densities and spic_maxwell_image need for
 * ns species on grd. */
size_t spic_field_workspace_bytes(const spic_grid *grd, int ns);

/* Ghost update of one scalar array, in place: kind 0 applyBCscalarDensN
 * (nodes, src/BC.cpp:11-123), 1 applyBCscalarFieldNzero (nodes, zeroed
 * non-periodic faces, src/BC.cpp:242-352), 2 applyBCscalarDensC (cells,
 * src/BC.cpp:623-735), 3 applyBCscalarFieldC (cells, src/BC.cpp:737-850). */
int spic_bc_scalar(const spic_grid *grd, double *a, int kind,
                   const spic_solve_params *p, void *stream);

/* The driver's step before the solver (src/sputniPIC_CPU.cpp:276-277):
 * applyBCscalarDensN on the net rhon (in place), then rhoc =
 * interpN2Cinterp(rhon) (src/Grid.cpp:164-185) on the interior cells. */
int spic_rhoc_from_rhon(const spic_grid *grd, double *rhon, double *rhoc,
                        const spic_solve_params *p, void *stream);

/* calculateHatDensities (src/MaxwellSolver.cpp:465-604): from the ns
 * (1..8) species' currents and pressure tensors, the net rhoc and B,
 * writes J hat (nodes) and rho hat (cells), ghosts included. */
int spic_hat_densities(const spic_grid *grd, const spic_field *field,
                       const spic_moments *ids, int ns, const double *rhoc,
                       const spic_solve_params *p, const spic_hat *out,
                       void *work, size_t work_bytes, void *stream);

/* MaxwellImage (src/MaxwellSolver.cpp:95-257): im = A vec for the GMRES
 * field solve; vec and im hold the interior nodes' (x, y, z) components
 * interleaved, i slowest (solver2phys / phys2solver, src/Basic.cpp:
 * 365-392), 3 (nxn-2)(nyn-2)(nzn-2) values; ns (0..8) species' rhon. */
int spic_maxwell_image(const spic_grid *grd, const spic_field *field,
                       const spic_moments *ids, int ns,
                       const spic_solve_params *p, const double *vec,
                       double *im, void *work, size_t work_bytes,
                       void *stream);


/* ---------------------------------------------------------------------- */
/* Tile-resident engine: the whole particle phase of a cycle -- mover_PC   */
/* then interpP2G for every species, the loop of                          */
/* src/sputniPIC_CPU.cpp:232-256 -- in one launch over HBM-resident        */
/* particles binned by tile.  The engine owns its particle storage: per    */
/* species two buffers with one segment per tile (its particles plus free  */
/* slots) and an overflow region, and a particle id per slot, so particles */
/* can be copied out in the order they were loaded.  Each phase writes     */
/* every particle into its destination tile's segment of the other buffer, */
/* so the binning stays current without a separate sort.  A sort (on load, */
/* every `sort_every` phases, or on request) rebuilds the segments.        */
/* Neither changes a particle's trajectory, only the summation order of    */
/* the moments.                                                            */
/* ---------------------------------------------------------------------- */
typedef struct spic_engine spic_engine;

#define SPIC_SORT_TILE 0
#define SPIC_SORT_RADIX 1
#define SPIC_POS_FLOAT 1
#define SPIC_POS_FIXED 2
#define SPIC_ORDER_LOAD 1
#define SPIC_ORDER_ENGINE 2

typedef struct spic_engine_opts {
  int tile_x, tile_y, tile_z; /* cells per tile: 4x4x4 (default, 0s) or 8x4x4 */
  int sort_every;             /* rebuild the segments every N phases; <= 0:
                                 only on load and spic_engine_sort */
  int lds_partials;           /* precision of the per-tile LDS moment block
                                 of an FP32 engine: 0 = default (FP64),
                                 1 = FP32, 2 = FP64.  An FP64 engine always
                                 uses FP64. */
  int slack_pct;              /* free slots per segment after a sort, in
                                 percent of the mean particles per tile
                                 (at least 16 slots); 0 = default (12) */
  int set_moments;            /* 1: each phase WRITES the ten moments of
                                 every loaded species (the deposit of this
                                 phase alone, zero where no particle
                                 deposits), which spares the caller the
                                 zeroing that precedes interpP2G
                                 (src/sputniPIC_CPU.cpp:247; rhoc is not
                                 touched); 0 (default): ADD them, as
                                 interpP2G does */
  int sort_mode;              /* re-sorts (sort_every, spic_engine_sort):
                                 SPIC_SORT_TILE (0, default) counts within
                                 each tile; SPIC_SORT_RADIX a radix sort
                                 of all particles by (tile, rank within
                                 cell, cell).  Both give the same segments
                                 and the same order within each tile,
                                 except that the tile sort puts particles
                                 arriving from other tiles after the
                                 tile's own.  Trajectories do not depend
                                 on the mode; loads always radix sort */
  int positions;              /* storage of positions: SPIC_POS_FLOAT the
                                 reference's (FPpart lengths), SPIC_POS_FIXED
                                 int32 fixed point in cell units (FP32
                                 engines only: 2^-k cell resolution, k <= 24
                                 chosen so the grid plus 64 cells fits, e.g.
                                 2^-23 cells on 128 cells, against float's
                                 2^-24 relative; exact wraps and
                                 reflections).  0 (default): fixed point for
                                 an FP32 engine created while strict
                                 arithmetic is off, else float.  Strict
                                 arithmetic refuses a fixed-point engine */
  int order;                  /* SPIC_ORDER_LOAD (0, default): each particle
                                 carries its load index (4 bytes read and
                                 written per phase) so that store and the
                                 snapshot return the load order;
                                 SPIC_ORDER_ENGINE: no index, store and the
                                 snapshot return the particles in the
                                 engine's own order (tile by tile), as the
                                 reference's particle_sort reorders its
                                 arrays.  Trajectories and moments do not
                                 depend on it */
  int overflow_pct;           /* overflow slots of each buffer in percent of
                                 the species' particles: 0 = default (100:
                                 every particle fits, none can be lost);
                                 1..99 bound the memory (at least 1024 slots)
                                 and a phase whose leaving particles do not
                                 fit their segments and the region reports
                                 SPIC_ENOMEM (its output is then invalid:
                                 load the species again) */
  int particles_per_thread;   /* particles each thread pushes and bins per
                                 batch of the phase kernel before the
                                 batch's bucket step: 0 or 1 (default), or
                                 2 (FP32 engines with fixed-point positions;
                                 the buckets double, two workgroups per CU
                                 instead of three).  Trajectories do not
                                 depend on it.  Batches (spic_batch_create)
                                 and spic_engine_deposit use one */
} spic_engine_opts;

typedef struct spic_engine_stats {
  int64_t phases;       /* phases run */
  int64_t sorts;        /* species sorts run */
  int ntiles;           /* tiles of the grid */
  int tile_x, tile_y, tile_z;
  size_t device_bytes;  /* device memory held by the engine */
  /* cumulative over all phases: predictor-corrector stencil gathers and
   * particle deposits served from the LDS-staged tile, and those that fell
   * outside it (a particle more than one cell out of its tile, or in an
   * overflow region) and went to global memory */
  int64_t lds_gathers, global_gathers;
  int64_t lds_deposits, global_deposits;
  int64_t overflow;     /* particles now in overflow regions (all species) */
  int positions;        /* SPIC_POS_FLOAT or SPIC_POS_FIXED */
  size_t workspace_bytes; /* of device_bytes: the sort workspace, kept
                             between sorts and grown on demand */
} spic_engine_stats;

/* Create an engine for `ns` species (1..8) on `grd` (its node tables are
 * copied).  opts may be NULL. */
int spic_engine_create(spic_engine **eng, const spic_grid *grd, int ns,
                       const spic_engine_opts *opts);
int spic_engine_destroy(spic_engine *eng);
/* Copy species `is` in (device arrays, caller order; nop particles, qom,
 * NiterMover, n_sub_cycles) and sort it into tile segments. */
int spic_engine_load(spic_engine *eng, int is, const spic_species *sp,
                     void *stream);
/* Copy species `is` out into sp's arrays in the order it was loaded
 * (sp->nop must equal the loaded count). */
int spic_engine_store(spic_engine *eng, int is, spic_species *sp,
                      void *stream);
/* spic_engine_load / spic_engine_store with the caller's arrays in
 * precision `prec` (SPIC_F64 or SPIC_F32), which may differ from the
 * engine's: an FP32 engine loads FP64 arrays (positions converted from the
 * FP64 values, so fixed-point storage keeps their precision; velocities and
 * charges rounded to float) and stores into FP64 arrays.  An FP64 engine
 * takes FP64 arrays only. */
int spic_engine_load_as(spic_engine *eng, int is, const spic_species *sp,
                        int prec, void *stream);
int spic_engine_store_as(spic_engine *eng, int is, spic_species *sp, int prec,
                         void *stream);
/* One particle phase: for every species, mover_PC then interpP2G, ADDING
 * the moments into mom[is] (ns moment sets; zero them first as the
 * reference driver does, src/sputniPIC_CPU.cpp:247). */
int spic_engine_phase(spic_engine *eng, const spic_field *field,
                      const spic_moments *mom, const spic_push *push,
                      void *stream);
/* interpP2G of every loaded species from its binned particles, without a
 * push (src/Mover.cpp:385-572): the moments are added into mom[is] (or
 * written, with set_moments).  Not a phase for the sort schedule. */
int spic_engine_deposit(spic_engine *eng, const spic_moments *mom, void *stream);
/* Particle snapshot of species `is` in load order: x y z u v w as
 * big-endian IEEE float32 words (values rounded to float) into six device
 * arrays of nop uint32 -- the datasets HDF5_Write_Particles writes per
 * species as H5T_IEEE_F32BE (src/RW_IO.cpp:376-460).  One kernel; the
 * caller copies the words out (spic_memcpy_d2h_async on a side stream lets
 * the next phase run meanwhile). */
int spic_engine_snapshot_be32(spic_engine *eng, int is, void *const dst[6],
                              void *stream);
/* Rebuild every species' tile segments now. */
int spic_engine_sort(spic_engine *eng, void *stream);
int spic_engine_get_stats(const spic_engine *eng, spic_engine_stats *st);

/* ---------------------------------------------------------------------- */
/* Batches: up to SPIC_MAX_INSTANCES independent PIC instances (engines    */
/* with their own grids, fields, species and moments) advanced together,  */
/* one launch per kernel and phase: every instance's tiles form one grid   */
/* of the tile kernel, and the count reset, overflow pass and node reduce  */
/* are one launch each.  The reference runs independent simulations as    */
/* separate processes; this is the same work in one stream of launches.    */
/* Each instance's results equal those of spic_engine_phase on its own.    */
/* Members share precision, tile shape, LDS partials and sort mode.  The  */
/* moment and push descriptors are copied at creation (the arrays they     */
/* point to must stay allocated).  Load every species before creating the  */
/* batch; afterwards advance the members only through spic_batch_phase     */
/* (a load re-reads the members at the next phase, outside graph capture). */
/* ---------------------------------------------------------------------- */
#define SPIC_MAX_INSTANCES 16
typedef struct spic_batch spic_batch;

typedef struct spic_batch_member {
  spic_engine *engine;
  const spic_field *field;
  const spic_moments *moments; /* one set per engine species */
  const spic_push *push;
} spic_batch_member;

/* Batch flags: SPIC_BATCH_SET_MOMENTS makes each phase WRITE the bound
 * moment sets (as spic_engine_opts.set_moments does for one engine). */
#define SPIC_BATCH_SET_MOMENTS 1

/* Bind n (1..SPIC_MAX_INSTANCES) distinct engines.  sort_every > 0 sorts
 * every species of every member before every sort_every-th phase; the
 * members' own sort_every and set_moments do not apply to batch phases. */
int spic_batch_create(spic_batch **batch, const spic_batch_member *members,
                      int n, int sort_every, int flags);
/* One particle phase of every member: mover_PC then interpP2G of each
 * species, ADDING the moments into the bound moment sets (or writing
 * them, with SPIC_BATCH_SET_MOMENTS). */
int spic_batch_phase(spic_batch *batch, void *stream);
int spic_batch_destroy(spic_batch *batch);

/* ---------------------------------------------------------------------- */
/* Synthetic Harris-sheet (GEM) input on the device, for runs too large    */
/* for a CPU initial condition.  Restates initGEM (src/IC.cpp:38-177) and  */
/* interpN2Crho (src/InterpDensSpecies.cpp:64-77) with one change: the     */
/* Maxwellian draws come from a counter-based generator (SplitMix64 of     */
/* seed, species, particle, draw) instead of glibc rand(), so particles    */
/* are generated independently.  Positions, charges and fields follow the  */
/* reference formulas; the sub-lattice is cell-major as the reference's.   */
/* ---------------------------------------------------------------------- */
typedef struct spic_gem_params {
  double B0x, B0y, B0z, delta; /* Harris sheet (inputfiles/GEM_3D.inp) */
  double pertX;                /* X-point perturbation amplitude; 0.4 in
                                  src/IC.cpp:11 */
  double fourpi;               /* 12.5663706144 (src/RW_IO.cpp:39) */
} spic_gem_params;

typedef struct spic_gem_species {
  int species;      /* index, part of the random counter */
  uint32_t seed;    /* run seed, part of the random counter */
  int sheet;        /* 1: current-sheet density profile (species 0,1 of the
                       reference); 0: uniform background */
  int npcelx, npcely, npcelz;
  double qom, rhoINIT;
  double u0, v0, w0, uth, vth, wth;
} spic_gem_species;

/* Fill the six node fields (all nodes, ghosts included). */
int spic_gem_fields(const spic_grid *grd, const spic_gem_params *gem,
                    spic_field *out, void *stream);
/* Fill out->npmax particles of one species: npcelx*npcely*npcely per
 * physical cell in cell-major order (kk < npcely, src/IC.cpp:128).
 * out->npmax must equal that count times the physical cell count. */
int spic_gem_particles(const spic_grid *grd, const spic_gem_params *gem,
                       const spic_gem_species *sp, spic_species *out,
                       void *stream);

/* ---------------------------------------------------------------------- */
/* HIP event timing on a stream (for callers without HIP headers).        */
/* ---------------------------------------------------------------------- */
int spic_event_create(void **event);
int spic_event_destroy(void *event);
int spic_event_record(void *event, void *stream);
int spic_event_elapsed_ms(float *ms, void *start, void *end);

/* ---------------------------------------------------------------------- */
/* Modes, errors, device memory helpers (thin wrappers over HIP so callers */
/* without HIP headers -- the C++ layer, ctypes -- can stage data).        */
/* ---------------------------------------------------------------------- */
const char *spic_last_error(void);
int spic_abi_version(void);

/* Synchronous (0, default) or asynchronous (1) reporting, see above. */
int spic_set_async(int async_mode);
/* Arithmetic mode.  0 (default): FMA contraction allowed (production).
 * 1: no FMA contraction, every operation rounded in the reference's order
 * and operand types, so the particle state matches the CPU reference bit
 * for bit. */
int spic_set_strict_fp(int strict);
/* Sub-cycling of the mover, for species with n_sub_cycles != 1.  The
 * reference's two movers differ here: the CPU mover has no sub-cycle loop
 * and pushes once with dt / n_sub_cycles (src/Mover.cpp:11-12); the GPU
 * mover pushes n_sub_cycles times with that step, applying the particle BC
 * after each push (src/gpu/Particles_gpu.cu:319-437).
 *   SPIC_SUBCYCLE_REFUSE (default): n_sub_cycles != 1 is refused with
 *     SPIC_EUNSUPPORTED;
 *   SPIC_SUBCYCLE_LOOP: the GPU mover's loop (step (FPpart)dt /
 *     n_sub_cycles, in the particle precision as that mover forms it);
 *   SPIC_SUBCYCLE_SINGLE: the CPU mover's single push of dt / n_sub_cycles.
 * Both BC tests of an axis are applied after each push, as the CPU mover
 * does (the GPU mover's `else if` gives the same result for any
 * displacement shorter than the box).  Read whenever push constants are
 * formed: by each mover call and engine phase, and by a batch when it is
 * created or its tables are rebuilt.  n_sub_cycles < 1 is always refused. */
#define SPIC_SUBCYCLE_REFUSE 0
#define SPIC_SUBCYCLE_LOOP 1
#define SPIC_SUBCYCLE_SINGLE 2
int spic_set_subcycling(int mode);
/* In asynchronous mode: wait for `stream`, return SPIC_EDOMAIN if any
 * out-of-domain particle was seen since the last check, and clear it. */
int spic_check_domain(void *stream);

/* Testing hook (no counterpart in the reference): the engine device
 * allocation `after` allocations from now (0: the next one) fails as an
 * out-of-memory error would, without allocating; negative switches it off.
 * Lets a test reach the engine's allocation-failure paths. */
int spic_test_fail_engine_alloc(int after);

int spic_device_count(int *count);
int spic_malloc(void **dptr, size_t bytes);
int spic_free(void *dptr);
int spic_memcpy_h2d(void *dst, const void *src, size_t bytes);
int spic_memcpy_d2h(void *dst, const void *src, size_t bytes);
int spic_memcpy_d2d(void *dst, const void *src, size_t bytes);
int spic_memset(void *dptr, int value, size_t bytes);
int spic_synchronize(void);
/* Pinned host memory (hipHostMalloc) and an asynchronous device-to-host
 * copy, for snapshots copied out on a side stream. */
int spic_host_alloc(void **hptr, size_t bytes);
int spic_host_free(void *hptr);
int spic_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream);
/* Make `stream` wait for `event` (recorded on another stream). */
int spic_stream_wait_event(void *stream, void *event);
int spic_stream_create(void **stream);
int spic_stream_destroy(void *stream);
int spic_stream_synchronize(void *stream);

/* HIP graphs: capture a run of calls on a created stream (in async mode:
 * spic_set_async(1)) once and replay it with one launch.  Every entry point
 * above that takes a stream can be captured once its buffers exist (an
 * engine after spic_engine_load, the moment arrays after allocation); none
 * allocates or synchronizes in async mode.  A replay repeats the captured
 * work on the same device buffers: an engine phase captures which of its
 * two particle buffers it reads, so capture an even number of phases and
 * sorts to replay a run more than once.  Replays do not advance the host
 * counters of spic_engine_get_stats. */
typedef struct spic_graph spic_graph;
int spic_graph_begin(void *stream);
int spic_graph_end(void *stream, spic_graph **graph);
int spic_graph_launch(spic_graph *graph, void *stream);
int spic_graph_destroy(spic_graph *graph);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* SPIC_H */
