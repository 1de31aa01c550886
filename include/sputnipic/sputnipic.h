// sputnipic.h -- the reference's own C++ interface to the particle hot path,
// served by libspic_dropin_{f32,f64}.so on the MI355X.
//
// The five structs below are layout-compatible declarations of the
// reference's (field for field, same types and order), so that an object
// compiled against the reference headers can call this library in place of
// the reference's src/Mover.cpp and src/Particles.cpp:
//   struct particles           include/Particles.h:11-56
//   struct EMfield             include/EMfield.h:8-25
//   struct grid                include/Grid.h:11-63
//   struct interpDensSpecies   include/InterpDensSpecies.h:9-40
//   struct parameters          include/Parameters.h:10-235
// and the four entry points keep the reference's C++ signatures (hence its
// mangled names):
//   mover_PC       include/Mover.h:11-12        src/Mover.cpp:4-162
//   interpP2G      include/Mover.h:19-21        src/Mover.cpp:385-572
//   mover_interp   include/Mover.h:24-26        src/Mover.cpp:575-862
//   particle_sort  include/Particles.h:95-97    src/Particles.cpp:72-234
//
// Precision follows the reference's compile-time choice
// (include/PrecisionTypes.h:6-10): float by default, double when
// SPIC_REF_FP64 is defined; use the library built the same way
// (libspic_dropin_f32.so / libspic_dropin_f64.so).
//
// Residency.  By default every call is coherent with the host arrays, as the
// reference is: particles and fields are uploaded, the result is copied
// back, and interpP2G ADDS its moments into the host arrays.  Callers that
// keep particles on the device between calls switch to resident mode
// (spic_dropin_set_resident(1)): particles then stay in HBM between calls,
// and the caller brackets host access with spic_dropin_sync_to_host() /
// spic_dropin_host_modified().
#ifndef SPUTNIPIC_DROPIN_H
#define SPUTNIPIC_DROPIN_H

#include <string>

#ifndef NS_MAX
#define NS_MAX 6
#endif

#ifdef SPIC_REF_FP64
typedef double FPpart;
typedef double FPfield;
typedef double FPinterp;
#else
typedef float FPpart;
typedef float FPfield;
typedef float FPinterp;
#endif

struct particles {
  int species_ID;
  long npmax;
  long nop;
  int NiterMover;
  int n_sub_cycles;
  int npcel;
  int npcelx;
  int npcely;
  int npcelz;
  FPpart qom;
  FPpart u0, v0, w0;
  FPpart uth, vth, wth;
  FPpart *x;
  FPpart *y;
  FPpart *z;
  FPpart *u;
  FPpart *v;
  FPpart *w;
  FPinterp *q;
  bool *track_particle;
};

struct EMfield {
  FPfield ***Ex;
  FPfield *Ex_flat;
  FPfield ***Ey;
  FPfield *Ey_flat;
  FPfield ***Ez;
  FPfield *Ez_flat;
  FPfield ***Bxn;
  FPfield *Bxn_flat;
  FPfield ***Byn;
  FPfield *Byn_flat;
  FPfield ***Bzn;
  FPfield *Bzn_flat;
};

struct grid {
  int nxc;
  int nxn;
  int nyc;
  int nyn;
  int nzc;
  int nzn;
  double dx;
  double dy;
  double dz;
  FPfield invdx;
  FPfield invdy;
  FPfield invdz;
  FPfield invVOL;
  double xStart, xEnd, yStart, yEnd, zStart, zEnd;
  double Lx, Ly, Lz;
  bool PERIODICX;
  bool PERIODICY;
  bool PERIODICZ;
  FPfield *XN_flat;
  FPfield ***XN;
  FPfield *YN_flat;
  FPfield ***YN;
  FPfield *ZN_flat;
  FPfield ***ZN;
};

struct interpDensSpecies {
  int species_ID;
  FPinterp ***rhon;
  FPinterp *rhon_flat;
  FPinterp ***rhoc;
  FPinterp *rhoc_flat;
  FPinterp ***Jx;
  FPinterp *Jx_flat;
  FPinterp ***Jy;
  FPinterp *Jy_flat;
  FPinterp ***Jz;
  FPinterp *Jz_flat;
  FPinterp ***pxx;
  FPinterp *pxx_flat;
  FPinterp ***pxy;
  FPinterp *pxy_flat;
  FPinterp ***pxz;
  FPinterp *pxz_flat;
  FPinterp ***pyy;
  FPinterp *pyy_flat;
  FPinterp ***pyz;
  FPinterp *pyz_flat;
  FPinterp ***pzz;
  FPinterp *pzz_flat;
};

// include/InterpDensNet.h:11-37: net densities (sum over species).
struct interpDensNet {
  FPinterp ***rhon;
  FPinterp *rhon_flat;
  FPinterp ***rhoc;
  FPinterp *rhoc_flat;
  FPinterp ***Jx;
  FPinterp *Jx_flat;
  FPinterp ***Jy;
  FPinterp *Jy_flat;
  FPinterp ***Jz;
  FPinterp *Jz_flat;
  FPinterp ***pxx;
  FPinterp *pxx_flat;
  FPinterp ***pxy;
  FPinterp *pxy_flat;
  FPinterp ***pxz;
  FPinterp *pxz_flat;
  FPinterp ***pyy;
  FPinterp *pyy_flat;
  FPinterp ***pyz;
  FPinterp *pyz_flat;
  FPinterp ***pzz;
  FPinterp *pzz_flat;
};

// include/InterpDens_aux.h:6-13: the hat densities.  Only chained views
// exist; the library reads and writes the contiguous block behind them, as
// newArr3 (include/Alloc.h:52-60) allocates it.
struct interpDens_aux {
  FPinterp ***rhoh;
  FPinterp ***Jxh;
  FPinterp ***Jyh;
  FPinterp ***Jzh;
};

struct parameters {
  double c;
  double fourpi;
  double dt;
  double th;
  int ncycles;
  int NiterMover;
  int n_sub_cycles;
  bool sort;
  long sort_every_n;
  long sort_cps;
  int number_of_batches;
  long threads_per_block;
  bool track_particles;
  long n_tracked_particles;
  long tracking_start_cycle;
  long tracking_end_cycle;
  long tracking_Lx;
  long tracking_Ly;
  long tracking_Lz;
  std::string tracked_particles_filename;
  double Lx;
  double Ly;
  double Lz;
  int nxc;
  int nyc;
  int nzc;
  double x_center;
  double y_center;
  double z_center;
  double L_square;
  int ns;
  int npcelx[NS_MAX];
  int npcely[NS_MAX];
  int npcelz[NS_MAX];
  long np[NS_MAX];
  long npMax[NS_MAX];
  double NpMaxNpRatio;
  double qom[NS_MAX];
  double rhoINIT[NS_MAX];
  double uth[NS_MAX];
  double vth[NS_MAX];
  double wth[NS_MAX];
  double u0[NS_MAX];
  double v0[NS_MAX];
  double w0[NS_MAX];
  bool PERIODICX;
  bool PERIODICY;
  bool PERIODICZ;
  bool PERIODICX_P;
  bool PERIODICY_P;
  bool PERIODICZ_P;
  int bcPfaceXright;
  int bcPfaceXleft;
  int bcPfaceYright;
  int bcPfaceYleft;
  int bcPfaceZright;
  int bcPfaceZleft;
  int bcPHIfaceXright;
  int bcPHIfaceXleft;
  int bcPHIfaceYright;
  int bcPHIfaceYleft;
  int bcPHIfaceZright;
  int bcPHIfaceZleft;
  int bcEMfaceXright;
  int bcEMfaceXleft;
  int bcEMfaceYright;
  int bcEMfaceYleft;
  int bcEMfaceZright;
  int bcEMfaceZleft;
  double Vinj;
  double delta;
  double B0x;
  double B0y;
  double B0z;
  int Nwaves;
  double dBoB0;
  double pitch_angle;
  double energy;
  bool SmoothON;
  double SmoothValue;
  int SmoothTimes;
  bool verbose;
  bool RESTART;
  bool PoissonCorrection;
  double CGtol;
  double GMREStol;
  int first_cycle_n;
  int FieldOutputCycle;
  int ParticlesOutputCycle;
  int RestartOutputCycle;
  int DiagnosticsOutputCycle;
  std::string inputfile;
  std::string SaveDirName;
  std::string RestartDirName;
  std::string WaveFile;
};

// The reference's entry points.  mover_PC / mover_interp return 0 on success
// like the reference (src/Mover.cpp:161, 861) and a nonzero SPIC_* status on
// failure; interpP2G and particle_sort are void as in the reference and
// report a failure on stderr and through spic_dropin_last_status().
int mover_PC(struct particles *part, struct EMfield *field, struct grid *grd,
             struct parameters *param);
void interpP2G(struct particles *part, struct interpDensSpecies *ids,
               struct grid *grd);
int mover_interp(struct particles *part, struct EMfield *field,
                 struct interpDensSpecies *ids, struct grid *grd,
                 struct parameters *param);
void particle_sort(struct parameters *params, struct particles *particles,
                   struct grid *grd);

// The field-solve consumers of the moments (include/MaxwellSolver.h:12,
// 37): computed on the device in FP64 and written back into the host
// arrays.  With the FP64 build the results are bit-identical to the
// reference's; with the float build the inputs are widened, the arithmetic
// is FP64, and the results are rounded to float once.
void calculateHatDensities(struct interpDens_aux *id_aux, struct interpDensNet *idn,
                           struct interpDensSpecies *ids, struct EMfield *field,
                           struct grid *grd, struct parameters *param);
void MaxwellImage(FPfield *im, FPfield *vector, struct EMfield *field,
                  struct interpDensSpecies *ids, struct grid *grd,
                  struct parameters *param);

// Extensions with no counterpart in the reference.
extern "C" {
// 0 (default): coherent with host arrays on every call; 1: particles stay
// resident in HBM between calls.
int spic_dropin_set_resident(int resident);
// Copy a resident species back into its host arrays.
int spic_dropin_sync_to_host(const struct particles *part);
// Declare that the host arrays of a species were written since the last
// call; the next call re-uploads them.
int spic_dropin_host_modified(const struct particles *part);
// Status of the last call made by this thread (0 = success).
int spic_dropin_last_status(void);
// Release every device mirror held by the library.
int spic_dropin_release_all(void);
}

#endif  // SPUTNIPIC_DROPIN_H
