// spic_dropin.cpp -- the reference's C++ entry points (mover_PC, interpP2G,
// mover_interp, particle_sort, and the field-solve side's
// calculateHatDensities and MaxwellImage) over the C ABI of libspic.so, for host
// structs laid out as the reference's (include/sputnipic/sputnipic.h).
// Built twice by the Makefile: libspic_dropin_f32.so (the reference's
// default float precision) and libspic_dropin_f64.so (-DSPIC_REF_FP64).
//
// Device mirrors are created lazily on first use and keyed by the host
// array they mirror (grid: XN_flat, fields: Ex_flat, species: x, moments:
// rhon_flat); the caller owns and never loses its host memory.  Calls are
// thread-safe: the registry is mutex-protected and every calling thread
// uses its own stream, as the reference calls interpP2G from an OpenMP team
// with one thread per species (src/sputniPIC_CPU.cpp:250-256).
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "spic.h"
#include "sputnipic/sputnipic.h"

namespace {

constexpr int kPrec = sizeof(FPpart) == 8 ? SPIC_F64 : SPIC_F32;
static_assert(sizeof(FPpart) == sizeof(FPfield) && sizeof(FPpart) == sizeof(FPinterp),
              "the drop-in is built for one precision throughout");

thread_local int t_status = SPIC_OK;
thread_local void *t_stream = nullptr;
std::mutex g_mu;
int g_resident = 0;

struct GridMirror {
  spic_grid g;
  void *xn = nullptr, *yn = nullptr, *zn = nullptr;
  spic_engine *eng = nullptr;  // single-species engine for the deposits
  bool eng_failed = false;
};
struct FieldMirror {
  void *d[6] = {};
  size_t nn = 0;
};
struct PartMirror {
  void *d[7] = {};
  long npmax = 0;
  bool device_current = false;  // device holds the newest particle state
  bool host_current = true;      // host arrays hold the newest state
};
struct MomMirror {
  double *d[10] = {};
  size_t nn = 0;
};

std::map<const void *, GridMirror> g_grids;
std::map<const void *, FieldMirror> g_fields;
std::map<const void *, PartMirror> g_parts;
std::map<const void *, MomMirror> g_moms;

int report(const char *fn, int rc) {
  t_status = rc;
  if (rc != SPIC_OK)
    std::fprintf(stderr, "sputnipic drop-in: %s: %s\n", fn, spic_last_error());
  return rc;
}

void *stream() {
  if (!t_stream && spic_stream_create(&t_stream) != SPIC_OK) t_stream = nullptr;
  return t_stream;
}

// The reference's node arrays depend on one index each
// (src/Grid.cpp:53-61); the device uses the 1-D tables.
int grid_mirror(const grid *grd, GridMirror **out) {
  auto it = g_grids.find(grd->XN_flat);
  if (it != g_grids.end()) {
    *out = &it->second;
    return SPIC_OK;
  }
  const int nxn = grd->nxn, nyn = grd->nyn, nzn = grd->nzn;
  std::vector<FPfield> xn(nxn), yn(nyn), zn(nzn);
  for (int i = 0; i < nxn; ++i) xn[i] = grd->XN_flat[(long)i * nyn * nzn];
  for (int j = 0; j < nyn; ++j) yn[j] = grd->YN_flat[(long)j * nzn];
  for (int k = 0; k < nzn; ++k) zn[k] = grd->ZN_flat[k];
  for (int i = 0; i < nxn; ++i)
    for (int j = 0; j < nyn; ++j)
      for (int k = 0; k < nzn; ++k) {
        const long n = ((long)i * nyn + j) * nzn + k;
        if (grd->XN_flat[n] != xn[i] || grd->YN_flat[n] != yn[j] ||
            grd->ZN_flat[n] != zn[k])
          return SPIC_EINVAL;  // not the separable lattice setGrid builds
      }
  GridMirror m;
  int rc;
  if ((rc = spic_malloc(&m.xn, sizeof(FPfield) * nxn)) ||
      (rc = spic_malloc(&m.yn, sizeof(FPfield) * nyn)) ||
      (rc = spic_malloc(&m.zn, sizeof(FPfield) * nzn)) ||
      (rc = spic_memcpy_h2d(m.xn, xn.data(), sizeof(FPfield) * nxn)) ||
      (rc = spic_memcpy_h2d(m.yn, yn.data(), sizeof(FPfield) * nyn)) ||
      (rc = spic_memcpy_h2d(m.zn, zn.data(), sizeof(FPfield) * nzn)))
    return rc;
  spic_grid &g = m.g;
  g.nxc = grd->nxc; g.nyc = grd->nyc; g.nzc = grd->nzc;
  g.nxn = nxn; g.nyn = nyn; g.nzn = nzn;
  g.dx = grd->dx; g.dy = grd->dy; g.dz = grd->dz;
  g.invdx = grd->invdx; g.invdy = grd->invdy; g.invdz = grd->invdz;
  g.invVOL = grd->invVOL;
  g.xStart = grd->xStart; g.yStart = grd->yStart; g.zStart = grd->zStart;
  g.Lx = grd->Lx; g.Ly = grd->Ly; g.Lz = grd->Lz;
  g.prec = kPrec;
  g.xn = m.xn; g.yn = m.yn; g.zn = m.zn;
  auto res = g_grids.emplace(grd->XN_flat, m);
  *out = &res.first->second;
  return SPIC_OK;
}

// Fields are uploaded on every call: the reference's field solver rewrites
// them on the host between cycles (src/sputniPIC_CPU.cpp:288-289).
int field_upload(const EMfield *f, size_t nn, spic_field *out) {
  FieldMirror &m = g_fields[f->Ex_flat];
  int rc;
  if (!m.d[0]) {
    for (int a = 0; a < 6; ++a)
      if ((rc = spic_malloc(&m.d[a], sizeof(FPfield) * nn))) return rc;
    m.nn = nn;
  }
  const FPfield *src[6] = {f->Ex_flat, f->Ey_flat, f->Ez_flat,
                           f->Bxn_flat, f->Byn_flat, f->Bzn_flat};
  for (int a = 0; a < 6; ++a)
    if ((rc = spic_memcpy_h2d(m.d[a], src[a], sizeof(FPfield) * nn))) return rc;
  out->Ex = m.d[0]; out->Ey = m.d[1]; out->Ez = m.d[2];
  out->Bxn = m.d[3]; out->Byn = m.d[4]; out->Bzn = m.d[5];
  return SPIC_OK;
}

FPpart *const *host_arrays(const particles *p, FPpart *arr[7]) {
  arr[0] = p->x; arr[1] = p->y; arr[2] = p->z;
  arr[3] = p->u; arr[4] = p->v; arr[5] = p->w;
  arr[6] = (FPpart *)p->q;
  return arr;
}

// Device copy of a species, current with respect to the host arrays.
int part_mirror(const particles *p, PartMirror **out, spic_species *sp) {
  PartMirror &m = g_parts[p->x];
  int rc;
  if (!m.d[0] || m.npmax != p->npmax) {
    for (int a = 0; a < 7; ++a)
      if (m.d[a]) spic_free(m.d[a]);
    for (int a = 0; a < 7; ++a)
      if ((rc = spic_malloc(&m.d[a], sizeof(FPpart) * (size_t)p->npmax))) return rc;
    m.npmax = p->npmax;
    m.device_current = false;
    m.host_current = true;
  }
  // coherent mode: the host arrays are the reference copy on every call
  if (!m.device_current || !g_resident) {
    FPpart *arr[7];
    host_arrays(p, arr);
    for (int a = 0; a < 7; ++a)
      if ((rc = spic_memcpy_h2d(m.d[a], arr[a], sizeof(FPpart) * (size_t)p->npmax)))
        return rc;
    m.device_current = true;
  }
  sp->x = m.d[0]; sp->y = m.d[1]; sp->z = m.d[2];
  sp->u = m.d[3]; sp->v = m.d[4]; sp->w = m.d[5]; sp->q = m.d[6];
  sp->nop = p->nop;
  sp->npmax = p->npmax;
  sp->qom = p->qom;
  sp->NiterMover = p->NiterMover;
  sp->n_sub_cycles = p->n_sub_cycles;
  *out = &m;
  return SPIC_OK;
}

// After a call that changed particles: coherent mode copies them back.
int part_written(const particles *p, PartMirror *m, bool q_too) {
  m->host_current = false;
  if (g_resident) return SPIC_OK;
  FPpart *arr[7];
  host_arrays(p, arr);
  const int na = q_too ? 7 : 6;
  for (int a = 0; a < na; ++a) {
    int rc = spic_memcpy_d2h(arr[a], m->d[a], sizeof(FPpart) * (size_t)p->npmax);
    if (rc) return rc;
  }
  m->host_current = true;
  return SPIC_OK;
}

int mom_mirror(const interpDensSpecies *ids, size_t nn, MomMirror **out,
               spic_moments *mom) {
  MomMirror &m = g_moms[ids->rhon_flat];
  int rc;
  if (!m.d[0]) {
    for (int a = 0; a < 10; ++a)
      if ((rc = spic_malloc((void **)&m.d[a], sizeof(double) * nn))) return rc;
    m.nn = nn;
  }
  mom->rhon = m.d[0]; mom->Jx = m.d[1]; mom->Jy = m.d[2]; mom->Jz = m.d[3];
  mom->pxx = m.d[4]; mom->pxy = m.d[5]; mom->pxz = m.d[6];
  mom->pyy = m.d[7]; mom->pyz = m.d[8]; mom->pzz = m.d[9];
  *out = &m;
  return SPIC_OK;
}

// interpP2G adds into the caller's arrays (src/Mover.cpp:469-570): the
// device sum of this call is added to each host node value.
int mom_add_to_host(const interpDensSpecies *ids, MomMirror *m) {
  FPinterp *dst[10] = {ids->rhon_flat, ids->Jx_flat,  ids->Jy_flat,  ids->Jz_flat,
                       ids->pxx_flat,  ids->pxy_flat, ids->pxz_flat, ids->pyy_flat,
                       ids->pyz_flat,  ids->pzz_flat};
  std::vector<double> buf(m->nn);
  for (int a = 0; a < 10; ++a) {
    int rc = spic_memcpy_d2h(buf.data(), m->d[a], sizeof(double) * m->nn);
    if (rc) return rc;
    for (size_t n = 0; n < m->nn; ++n) dst[a][n] = (FPinterp)(dst[a][n] + buf[n]);
  }
  return SPIC_OK;
}

spic_push push_of(const parameters *param) {
  spic_push p;
  p.dt = param->dt;
  p.c = param->c;
  p.PERIODICX = param->PERIODICX;
  p.PERIODICY = param->PERIODICY;
  p.PERIODICZ = param->PERIODICZ;
  return p;
}

size_t nodes(const grid *grd) { return (size_t)grd->nxn * grd->nyn * grd->nzn; }

// The grid's single-species engine: interpP2G and mover_interp bin the
// caller's particles into it and deposit through the LDS tiles of the
// binned kernels instead of the caller-order kernels' global atomics.  Each
// call loads the particles anew (they are the caller's, in the caller's
// order) and finishes with a blocking copy of the moments, so calls do not
// overlap in it.  Without an engine (its creation failed) the calls use the
// caller-order kernels.
spic_engine *grid_engine(GridMirror *gm) {
  if (!gm->eng && !gm->eng_failed) {
    spic_engine_opts o;
    std::memset(&o, 0, sizeof(o));
    // the caller's arrays hold the particles between calls in the
    // reference's own representation, so the engine keeps it too
    o.positions = SPIC_POS_FLOAT;
    if (spic_engine_create(&gm->eng, &gm->g, 1, &o) != SPIC_OK) {
      gm->eng = nullptr;
      gm->eng_failed = true;
    }
  }
  return gm->eng;
}

}  // namespace

int mover_PC(struct particles *part, struct EMfield *field, struct grid *grd,
             struct parameters *param) {
  std::lock_guard<std::mutex> lock(g_mu);
  GridMirror *gm;
  PartMirror *pm;
  spic_field f;
  spic_species sp;
  int rc = grid_mirror(grd, &gm);
  if (rc == SPIC_OK) rc = field_upload(field, nodes(grd), &f);
  if (rc == SPIC_OK) rc = part_mirror(part, &pm, &sp);
  if (rc == SPIC_OK) {
    const spic_push p = push_of(param);
    rc = spic_mover_pc(&gm->g, &f, &sp, &p, stream());
  }
  if (rc == SPIC_OK) rc = part_written(part, pm, false);
  return report("mover_PC", rc);
}

void interpP2G(struct particles *part, struct interpDensSpecies *ids,
               struct grid *grd) {
  std::lock_guard<std::mutex> lock(g_mu);
  GridMirror *gm;
  PartMirror *pm;
  MomMirror *mm;
  spic_species sp;
  spic_moments mom;
  int rc = grid_mirror(grd, &gm);
  if (rc == SPIC_OK) rc = part_mirror(part, &pm, &sp);
  if (rc == SPIC_OK) rc = mom_mirror(ids, nodes(grd), &mm, &mom);
  if (rc == SPIC_OK) rc = spic_zero_moments(&gm->g, &mom, stream());
  if (rc == SPIC_OK) {
    spic_engine *e = grid_engine(gm);
    if (e) {
      rc = spic_engine_load(e, 0, &sp, stream());
      if (rc == SPIC_OK) rc = spic_engine_deposit(e, &mom, stream());
    } else {
      rc = spic_interp_p2g(&gm->g, &sp, &mom, stream());
    }
  }
  if (rc == SPIC_OK) rc = mom_add_to_host(ids, mm);
  report("interpP2G", rc);
}

int mover_interp(struct particles *part, struct EMfield *field,
                 struct interpDensSpecies *ids, struct grid *grd,
                 struct parameters *param) {
  std::lock_guard<std::mutex> lock(g_mu);
  GridMirror *gm;
  PartMirror *pm;
  MomMirror *mm;
  spic_field f;
  spic_species sp;
  spic_moments mom;
  int rc = grid_mirror(grd, &gm);
  if (rc == SPIC_OK) rc = field_upload(field, nodes(grd), &f);
  if (rc == SPIC_OK) rc = part_mirror(part, &pm, &sp);
  if (rc == SPIC_OK) rc = mom_mirror(ids, nodes(grd), &mm, &mom);
  if (rc == SPIC_OK) rc = spic_zero_moments(&gm->g, &mom, stream());
  if (rc == SPIC_OK) {
    const spic_push p = push_of(param);
    spic_engine *e = grid_engine(gm);
    if (e) {
      rc = spic_engine_load(e, 0, &sp, stream());
      if (rc == SPIC_OK) rc = spic_engine_phase(e, &f, &mom, &p, stream());
      if (rc == SPIC_OK) rc = spic_engine_store(e, 0, &sp, stream());
    } else {
      rc = spic_mover_interp(&gm->g, &f, &sp, &mom, &p, stream());
    }
  }
  if (rc == SPIC_OK) rc = part_written(part, pm, false);
  if (rc == SPIC_OK) rc = mom_add_to_host(ids, mm);
  return report("mover_interp", rc);
}

void particle_sort(struct parameters *params, struct particles *particles,
                   struct grid *grd) {
  std::lock_guard<std::mutex> lock(g_mu);
  GridMirror *gm;
  PartMirror *pm;
  spic_species sp;
  int rc = grid_mirror(grd, &gm);
  if (rc == SPIC_OK) rc = part_mirror(particles, &pm, &sp);
  if (rc == SPIC_OK) rc = spic_particle_sort(&gm->g, params->sort_cps, &sp, stream());
  if (rc == SPIC_OK) rc = part_written(particles, pm, true);
  report("particle_sort", rc);
}

extern "C" {

int spic_dropin_set_resident(int resident) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_resident = resident ? 1 : 0;
  return SPIC_OK;
}

int spic_dropin_sync_to_host(const struct particles *part) {
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_parts.find(part->x);
  if (it == g_parts.end() || it->second.host_current) return SPIC_OK;
  FPpart *arr[7];
  host_arrays(part, arr);
  for (int a = 0; a < 7; ++a) {
    int rc = spic_memcpy_d2h(arr[a], it->second.d[a],
                             sizeof(FPpart) * (size_t)part->npmax);
    if (rc) return report("spic_dropin_sync_to_host", rc);
  }
  it->second.host_current = true;
  return SPIC_OK;
}

int spic_dropin_host_modified(const struct particles *part) {
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_parts.find(part->x);
  if (it != g_parts.end()) it->second.device_current = false;
  return SPIC_OK;
}

int spic_dropin_last_status(void) { return t_status; }

int spic_dropin_release_all(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  spic_synchronize();
  for (auto &kv : g_grids) {
    if (kv.second.eng) spic_engine_destroy(kv.second.eng);
    spic_free(kv.second.xn);
    spic_free(kv.second.yn);
    spic_free(kv.second.zn);
  }
  for (auto &kv : g_fields)
    for (void *p : kv.second.d) spic_free(p);
  for (auto &kv : g_parts)
    for (void *p : kv.second.d) spic_free(p);
  for (auto &kv : g_moms)
    for (double *p : kv.second.d) spic_free(p);
  g_grids.clear();
  g_fields.clear();
  g_parts.clear();
  g_moms.clear();
  return SPIC_OK;
}

}  // extern "C"

// ---- field-solve side ------------------------------------------------------

namespace {

// FP64 device buffers of the field-solve calls, grown on demand.
struct SolveBuffers {
  std::vector<double *> arr;  // named slots below
  size_t nn = 0, nc = 0, nvec = 0;
  int ns = 0;
  void *work = nullptr;
  size_t work_bytes = 0;
};
SolveBuffers g_solve;

// Slots: 0..2 B, 3 rhoc, 4 rhoh, 5..7 J hat, 8 vector, 9 image,
// 10 + 9 s .. 18 + 9 s: species s' Jx Jy Jz pxx pxy pxz pyy pyz pzz, then
// 10 + 9 ns + s: species s' rhon.
int solve_buffers(const grid *grd, int ns) {
  const size_t nn = nodes(grd), nc = (size_t)grd->nxc * grd->nyc * grd->nzc;
  const size_t nvec = 3 * (size_t)(grd->nxn - 2) * (grd->nyn - 2) * (grd->nzn - 2);
  SolveBuffers &b = g_solve;
  if (b.nn >= nn && b.nc >= nc && b.nvec >= nvec && b.ns >= ns) return SPIC_OK;
  for (double *p : b.arr)
    if (p) spic_free(p);
  b.arr.assign(10 + 10 * (size_t)ns, nullptr);
  int rc;
  for (size_t i = 0; i < b.arr.size(); ++i) {
    size_t n = nn;
    if (i == 3 || i == 4) n = nc;
    if (i == 8 || i == 9) n = nvec;
    if ((rc = spic_malloc((void **)&b.arr[i], sizeof(double) * n))) return rc;
  }
  if (b.work) spic_free(b.work);
  spic_grid g;
  std::memset(&g, 0, sizeof(g));
  g.nxc = grd->nxc; g.nyc = grd->nyc; g.nzc = grd->nzc;
  g.nxn = grd->nxn; g.nyn = grd->nyn; g.nzn = grd->nzn;
  b.work_bytes = spic_field_workspace_bytes(&g, ns);
  if ((rc = spic_malloc(&b.work, b.work_bytes))) return rc;
  b.nn = nn; b.nc = nc; b.nvec = nvec; b.ns = ns;
  return SPIC_OK;
}

template <class T>
int up64(double *dst, const T *src, size_t n) {
  std::vector<double> buf(src, src + n);
  return spic_memcpy_h2d(dst, buf.data(), sizeof(double) * n);
}

template <class T>
int down64(T *dst, const double *src, size_t n) {
  std::vector<double> buf(n);
  int rc = spic_memcpy_d2h(buf.data(), src, sizeof(double) * n);
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i) dst[i] = (T)buf[i];
  return SPIC_OK;
}

// The FP64 grid and parameters of the field-solve calls.
int solve_setup(const grid *grd, const parameters *param, spic_grid *g,
                spic_solve_params *sp) {
  GridMirror *gm;
  int rc = grid_mirror(grd, &gm);
  if (rc) return rc;
  *g = gm->g;
  g->prec = SPIC_F64;  // the operators read no node coordinates
  std::memset(sp, 0, sizeof(*sp));
  sp->dt = param->dt;
  sp->c = param->c;
  sp->th = param->th;
  sp->fourpi = param->fourpi;
  sp->PERIODICX = param->PERIODICX;
  sp->PERIODICY = param->PERIODICY;
  sp->PERIODICZ = param->PERIODICZ;
  if (param->ns < 1 || param->ns > 8) return SPIC_EINVAL;
  for (int s = 0; s < param->ns; ++s) sp->qom[s] = param->qom[s];
  return SPIC_OK;
}

int upload_b(const EMfield *field, size_t nn, spic_field *f) {
  double **a = g_solve.arr.data();
  int rc;
  if ((rc = up64(a[0], field->Bxn_flat, nn)) || (rc = up64(a[1], field->Byn_flat, nn)) ||
      (rc = up64(a[2], field->Bzn_flat, nn)))
    return rc;
  // the field-solve operators read B only; E slots point at B
  f->Ex = a[0]; f->Ey = a[1]; f->Ez = a[2];
  f->Bxn = a[0]; f->Byn = a[1]; f->Bzn = a[2];
  return SPIC_OK;
}

}  // namespace

void calculateHatDensities(struct interpDens_aux *id_aux, struct interpDensNet *idn,
                           struct interpDensSpecies *ids, struct EMfield *field,
                           struct grid *grd, struct parameters *param) {
  std::lock_guard<std::mutex> lock(g_mu);
  spic_grid g;
  spic_solve_params sp;
  int rc = solve_setup(grd, param, &g, &sp);
  const int ns = param->ns;
  if (rc == SPIC_OK) rc = solve_buffers(grd, ns);
  if (rc) {
    report("calculateHatDensities", rc);
    return;
  }
  double **a = g_solve.arr.data();
  const size_t nn = nodes(grd), nc = (size_t)grd->nxc * grd->nyc * grd->nzc;
  spic_field f;
  std::vector<spic_moments> mom(ns);
  rc = upload_b(field, nn, &f);
  if (rc == SPIC_OK) rc = up64(a[3], idn->rhoc_flat, nc);
  for (int s = 0; s < ns && rc == SPIC_OK; ++s) {
    const FPinterp *src[9] = {ids[s].Jx_flat,  ids[s].Jy_flat,  ids[s].Jz_flat,
                              ids[s].pxx_flat, ids[s].pxy_flat, ids[s].pxz_flat,
                              ids[s].pyy_flat, ids[s].pyz_flat, ids[s].pzz_flat};
    for (int m = 0; m < 9 && rc == SPIC_OK; ++m) rc = up64(a[10 + 9 * s + m], src[m], nn);
    double *const *d = &a[10 + 9 * s];
    mom[s].rhon = d[0];  // unread by the hat densities
    mom[s].Jx = d[0]; mom[s].Jy = d[1]; mom[s].Jz = d[2];
    mom[s].pxx = d[3]; mom[s].pxy = d[4]; mom[s].pxz = d[5];
    mom[s].pyy = d[6]; mom[s].pyz = d[7]; mom[s].pzz = d[8];
  }
  spic_hat hat;
  hat.rhoh = a[4]; hat.Jxh = a[5]; hat.Jyh = a[6]; hat.Jzh = a[7];
  if (rc == SPIC_OK)
    rc = spic_hat_densities(&g, &f, mom.data(), ns, a[3], &sp, &hat, g_solve.work,
                            g_solve.work_bytes, stream());
  if (rc == SPIC_OK) rc = spic_stream_synchronize(stream());
  if (rc == SPIC_OK) rc = down64(&id_aux->rhoh[0][0][0], a[4], nc);
  if (rc == SPIC_OK) rc = down64(&id_aux->Jxh[0][0][0], a[5], nn);
  if (rc == SPIC_OK) rc = down64(&id_aux->Jyh[0][0][0], a[6], nn);
  if (rc == SPIC_OK) rc = down64(&id_aux->Jzh[0][0][0], a[7], nn);
  report("calculateHatDensities", rc);
}

void MaxwellImage(FPfield *im, FPfield *vector, struct EMfield *field,
                  struct interpDensSpecies *ids, struct grid *grd,
                  struct parameters *param) {
  std::lock_guard<std::mutex> lock(g_mu);
  spic_grid g;
  spic_solve_params sp;
  int rc = solve_setup(grd, param, &g, &sp);
  const int ns = param->ns;
  if (rc == SPIC_OK) rc = solve_buffers(grd, ns);
  if (rc) {
    report("MaxwellImage", rc);
    return;
  }
  double **a = g_solve.arr.data();
  const size_t nn = nodes(grd);
  const size_t nvec = 3 * (size_t)(grd->nxn - 2) * (grd->nyn - 2) * (grd->nzn - 2);
  spic_field f;
  std::vector<spic_moments> mom(ns);
  rc = upload_b(field, nn, &f);
  if (rc == SPIC_OK) rc = up64(a[8], vector, nvec);
  for (int s = 0; s < ns && rc == SPIC_OK; ++s) {
    double *r = a[10 + 9 * ns + s];
    rc = up64(r, ids[s].rhon_flat, nn);
    mom[s].rhon = r;
    mom[s].Jx = mom[s].Jy = mom[s].Jz = r;  // unread by MaxwellImage
    mom[s].pxx = mom[s].pxy = mom[s].pxz = mom[s].pyy = mom[s].pyz = mom[s].pzz = r;
  }
  if (rc == SPIC_OK)
    rc = spic_maxwell_image(&g, &f, mom.data(), ns, &sp, a[8], a[9], g_solve.work,
                            g_solve.work_bytes, stream());
  if (rc == SPIC_OK) rc = spic_stream_synchronize(stream());
  if (rc == SPIC_OK) rc = down64(im, a[9], nvec);
  report("MaxwellImage", rc);
}
