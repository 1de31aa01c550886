// spic_engine.cpp -- the tile-resident engine of include/spic.h: particle
// storage in tile order, the sort schedule, and the one-launch particle
// phase.  Host code only; the kernels are in spic_tiled.inl and
// spic_tilesort.hip.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "spic.h"
#include "spic_internal.hpp"
#include "spic_tiled.hpp"

using spic::fail;
using spic::hip_fail;

struct spic_engine {
  spic_grid grd;                   // with engine-owned node tables
  void *xn = nullptr, *yn = nullptr, *zn = nullptr;
  int ns = 0;
  int esz = 8;                     // bytes per particle scalar
  spic::TileLayout L{};
  int sort_every = 0;
  int lds_fp64 = 1;
  int64_t phases = 0, sorts = 0;
  size_t bytes = 0;
  struct Species {
    bool loaded = false;
    int64_t n = 0;
    double qom = 0;
    int niter = 3, nsub = 1;
    void *buf[2][7] = {};          // two sets of x y z u v w q
    uint32_t *id[2] = {nullptr, nullptr};
    int cur = 0;                   // which set holds the particles
    int64_t *tile_start = nullptr; // [ntiles + 1]
  } sp[spic::kMaxSpecies];
  unsigned long long *counters = nullptr;  // see TiledArgs::counters
  double *partials = nullptr;      // see TiledArgs::partials
  size_t partials_bytes = 0;
  void *ws = nullptr;              // sort workspace
  size_t ws_bytes = 0;
  std::mutex mu;
};

namespace {

int ceil_log2(long v) {
  int b = 0;
  while ((1L << b) < v) ++b;
  return b;
}

hipError_t dmalloc(spic_engine *e, void **p, size_t bytes) {
  hipError_t err = hipMalloc(p, bytes ? bytes : 1);
  if (err == hipSuccess) e->bytes += bytes;
  return err;
}

void free_species(spic_engine *e, spic_engine::Species &s) {
  for (int b = 0; b < 2; ++b) {
    for (int a = 0; a < 7; ++a)
      if (s.buf[b][a]) (void)hipFree(s.buf[b][a]);
    if (s.id[b]) (void)hipFree(s.id[b]);
  }
  if (s.tile_start) (void)hipFree(s.tile_start);
  s = spic_engine::Species();
  (void)e;
}

template <typename T>
hipError_t sort_species(spic_engine *e, int is, hipStream_t s) {
  auto &sp = e->sp[is];
  const spic::KGrid<T> g = spic::kgrid<T>(&e->grd);
  T *src[7], *dst[7];
  for (int a = 0; a < 7; ++a) {
    src[a] = (T *)sp.buf[sp.cur][a];
    dst[a] = (T *)sp.buf[1 - sp.cur][a];
  }
  size_t need = 0;
  hipError_t err = spic::tile_sort<T>(g, e->L, (int)sp.n, src, sp.id[sp.cur], dst,
                                      sp.id[1 - sp.cur], sp.tile_start, nullptr, 0,
                                      &need, s);
  if (err != hipSuccess) return err;
  if (need > e->ws_bytes) {
    if (e->ws) {
      if ((err = hipStreamSynchronize(s)) != hipSuccess) return err;
      (void)hipFree(e->ws);
      e->bytes -= e->ws_bytes;
      e->ws = nullptr;
      e->ws_bytes = 0;
    }
    if ((err = dmalloc(e, &e->ws, need)) != hipSuccess) return err;
    e->ws_bytes = need;
  }
  err = spic::tile_sort<T>(g, e->L, (int)sp.n, src, sp.id[sp.cur], dst,
                           sp.id[1 - sp.cur], sp.tile_start, e->ws, e->ws_bytes,
                           nullptr, s);
  if (err != hipSuccess) return err;
  sp.cur = 1 - sp.cur;
  e->sorts++;
  return hipSuccess;
}

hipError_t sort_any(spic_engine *e, int is, hipStream_t s) {
  return e->grd.prec == SPIC_F64 ? sort_species<double>(e, is, s)
                                 : sort_species<float>(e, is, s);
}

template <typename T>
hipError_t launch_phase(spic_engine *e, const spic_field *field,
                        const spic_moments *mom, const spic_push *push,
                        unsigned int *status, hipStream_t s) {
  spic::TiledArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  a.g = spic::kgrid<T>(&e->grd);
  a.f = spic::kfield<T>(field);
  a.ntx = e->L.ntx; a.nty = e->L.nty; a.ntz = e->L.ntz;
  a.tx = e->L.tx; a.ty = e->L.ty; a.tz = e->L.tz;
  a.status = status;
  a.counters = e->counters;
  a.lds_fp64 = e->lds_fp64;
  a.partials = e->partials;
  int k = 0;
  for (int is = 0; is < e->ns; ++is) {
    auto &sp = e->sp[is];
    if (!sp.loaded || sp.n == 0) continue;
    auto &t = a.sp[k++];
    t.x = (T *)sp.buf[sp.cur][0]; t.y = (T *)sp.buf[sp.cur][1];
    t.z = (T *)sp.buf[sp.cur][2]; t.u = (T *)sp.buf[sp.cur][3];
    t.v = (T *)sp.buf[sp.cur][4]; t.w = (T *)sp.buf[sp.cur][5];
    t.q = (const T *)sp.buf[sp.cur][6];
    t.tile_start = sp.tile_start;
    spic_species desc;
    std::memset(&desc, 0, sizeof(desc));
    desc.qom = sp.qom;
    desc.NiterMover = sp.niter;
    desc.n_sub_cycles = sp.nsub;
    t.k = spic::kpush<T>(&desc, push);
    t.m = spic::kmom(&mom[is]);
  }
  a.ns = k;
  if (k == 0) return hipSuccess;
  const int nn = e->grd.nxn * e->grd.nyn * e->grd.nzn;
  hipError_t err = spic::strict_mode()
                       ? spic::spic_strict::launch_tiled_phase<T>(a, e->L.ntiles, s)
                       : spic::spic_fast::launch_tiled_phase<T>(a, e->L.ntiles, s);
  if (err != hipSuccess) return err;
  return spic::spic_fast::launch_tile_reduce<T>(a, nn, s);
}

}  // namespace

extern "C" {

int spic_engine_create(spic_engine **out, const spic_grid *grd, int ns,
                       const spic_engine_opts *opts) {
  if (!out) return fail(SPIC_EINVAL, "engine pointer is NULL");
  *out = nullptr;
  int rc = spic::check_grid(grd);
  if (rc) return rc;
  if (ns < 1 || ns > spic::kMaxSpecies)
    return fail(SPIC_EINVAL, "engine species count must be 1..%d", spic::kMaxSpecies);
  int tx = 4, ty = 4, tz = 4, sort_every = 0, lds_fp64 = 1;
  if (opts) {
    if (opts->tile_x || opts->tile_y || opts->tile_z) {
      tx = opts->tile_x; ty = opts->tile_y; tz = opts->tile_z;
    }
    sort_every = opts->sort_every;
    if (opts->lds_partials == 1) lds_fp64 = 0;
    else if (opts->lds_partials != 0 && opts->lds_partials != 2)
      return fail(SPIC_EINVAL, "lds_partials must be 0, 1 or 2");
  }
  const bool shape_ok = (tx == 4 && ty == 4 && tz == 4) ||
                        (tx == 8 && ty == 4 && tz == 4) ||
                        (tx == 2 && ty == 2 && tz == 2);
  if (!shape_ok)
    return fail(SPIC_EINVAL, "tile shape %dx%dx%d not built (4x4x4, 8x4x4, 2x2x2)",
                tx, ty, tz);
  spic_engine *e = new (std::nothrow) spic_engine();
  if (!e) return fail(SPIC_ENOMEM, "engine allocation");
  e->grd = *grd;
  e->ns = ns;
  e->esz = grd->prec == SPIC_F64 ? 8 : 4;
  e->sort_every = sort_every;
  e->lds_fp64 = lds_fp64;
  auto &L = e->L;
  L.ncx = grd->nxc - 2; L.ncy = grd->nyc - 2; L.ncz = grd->nzc - 2;
  L.tx = tx; L.ty = ty; L.tz = tz;
  L.ntx = (L.ncx + tx - 1) / tx;
  L.nty = (L.ncy + ty - 1) / ty;
  L.ntz = (L.ncz + tz - 1) / tz;
  L.ntiles = L.ntx * L.nty * L.ntz;
  L.cell_bits = ceil_log2((long)tx * ty * tz);
  L.tile_bits = ceil_log2(L.ntiles);
  L.rank_bits = 32 - L.cell_bits - L.tile_bits;
  if (L.rank_bits < 6) {
    delete e;
    return fail(SPIC_EUNSUPPORTED, "grid has too many tiles (%d) for 32-bit sort keys",
                L.ntiles);
  }
  // copy the node tables
  hipError_t err;
  const size_t sx = (size_t)e->esz * grd->nxn, sy = (size_t)e->esz * grd->nyn,
               sz = (size_t)e->esz * grd->nzn;
  if ((err = dmalloc(e, &e->xn, sx)) != hipSuccess ||
      (err = dmalloc(e, &e->yn, sy)) != hipSuccess ||
      (err = dmalloc(e, &e->zn, sz)) != hipSuccess ||
      (err = hipMemcpy(e->xn, grd->xn, sx, hipMemcpyDeviceToDevice)) != hipSuccess ||
      (err = hipMemcpy(e->yn, grd->yn, sy, hipMemcpyDeviceToDevice)) != hipSuccess ||
      (err = hipMemcpy(e->zn, grd->zn, sz, hipMemcpyDeviceToDevice)) != hipSuccess) {
    spic_engine_destroy(e);
    return hip_fail(err, "engine node tables");
  }
  e->grd.xn = e->xn;
  e->grd.yn = e->yn;
  e->grd.zn = e->zn;
  e->partials_bytes = sizeof(double) * 10 * (size_t)ns * L.ntiles *
                     spic::block_nodes(tx, ty, tz);
  if ((err = dmalloc(e, (void **)&e->partials, e->partials_bytes)) != hipSuccess) {
    spic_engine_destroy(e);
    return hip_fail(err, "engine tile partials");
  }
  if ((err = dmalloc(e, (void **)&e->counters, 4 * sizeof(unsigned long long))) !=
          hipSuccess ||
      (err = hipMemset(e->counters, 0, 4 * sizeof(unsigned long long))) != hipSuccess) {
    spic_engine_destroy(e);
    return hip_fail(err, "engine counters");
  }
  *out = e;
  return SPIC_OK;
}

int spic_engine_destroy(spic_engine *e) {
  if (!e) return SPIC_OK;
  (void)hipDeviceSynchronize();
  for (int is = 0; is < spic::kMaxSpecies; ++is) free_species(e, e->sp[is]);
  if (e->xn) (void)hipFree(e->xn);
  if (e->yn) (void)hipFree(e->yn);
  if (e->zn) (void)hipFree(e->zn);
  if (e->ws) (void)hipFree(e->ws);
  if (e->counters) (void)hipFree(e->counters);
  if (e->partials) (void)hipFree(e->partials);
  delete e;
  return SPIC_OK;
}

int spic_engine_load(spic_engine *e, int is, const spic_species *src,
                     void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  if (is < 0 || is >= e->ns) return fail(SPIC_EINVAL, "species %d out of range", is);
  int rc = spic::check_species(src, true, true);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  auto &sp = e->sp[is];
  const int64_t n = src->nop;
  hipError_t err;
  if (!sp.loaded || sp.n != n) {
    if ((err = hipDeviceSynchronize()) != hipSuccess) return hip_fail(err, "engine load");
    free_species(e, sp);
    for (int b = 0; b < 2; ++b) {
      for (int a = 0; a < 7; ++a)
        if ((err = dmalloc(e, &sp.buf[b][a], (size_t)e->esz * n)) != hipSuccess)
          return hip_fail(err, "engine particle storage");
      if ((err = dmalloc(e, (void **)&sp.id[b], sizeof(uint32_t) * n)) != hipSuccess)
        return hip_fail(err, "engine id storage");
    }
    if ((err = dmalloc(e, (void **)&sp.tile_start,
                       sizeof(int64_t) * (e->L.ntiles + 1))) != hipSuccess)
      return hip_fail(err, "engine tile table");
  }
  sp.n = n;
  sp.qom = src->qom;
  sp.niter = src->NiterMover;
  sp.nsub = src->n_sub_cycles;
  sp.cur = 0;
  void *srcs[7] = {src->x, src->y, src->z, src->u, src->v, src->w, src->q};
  for (int a = 0; a < 7; ++a)
    if ((err = hipMemcpyAsync(sp.buf[0][a], srcs[a], (size_t)e->esz * n,
                              hipMemcpyDeviceToDevice, s)) != hipSuccess)
      return hip_fail(err, "engine load copy");
  if ((err = spic::iota_u32(sp.id[0], (int)n, s)) != hipSuccess)
    return hip_fail(err, "engine ids");
  sp.loaded = true;
  if ((err = sort_any(e, is, s)) != hipSuccess) return hip_fail(err, "engine sort");
  if (!spic::async_mode() && (err = hipStreamSynchronize(s)) != hipSuccess)
    return hip_fail(err, "engine load");
  return SPIC_OK;
}

int spic_engine_store(spic_engine *e, int is, spic_species *dst, void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  if (is < 0 || is >= e->ns || !e->sp[is].loaded)
    return fail(SPIC_EINVAL, "species %d not loaded", is);
  int rc = spic::check_species(dst, true, false);
  if (rc) return rc;
  auto &sp = e->sp[is];
  if (dst->nop != sp.n)
    return fail(SPIC_EINVAL, "store: nop %lld != loaded %lld", (long long)dst->nop,
                (long long)sp.n);
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  void *dsts[7] = {dst->x, dst->y, dst->z, dst->u, dst->v, dst->w, dst->q};
  hipError_t err = hipSuccess;
  for (int a = 0; a < 7 && err == hipSuccess; ++a)
    err = e->esz == 8
              ? spic::scatter_by_id_f64((const double *)sp.buf[sp.cur][a],
                                        sp.id[sp.cur], (double *)dsts[a], (int)sp.n, s)
              : spic::scatter_by_id_f32((const float *)sp.buf[sp.cur][a],
                                        sp.id[sp.cur], (float *)dsts[a], (int)sp.n, s);
  if (err == hipSuccess && !spic::async_mode()) err = hipStreamSynchronize(s);
  return err == hipSuccess ? SPIC_OK : hip_fail(err, "engine store");
}

int spic_engine_sort(spic_engine *e, void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  for (int is = 0; is < e->ns; ++is) {
    if (!e->sp[is].loaded) continue;
    hipError_t err = sort_any(e, is, s);
    if (err != hipSuccess) return hip_fail(err, "engine sort");
  }
  if (!spic::async_mode()) {
    hipError_t err = hipStreamSynchronize(s);
    if (err != hipSuccess) return hip_fail(err, "engine sort");
  }
  return SPIC_OK;
}

int spic_engine_phase(spic_engine *e, const spic_field *field,
                      const spic_moments *mom, const spic_push *push,
                      void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  int rc;
  if ((rc = spic::check_field(field)) || (rc = spic::check_push(push))) return rc;
  if (!mom) return fail(SPIC_EINVAL, "moments are NULL");
  for (int is = 0; is < e->ns; ++is) {
    if (!e->sp[is].loaded) continue;
    if ((rc = spic::check_moments(&mom[is]))) return rc;
  }
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  hipError_t err;
  if (e->sort_every > 0 && e->phases > 0 && e->phases % e->sort_every == 0) {
    for (int is = 0; is < e->ns; ++is) {
      if (!e->sp[is].loaded) continue;
      if ((err = sort_any(e, is, s)) != hipSuccess) return hip_fail(err, "engine sort");
    }
  }
  unsigned int *st;
  if ((rc = spic::begin_call(s, &st))) return rc;
  err = e->grd.prec == SPIC_F64 ? launch_phase<double>(e, field, mom, push, st, s)
                                : launch_phase<float>(e, field, mom, push, st, s);
  e->phases++;
  return spic::end_call(s, st, err);
}

int spic_engine_get_stats(const spic_engine *e, spic_engine_stats *st) {
  if (!e || !st) return fail(SPIC_EINVAL, "NULL argument");
  st->phases = e->phases;
  st->sorts = e->sorts;
  st->ntiles = e->L.ntiles;
  st->tile_x = e->L.tx;
  st->tile_y = e->L.ty;
  st->tile_z = e->L.tz;
  st->device_bytes = e->bytes;
  unsigned long long c[4] = {0, 0, 0, 0};
  hipError_t err = hipMemcpy(c, e->counters, sizeof(c), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return hip_fail(err, "engine counters");
  st->global_gathers = (int64_t)c[0];
  st->global_deposits = (int64_t)c[1];
  st->lds_gathers = (int64_t)c[2];
  st->lds_deposits = (int64_t)c[3];
  return SPIC_OK;
}

}  // extern "C"
