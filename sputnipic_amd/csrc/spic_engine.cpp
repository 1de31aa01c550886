// spic_engine.cpp -- the tile-resident engine of include/spic.h: binned
// particle storage, the sort schedule, and the one-launch particle phase.
// Host code only; the kernels are in spic_binned.inl and spic_binsort.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "spic.h"
#include "spic_binned.hpp"
#include "spic_internal.hpp"

using spic::fail;
using spic::hip_fail;

struct spic_engine {
  spic_grid grd;                   // with engine-owned node tables
  void *xn = nullptr, *yn = nullptr, *zn = nullptr;
  int ns = 0;
  int esz = 8;                     // bytes per particle scalar
  spic::BinLayout L{};
  int slack_pct = 12;
  int sort_every = 0;
  int sort_mode = SPIC_SORT_TILE;  // re-sorts: tile_sort or bin_sort
  int lds_fp64 = 1;
  int set_moments = 0;             // phases write the moments (else add)
  int keep_ids = 1;                // a load index per particle (store in load order)
  int ovf_pct = 100;               // overflow slots per buffer, percent of n
  int ppt = 1;                     // particles per thread and batch of the phase
  int64_t phases = 0, sorts = 0;
  uint64_t gen = 0;                // bumped by every load (storage may move)
  size_t bytes = 0;
  struct Species {
    bool loaded = false;
    int64_t n = 0;
    double qom = 0;
    int niter = 3, nsub = 1;
    int64_t slack = 0;             // free slots per segment after a sort
    int64_t cap = 0;               // slots per buffer
    int64_t alloc_cap = 0;         // slots the buffers were allocated for
    unsigned int ovf_cap = 0;      // overflow slots (n by default: nothing is lost)
    void *block[2] = {};           // two buffers of x y z u v w q, one block each
    void *buf[2][7] = {};          // the components: block[b] + c * cap
    uint32_t *id[2] = {nullptr, nullptr};
    unsigned int *cnt[2] = {nullptr, nullptr};  // [3 * (ntiles + 1)] per buffer
    int64_t *seg[2] = {nullptr, nullptr};       // [ntiles + 1] per layout
    int cur = 0;                   // buffer holding the particles
    int segi = 0;                  // layout in use
  } sp[spic::kMaxSpecies];
  unsigned long long *counters = nullptr;  // see BinnedArgs::counters
  void *partials = nullptr;        // see BinnedArgs::partials
  size_t partials_bytes = 0;
  unsigned char *pflag = nullptr;  // see BinnedArgs::pflag
  void *ws = nullptr;              // sort workspace
  size_t ws_bytes = 0;
  std::mutex mu;
};

namespace {

constexpr int kOverflowBlocks = 256;

int ceil_log2(long v) {
  int b = 0;
  while ((1L << b) < v) ++b;
  return b;
}

// spic_test_fail_engine_alloc: allocations left before the one that fails
// (negative: none fails)
std::atomic<int> g_fail_after{-1};

hipError_t dmalloc(spic_engine *e, void **p, size_t bytes) {
  *p = nullptr;
  if (g_fail_after.load() >= 0 && g_fail_after.fetch_sub(1) == 0) {
    g_fail_after.store(-1);
    return hipErrorOutOfMemory;
  }
  hipError_t err = hipMalloc(p, bytes ? bytes : 1);
  if (err == hipSuccess) e->bytes += bytes;
  return err;
}

void free_species(spic_engine *e, spic_engine::Species &s) {
  const size_t nt1 = (size_t)e->L.ntiles + 1;
  for (int b = 0; b < 2; ++b) {
    if (s.block[b]) {
      (void)hipFree(s.block[b]);
      e->bytes -= (size_t)e->esz * 7 * s.alloc_cap;
    }
    if (s.id[b]) {
      (void)hipFree(s.id[b]);
      e->bytes -= sizeof(uint32_t) * s.alloc_cap;
    }
    if (s.cnt[b]) {
      (void)hipFree(s.cnt[b]);
      e->bytes -= sizeof(unsigned int) * 3 * nt1;
    }
    if (s.seg[b]) {
      (void)hipFree(s.seg[b]);
      e->bytes -= sizeof(int64_t) * nt1;
    }
  }
  s = spic_engine::Species();
}

hipError_t ensure_workspace(spic_engine *e, size_t need, hipStream_t s) {
  if (need <= e->ws_bytes) return hipSuccess;
  hipError_t err;
  if (e->ws) {
    if ((err = hipStreamSynchronize(s)) != hipSuccess) return err;
    (void)hipFree(e->ws);
    e->bytes -= e->ws_bytes;
    e->ws = nullptr;
    e->ws_bytes = 0;
  }
  if ((err = dmalloc(e, &e->ws, need)) != hipSuccess) return err;
  e->ws_bytes = need;
  return hipSuccess;
}

// Sort species `is` into the other buffer and the other layout.  With
// `src` given, the particles come densely from the caller's arrays instead.
// A re-sort in SPIC_SORT_TILE mode is a tile sort.
// The tile-sort job of species `is` from buffer `cur` in layout `segi` into
// the other ones; count / own / fill at ws.
template <typename T>
spic::TileSortJob<T> tile_sort_job(const spic_engine *e, int is, int cur, int segi,
                                   unsigned int *ws) {
  const auto &sp = e->sp[is];
  spic::TileSortJob<T> J;
  std::memset(&J, 0, sizeof(J));
  J.g = spic::kgrid<T>(&e->grd);
  J.L = e->L;
  J.L.slack = sp.slack;
  const int from = cur, to = 1 - cur;
  J.src = (const T *)sp.block[from];
  J.dst = (T *)sp.block[to];
  J.stride = sp.cap;
  J.src_id = sp.id[from];
  J.dst_id = sp.id[to];
  J.src_seg = sp.seg[segi];
  J.dst_seg = sp.seg[1 - segi];
  J.src_cnt = sp.cnt[from];
  J.dst_cnt = sp.cnt[to];
  J.ovf_cap = sp.ovf_cap;
  const size_t nt1 = (size_t)e->L.ntiles + 1;
  J.count = ws;
  J.own = ws + nt1;
  J.fill = ws + 2 * nt1;
  return J;
}

template <typename T>
hipError_t tile_sort_species(spic_engine *e, int is, hipStream_t s) {
  auto &sp = e->sp[is];
  hipError_t err =
      ensure_workspace(e, 3 * ((size_t)e->L.ntiles + 1) * sizeof(unsigned int), s);
  if (err != hipSuccess) return err;
  err = spic::tile_sort<T>(tile_sort_job<T>(e, is, sp.cur, sp.segi, (unsigned int *)e->ws), s);
  if (err != hipSuccess) return err;
  sp.cur = 1 - sp.cur;
  sp.segi = 1 - sp.segi;
  e->sorts++;
  return hipSuccess;
}

// `src`: the caller's arrays (src_f64: FP64 ones) for a load, else null.
template <typename T>
hipError_t sort_species(spic_engine *e, int is, void *const *src, int src_f64,
                        hipStream_t s) {
  auto &sp = e->sp[is];
  if (!src && e->sort_mode == SPIC_SORT_TILE) return tile_sort_species<T>(e, is, s);
  const spic::KGrid<T> g = spic::kgrid<T>(&e->grd);
  spic::BinLayout L = e->L;
  L.slack = sp.slack;
  const int from = sp.cur, to = src ? 0 : 1 - sp.cur;
  const int lay_from = sp.segi, lay_to = src ? 0 : 1 - sp.segi;
  const void *in[7];
  T *out[7];
  for (int a = 0; a < 7; ++a) {
    in[a] = src ? src[a] : sp.buf[from][a];
    out[a] = (T *)sp.buf[to][a];
  }
  const int in_f64 = src ? src_f64 : (sizeof(T) == 8 ? 1 : 0);
  const uint32_t *in_id = src ? nullptr : sp.id[from];
  const int64_t *in_seg = src ? nullptr : sp.seg[lay_from];
  const unsigned int *in_cnt = src ? nullptr : sp.cnt[from];
  size_t need = 0;
  hipError_t err = spic::bin_sort<T>(g, L, sp.n, in, in_f64, in_id, in_seg, in_cnt,
                                     sp.ovf_cap, out, sp.id[to], sp.seg[lay_to], sp.cnt[to],
                                     nullptr, 0, &need, s);
  if (err != hipSuccess) return err;
  if ((err = ensure_workspace(e, need, s)) != hipSuccess) return err;
  err = spic::bin_sort<T>(g, L, sp.n, in, in_f64, in_id, in_seg, in_cnt, sp.ovf_cap, out,
                          sp.id[to], sp.seg[lay_to], sp.cnt[to], e->ws, e->ws_bytes,
                          nullptr, s);
  if (err != hipSuccess) return err;
  sp.cur = to;
  sp.segi = lay_to;
  e->sorts++;
  return hipSuccess;
}

hipError_t sort_any(spic_engine *e, int is, void *const *src, hipStream_t s,
                    int src_f64 = 0) {
  return e->grd.prec == SPIC_F64 ? sort_species<double>(e, is, src, 1, s)
                                 : sort_species<float>(e, is, src, src_f64, s);
}

// The phase kernel's argument block for the loaded species of `e`, read
// from buffer `cur` in layout `segi` (the species' own when negative).
// idx[k] is the engine species of argument species k; returns the count.
template <typename T>
int fill_args(const spic_engine *e, const spic_field *field, const spic_moments *mom,
              const spic_push *push, unsigned int *status, int cur, int segi,
              spic::BinnedArgs<T> &a, int *idx) {
  std::memset(&a, 0, sizeof(a));
  a.g = spic::kgrid<T>(&e->grd);
  a.f = spic::kfield<T>(field);
  a.ncx = e->L.ncx; a.ncy = e->L.ncy; a.ncz = e->L.ncz;
  a.ntx = e->L.ntx; a.nty = e->L.nty; a.ntz = e->L.ntz;
  a.ntiles = e->L.ntiles;
  a.tx = e->L.tx; a.ty = e->L.ty; a.tz = e->L.tz;
  a.lds_fp64 = e->lds_fp64;
  a.ppt = e->ppt;
  a.fixed = e->L.fixed;
  a.fx = e->L.fx;
  a.n_ovf_blocks = kOverflowBlocks;
  a.status = status;
  a.counters = e->counters;
  a.partials = e->partials;
  a.pflag = e->pflag;
  int k = 0;
  for (int is = 0; is < e->ns; ++is) {
    const auto &sp = e->sp[is];
    if (!sp.loaded || sp.n == 0) continue;
    if (idx) idx[k] = is;
    auto &t = a.sp[k];
    const int in = cur < 0 ? sp.cur : cur, out = 1 - in;
    t.in = (const T *)sp.block[in];
    t.out = (T *)sp.block[out];
    t.stride = sp.cap;
    t.id_in = sp.id[in];
    t.id_out = sp.id[out];
    t.seg = sp.seg[segi < 0 ? sp.segi : segi];
    t.cnt_in = sp.cnt[in];
    t.cnt_out = sp.cnt[out];
    t.ovf_cap = sp.ovf_cap;
    spic_species desc;
    std::memset(&desc, 0, sizeof(desc));
    desc.qom = sp.qom;
    desc.NiterMover = sp.niter;
    desc.n_sub_cycles = sp.nsub;
    t.k = spic::kpush<T>(&desc, push);
    t.m = spic::kmom(&mom[is]);
    ++k;
  }
  a.ns = k;
  return k;
}

// One phase of every loaded species; with push_off, the deposit alone
// (the particles are copied to the other buffer unchanged).
template <typename T>
hipError_t launch_phase(spic_engine *e, const spic_field *field,
                        const spic_moments *mom, const spic_push *push,
                        unsigned int *status, hipStream_t s, int push_off = 0) {
  spic::BinnedArgs<T> a;
  int idx[spic::kMaxSpecies];
  const int k = fill_args<T>(e, field, mom, push, status, -1, -1, a, idx);
  spic::BinnedReduceArgs r;
  std::memset(&r, 0, sizeof(r));
  spic::CountArrays zc;
  std::memset(&zc, 0, sizeof(zc));
  for (int j = 0; j < k; ++j) {
    r.mom[j] = a.sp[j].m;
    zc.p[j] = a.sp[j].cnt_out;
    zc.mom[j] = a.sp[j].m;
  }
  r.ns = zc.ns = k;
  // writing the moments: zero them in the same launch; the deposits (LDS
  // partials through the reduce, global atomics directly) then add.  A
  // loaded species without particles deposits nothing, so its moments are
  // zeroed here too (its counters, which no kernel reads, along with them).
  if (e->set_moments) {
    for (int is = 0; is < e->ns; ++is) {
      const auto &sp = e->sp[is];
      if (!sp.loaded || sp.n != 0) continue;
      zc.p[zc.ns] = sp.cnt[1 - sp.cur];
      zc.mom[zc.ns] = spic::kmom(&mom[is]);
      ++zc.ns;
    }
  }
  if (zc.ns == 0) return hipSuccess;
  a.push_off = push_off;
  zc.n = 3 * (e->L.ntiles + 1);
  zc.nn = e->set_moments ? (int64_t)e->grd.nxn * e->grd.nyn * e->grd.nzn : 0;
  hipError_t err = spic::zero_counts(zc, s);
  if (err != hipSuccess || k == 0) return err;
  err = spic::strict_mode() ? spic::spic_strict::launch_binned_phase<T>(a, s)
                            : spic::spic_fast::launch_binned_phase<T>(a, s);
  if (err != hipSuccess) return err;
  r.nyn = e->grd.nyn; r.nzn = e->grd.nzn;
  r.ntx = e->L.ntx; r.nty = e->L.nty; r.ntz = e->L.ntz;
  r.ntiles = e->L.ntiles;
  r.partials = e->partials;
  r.partials_fp32 = e->esz == 4 ? 1 : 0;
  r.pflag = e->pflag;
  const int nn = e->grd.nxn * e->grd.nyn * e->grd.nzn;
  err = spic::launch_binned_reduce(r, e->L.tx, e->L.ty, e->L.tz, nn, s);
  if (err != hipSuccess) return err;
  for (int j = 0; j < k; ++j) e->sp[idx[j]].cur ^= 1;
  return hipSuccess;
}

// Dense store positions of species `is`'s ranges when the engine keeps no
// ids (*bases stays null when it does): range counts and a scan into the
// sort workspace.
hipError_t store_bases(spic_engine *e, int is, const uint32_t **bases, hipStream_t s) {
  *bases = nullptr;
  const auto &sp = e->sp[is];
  if (sp.id[sp.cur]) return hipSuccess;
  size_t need = 0;
  hipError_t err = spic::range_bases(sp.seg[sp.segi], sp.cnt[sp.cur], sp.ovf_cap, e->L.ntiles,
                                     nullptr, 0, &need, bases, s);
  if (err != hipSuccess) return err;
  if ((err = ensure_workspace(e, need, s)) != hipSuccess) return err;
  return spic::range_bases(sp.seg[sp.segi], sp.cnt[sp.cur], sp.ovf_cap, e->L.ntiles, e->ws,
                           e->ws_bytes, nullptr, bases, s);
}

// A species loaded with n_sub_cycles != 1 runs only while a sub-cycling
// mode is chosen (the mode may have changed since the load).
int check_subcycles(const spic_engine *e) {
  if (spic::subcycle_mode() != SPIC_SUBCYCLE_REFUSE) return SPIC_OK;
  for (int is = 0; is < e->ns; ++is)
    if (e->sp[is].loaded && e->sp[is].nsub != 1)
      return fail(SPIC_EUNSUPPORTED,
                  "species %d has n_sub_cycles = %d and sub-cycling is refused "
                  "(spic_set_subcycling)", is, e->sp[is].nsub);
  return SPIC_OK;
}

// Strict arithmetic reproduces the reference's rounding of lengths, which
// fixed-point storage does not hold.
int check_positions(const spic_engine *e) {
  if (e->L.fixed && spic::strict_mode())
    return fail(SPIC_EINVAL,
                "the engine stores fixed-point positions (an FP32 engine created in "
                "production arithmetic); strict arithmetic needs one created with "
                "positions = SPIC_POS_FLOAT");
  return SPIC_OK;
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
  int tx = 4, ty = 4, tz = 4, sort_every = 0, lds_fp64 = 1, slack_pct = 12, set_mom = 0,
      sort_mode = SPIC_SORT_TILE, positions = 0, keep_ids = 1, ovf_pct = 100, ppt = 1;
  if (opts) {
    if (opts->tile_x || opts->tile_y || opts->tile_z) {
      tx = opts->tile_x; ty = opts->tile_y; tz = opts->tile_z;
    }
    sort_every = opts->sort_every;
    if (opts->lds_partials == 1) lds_fp64 = 0;
    else if (opts->lds_partials != 0 && opts->lds_partials != 2)
      return fail(SPIC_EINVAL, "lds_partials must be 0, 1 or 2");
    if (opts->slack_pct < 0 || opts->slack_pct > 1000)
      return fail(SPIC_EINVAL, "slack_pct must be 0..1000");
    if (opts->slack_pct) slack_pct = opts->slack_pct;
    if (opts->set_moments != 0 && opts->set_moments != 1)
      return fail(SPIC_EINVAL, "set_moments must be 0 or 1");
    set_mom = opts->set_moments;
    if (opts->sort_mode != SPIC_SORT_TILE && opts->sort_mode != SPIC_SORT_RADIX)
      return fail(SPIC_EINVAL, "sort_mode must be SPIC_SORT_TILE or SPIC_SORT_RADIX");
    sort_mode = opts->sort_mode;
    if (opts->positions != 0 && opts->positions != SPIC_POS_FLOAT &&
        opts->positions != SPIC_POS_FIXED)
      return fail(SPIC_EINVAL, "positions must be 0, SPIC_POS_FLOAT or SPIC_POS_FIXED");
    positions = opts->positions;
    if (opts->order != 0 && opts->order != SPIC_ORDER_LOAD && opts->order != SPIC_ORDER_ENGINE)
      return fail(SPIC_EINVAL, "order must be 0, SPIC_ORDER_LOAD or SPIC_ORDER_ENGINE");
    keep_ids = opts->order == SPIC_ORDER_ENGINE ? 0 : 1;
    if (opts->overflow_pct < 0 || opts->overflow_pct > 100)
      return fail(SPIC_EINVAL, "overflow_pct must be 0..100");
    if (opts->overflow_pct) ovf_pct = opts->overflow_pct;
    if (opts->particles_per_thread < 0 || opts->particles_per_thread > 2)
      return fail(SPIC_EINVAL, "particles_per_thread must be 0, 1 or 2");
    if (opts->particles_per_thread) ppt = opts->particles_per_thread;
  }
  if (positions == 0)
    positions = (grd->prec == SPIC_F32 && !spic::strict_mode()) ? SPIC_POS_FIXED
                                                                 : SPIC_POS_FLOAT;
  if (positions == SPIC_POS_FIXED && grd->prec != SPIC_F32)
    return fail(SPIC_EUNSUPPORTED,
                "fixed-point positions are for FP32 engines (an FP64 length is finer)");
  if (ppt == 2 && positions != SPIC_POS_FIXED)
    return fail(SPIC_EUNSUPPORTED,
                "two particles per thread are built for FP32 fixed-point engines only");
  // the phase kernel packs a deposit cell into 12 + 10 + 10 bits
  if (grd->nxn > 4096 || grd->nyn > 1024 || grd->nzn > 1024)
    return fail(SPIC_EUNSUPPORTED,
                "engine grids hold at most 4096 x 1024 x 1024 nodes (got %d x %d x %d)",
                grd->nxn, grd->nyn, grd->nzn);
  const bool shape_ok = (tx == 4 && ty == 4 && tz == 4) || (tx == 8 && ty == 4 && tz == 4);
  if (!shape_ok)
    return fail(SPIC_EINVAL, "tile shape %dx%dx%d not built (4x4x4, 8x4x4)", tx, ty, tz);
  spic_engine *e = new (std::nothrow) spic_engine();
  if (!e) return fail(SPIC_ENOMEM, "engine allocation");
  e->grd = *grd;
  e->ns = ns;
  e->esz = grd->prec == SPIC_F64 ? 8 : 4;
  e->sort_every = sort_every;
  e->sort_mode = sort_mode;
  e->lds_fp64 = lds_fp64;
  e->slack_pct = slack_pct;
  e->set_moments = set_mom;
  e->keep_ids = keep_ids;
  e->ovf_pct = ovf_pct;
  e->ppt = ppt;
  auto &L = e->L;
  L.ncx = grd->nxc - 2; L.ncy = grd->nyc - 2; L.ncz = grd->nzc - 2;
  L.tx = tx; L.ty = ty; L.tz = tz;
  L.ntx = (L.ncx + tx - 1) / tx;
  L.nty = (L.ncy + ty - 1) / ty;
  L.ntz = (L.ncz + tz - 1) / tz;
  L.ntiles = L.ntx * L.nty * L.ntz;
  L.tile_bits = ceil_log2(L.ntiles);
  L.cell_bits = ceil_log2((long)tx * ty * tz);
  L.rank_bits = 32 - L.tile_bits - L.cell_bits;
  if (L.rank_bits < 4) {
    delete e;
    return fail(SPIC_EUNSUPPORTED, "grid has too many tiles (%d) for 32-bit sort keys",
                L.ntiles);
  }
  if (L.rank_bits > 16) L.rank_bits = 16;
  L.fixed = positions == SPIC_POS_FIXED ? 1 : 0;
  L.fx = spic::fixed_params(L.ncx, L.ncy, L.ncz, grd->dx, grd->dy, grd->dz, grd->xStart,
                            grd->yStart, grd->zStart);
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
  e->partials_bytes = (size_t)e->esz * 10 * (size_t)ns * L.ntiles *
                     spic::block_nodes(tx, ty, tz);
  if ((err = dmalloc(e, (void **)&e->partials, e->partials_bytes)) != hipSuccess ||
      (err = dmalloc(e, (void **)&e->pflag, (size_t)ns * L.ntiles)) != hipSuccess) {
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
  if (e->pflag) (void)hipFree(e->pflag);
  delete e;
  return SPIC_OK;
}

int spic_engine_load_as(spic_engine *e, int is, const spic_species *src, int prec,
                        void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  if (prec != SPIC_F64 && prec != SPIC_F32)
    return fail(SPIC_EINVAL, "precision must be SPIC_F64 or SPIC_F32");
  if (e->grd.prec == SPIC_F64 && prec != SPIC_F64)
    return fail(SPIC_EUNSUPPORTED, "an FP64 engine loads FP64 arrays only");
  if (is < 0 || is >= e->ns) return fail(SPIC_EINVAL, "species %d out of range", is);
  int rc = spic::check_species(src, true, true);
  if (rc) return rc;
  if (src->nop > 0x7FFFFFFFll)
    return fail(SPIC_EUNSUPPORTED, "engine species hold at most 2^31-1 particles");
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  auto &sp = e->sp[is];
  const int64_t n = src->nop;
  const int64_t nt = e->L.ntiles;
  int64_t slack = (n * e->slack_pct + 100 * nt - 1) / (100 * nt);
  if (slack < 16) slack = 16;
  // segments (n + nt * slack) then the overflow region: n slots by default
  // (no particle can be lost), else overflow_pct percent of n (at least
  // 1024 slots), and a phase whose particles do not fit reports SPIC_ENOMEM
  int64_t ovf = n;
  if (e->ovf_pct < 100) {
    ovf = (n * e->ovf_pct + 99) / 100;
    if (ovf < 1024) ovf = 1024;
    if (ovf > n) ovf = n;
  }
  if (ovf < 1) ovf = 1;
  const int64_t cap = n + nt * slack + ovf;
  // the phase kernels address a slot by its 32-bit byte offset within a
  // component array
  if ((uint64_t)cap * (uint64_t)e->esz > 0xffffffffull)
    return fail(SPIC_EUNSUPPORTED,
                "species of %lld particles: its %lld slots per component exceed 4 GiB "
                "(at most about %lld particles per species at this precision)",
                (long long)n, (long long)cap, (long long)(0xffffffffull / e->esz / 2.2));
  hipError_t err;
  // storage is kept while it holds `cap` slots (a smaller species loaded
  // after a larger one reuses it with the smaller component stride)
  if (!sp.block[0] || cap > sp.alloc_cap) {
    if ((err = hipDeviceSynchronize()) != hipSuccess) return hip_fail(err, "engine load");
    free_species(e, sp);
    sp.alloc_cap = cap;
    // a failed allocation frees every buffer this load allocated (free_species
    // releases the non-null ones), so the species is left unallocated and the
    // next load allocates all of them again instead of reusing a partial set
    const char *what = nullptr;
    for (int b = 0; b < 2 && !what; ++b) {
      if ((err = dmalloc(e, &sp.block[b], (size_t)e->esz * 7 * cap)) != hipSuccess)
        what = "engine particle storage";
      else if (e->keep_ids &&
               (err = dmalloc(e, (void **)&sp.id[b], sizeof(uint32_t) * cap)) != hipSuccess)
        what = "engine id storage";
      else if ((err = dmalloc(e, (void **)&sp.cnt[b],
                              sizeof(unsigned int) * 3 * (nt + 1))) != hipSuccess ||
               (err = dmalloc(e, (void **)&sp.seg[b], sizeof(int64_t) * (nt + 1))) !=
                   hipSuccess)
        what = "engine tile tables";
    }
    if (what) {
      free_species(e, sp);
      return hip_fail(err, what);
    }
  }
  for (int b = 0; b < 2; ++b)
    for (int a = 0; a < 7; ++a)
      sp.buf[b][a] = (char *)sp.block[b] + (size_t)e->esz * a * cap;
  sp.n = n;
  sp.cap = cap;
  sp.slack = slack;
  sp.ovf_cap = (unsigned int)ovf;
  sp.qom = src->qom;
  sp.niter = src->NiterMover;
  sp.nsub = src->n_sub_cycles;
  sp.cur = 0;
  sp.segi = 0;
  sp.loaded = true;
  e->gen++;
  void *srcs[7] = {src->x, src->y, src->z, src->u, src->v, src->w, src->q};
  if ((err = sort_any(e, is, srcs, s, prec == SPIC_F64 ? 1 : 0)) != hipSuccess) {
    sp.loaded = false;
    return hip_fail(err, "engine sort");
  }
  if (!spic::async_mode() && (err = hipStreamSynchronize(s)) != hipSuccess)
    return hip_fail(err, "engine load");
  return SPIC_OK;
}

int spic_engine_load(spic_engine *e, int is, const spic_species *src, void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  return spic_engine_load_as(e, is, src, e->grd.prec, stream);
}

int spic_engine_store_as(spic_engine *e, int is, spic_species *dst, int prec,
                         void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  if (prec != SPIC_F64 && prec != SPIC_F32)
    return fail(SPIC_EINVAL, "precision must be SPIC_F64 or SPIC_F32");
  if (e->grd.prec == SPIC_F64 && prec != SPIC_F64)
    return fail(SPIC_EUNSUPPORTED, "an FP64 engine stores FP64 arrays only");
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
  const int f64 = prec == SPIC_F64 ? 1 : 0;
  const uint32_t *bases = nullptr;
  hipError_t err = store_bases(e, is, &bases, s);
  if (err == hipSuccess)
    err = e->esz == 8
              ? spic::bin_store<double>(e->L, (double *const *)sp.buf[sp.cur], sp.id[sp.cur],
                                        bases, sp.seg[sp.segi], sp.cnt[sp.cur], sp.ovf_cap,
                                        dsts, f64, s)
              : spic::bin_store<float>(e->L, (float *const *)sp.buf[sp.cur], sp.id[sp.cur],
                                       bases, sp.seg[sp.segi], sp.cnt[sp.cur], sp.ovf_cap,
                                       dsts, f64, s);
  if (err == hipSuccess && !spic::async_mode()) err = hipStreamSynchronize(s);
  return err == hipSuccess ? SPIC_OK : hip_fail(err, "engine store");
}

int spic_engine_store(spic_engine *e, int is, spic_species *dst, void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  return spic_engine_store_as(e, is, dst, e->grd.prec, stream);
}

int spic_engine_snapshot_be32(spic_engine *e, int is, void *const dst[6], void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  if (is < 0 || is >= e->ns || !e->sp[is].loaded)
    return fail(SPIC_EINVAL, "species %d not loaded", is);
  if (!dst) return fail(SPIC_EINVAL, "snapshot arrays are NULL");
  for (int a = 0; a < 6; ++a)
    if (!dst[a]) return fail(SPIC_EINVAL, "snapshot array %d is NULL", a);
  std::lock_guard<std::mutex> lock(e->mu);
  auto &sp = e->sp[is];
  hipStream_t s = (hipStream_t)stream;
  uint32_t *const *d = (uint32_t *const *)dst;
  const uint32_t *bases = nullptr;
  hipError_t err = store_bases(e, is, &bases, s);
  if (err == hipSuccess)
    err = e->esz == 8
              ? spic::bin_store_be32<double>(e->L, (double *const *)sp.buf[sp.cur],
                                             sp.id[sp.cur], bases, sp.seg[sp.segi],
                                             sp.cnt[sp.cur], sp.ovf_cap, d, s)
              : spic::bin_store_be32<float>(e->L, (float *const *)sp.buf[sp.cur],
                                            sp.id[sp.cur], bases, sp.seg[sp.segi],
                                            sp.cnt[sp.cur], sp.ovf_cap, d, s);
  if (err == hipSuccess && !spic::async_mode()) err = hipStreamSynchronize(s);
  return err == hipSuccess ? SPIC_OK : hip_fail(err, "engine snapshot");
}

int spic_engine_sort(spic_engine *e, void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  for (int is = 0; is < e->ns; ++is) {
    if (!e->sp[is].loaded) continue;
    hipError_t err = sort_any(e, is, nullptr, s);
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
  if ((rc = check_positions(e)) || (rc = check_subcycles(e))) return rc;
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  hipError_t err;
  if (e->sort_every > 0 && e->phases > 0 && e->phases % e->sort_every == 0) {
    for (int is = 0; is < e->ns; ++is) {
      if (!e->sp[is].loaded) continue;
      if ((err = sort_any(e, is, nullptr, s)) != hipSuccess)
        return hip_fail(err, "engine sort");
    }
  }
  unsigned int *st;
  if ((rc = spic::begin_call(s, &st))) return rc;
  err = e->grd.prec == SPIC_F64 ? launch_phase<double>(e, field, mom, push, st, s)
                                : launch_phase<float>(e, field, mom, push, st, s);
  e->phases++;
  return spic::end_call(s, st, err);
}

int spic_engine_deposit(spic_engine *e, const spic_moments *mom, void *stream) {
  if (!e) return fail(SPIC_EINVAL, "engine is NULL");
  if (!mom) return fail(SPIC_EINVAL, "moments are NULL");
  int rc;
  for (int is = 0; is < e->ns; ++is) {
    if (!e->sp[is].loaded) continue;
    if ((rc = spic::check_moments(&mom[is]))) return rc;
  }
  if ((rc = check_positions(e))) return rc;
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t s = (hipStream_t)stream;
  // neither is read: the kernels skip the push and the field stage
  spic_field f0;
  std::memset(&f0, 0, sizeof(f0));
  spic_push p0;
  std::memset(&p0, 0, sizeof(p0));
  p0.c = 1.0;
  unsigned int *st;
  if ((rc = spic::begin_call(s, &st))) return rc;
  hipError_t err = e->grd.prec == SPIC_F64
                       ? launch_phase<double>(e, &f0, mom, &p0, st, s, 1)
                       : launch_phase<float>(e, &f0, mom, &p0, st, s, 1);
  return spic::end_call(s, st, err);
}

int spic_test_fail_engine_alloc(int after) {
  g_fail_after.store(after < 0 ? -1 : after);
  return SPIC_OK;
}

int spic_engine_get_stats(const spic_engine *e, spic_engine_stats *st) {
  if (!e || !st) return fail(SPIC_EINVAL, "NULL argument");
  std::memset(st, 0, sizeof(*st));
  st->phases = e->phases;
  st->sorts = e->sorts;
  st->ntiles = e->L.ntiles;
  st->tile_x = e->L.tx;
  st->tile_y = e->L.ty;
  st->tile_z = e->L.tz;
  st->device_bytes = e->bytes;
  st->positions = e->L.fixed ? SPIC_POS_FIXED : SPIC_POS_FLOAT;
  st->workspace_bytes = e->ws_bytes;
  unsigned long long c[4] = {0, 0, 0, 0};
  hipError_t err = hipMemcpy(c, e->counters, sizeof(c), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return hip_fail(err, "engine counters");
  st->global_gathers = (int64_t)c[0];
  st->global_deposits = (int64_t)c[1];
  st->lds_gathers = (int64_t)c[2];
  st->lds_deposits = (int64_t)c[3];
  for (int is = 0; is < e->ns; ++is) {
    const auto &sp = e->sp[is];
    if (!sp.loaded) continue;
    unsigned int ovf = 0;
    err = hipMemcpy(&ovf, sp.cnt[sp.cur] + e->L.ntiles, sizeof(ovf),
                    hipMemcpyDeviceToHost);
    if (err != hipSuccess) return hip_fail(err, "engine overflow count");
    st->overflow += ovf < sp.ovf_cap ? ovf : sp.ovf_cap;
  }
  return SPIC_OK;
}

}  // extern "C"

// ---- batches --------------------------------------------------------------

struct spic_batch {
  int n = 0;
  int prec = SPIC_F64;
  int sort_every = 0;
  int flags = 0;
  int64_t phases = 0;
  spic_engine *e[spic::kMaxInstances] = {};
  spic_field field[spic::kMaxInstances];
  spic_moments mom[spic::kMaxInstances][spic::kMaxSpecies];
  spic_push push[spic::kMaxInstances];
  // the argument tables were built for these engine generations and this
  // status buffer
  uint64_t gen[spic::kMaxInstances] = {};
  unsigned int *status = nullptr;
  // device: four tables of n argument blocks, one per (layout, buffer)
  // state: table (segi * 2 + cur)
  void *tab = nullptr;
  size_t arg_bytes = 0;
  spic::BatchMap nodes{}, counts{};
  int max_ns = 0;
  // re-sorts of the whole batch in four launches when every member sorts
  // by tile sort: device tables of the jobs (member species with
  // particles) per (layout, buffer) state, their block map, and the jobs'
  // count / own / fill words
  bool tile_sorts = false;
  void *sort_tab = nullptr;
  size_t job_bytes = 0;
  spic::TileSortMap sort_map{};
  unsigned int *sort_ws = nullptr;
  size_t sort_ws_words = 0, sort_ws_cap = 0;
  // the persistent phase launch: its work list (device; tiles of every
  // instance, heaviest first, then the overflow parts) and its two counters
  // (device, zero between launches)
  uint32_t *items = nullptr;
  size_t items_cap = 0;
  uint32_t nitems = 0;
  unsigned int *queue_ctr = nullptr;
};

namespace {

constexpr int kBatchOverflowBlocks = spic::kQueueOverflowParts;  // per instance

// The work list of the persistent phase launch: every tile of every
// instance with particles, in decreasing order of the particles its
// segments held at the last sort (all species), so that the heaviest tiles
// start first; then each instance's overflow parts.
hipError_t build_items(spic_batch *b) {
  std::vector<std::pair<int64_t, uint32_t>> tiles;
  std::vector<uint32_t> items;
  std::vector<int64_t> seg;
  for (int i = 0; i < b->n; ++i) {
    const spic_engine *e = b->e[i];
    const int nt = e->L.ntiles;
    std::vector<int64_t> load(nt, 0);
    bool any = false;
    for (int is = 0; is < e->ns; ++is) {
      const auto &sp = e->sp[is];
      if (!sp.loaded || sp.n == 0) continue;
      any = true;
      seg.resize((size_t)nt + 1);
      hipError_t err = hipMemcpy(seg.data(), sp.seg[sp.segi], sizeof(int64_t) * (nt + 1),
                                 hipMemcpyDeviceToHost);
      if (err != hipSuccess) return err;
      for (int t = 0; t < nt; ++t) load[t] += seg[t + 1] - seg[t];
    }
    if (!any) continue;
    for (int t = 0; t < nt; ++t)
      tiles.emplace_back(load[t], ((uint32_t)i << spic::kQueueInstShift) | (uint32_t)t);
    for (int p = 0; p < kBatchOverflowBlocks; ++p)
      items.push_back(((uint32_t)i << spic::kQueueInstShift) | spic::kQueueOverflowBit |
                      (uint32_t)p);
  }
  std::stable_sort(tiles.begin(), tiles.end(),
                   [](const std::pair<int64_t, uint32_t> &a,
                      const std::pair<int64_t, uint32_t> &c) { return a.first > c.first; });
  std::vector<uint32_t> all;
  all.reserve(tiles.size() + items.size());
  for (const auto &t : tiles) all.push_back(t.second);
  all.insert(all.end(), items.begin(), items.end());
  hipError_t err;
  if (all.size() > b->items_cap) {
    if (b->items) (void)hipFree(b->items);
    b->items = nullptr;
    b->items_cap = 0;
    if ((err = hipMalloc(&b->items, sizeof(uint32_t) * all.size())) != hipSuccess) return err;
    b->items_cap = all.size();
  }
  if (!b->queue_ctr) {
    if ((err = hipMalloc(&b->queue_ctr, 2 * sizeof(unsigned int))) != hipSuccess) return err;
    if ((err = hipMemset(b->queue_ctr, 0, 2 * sizeof(unsigned int))) != hipSuccess) return err;
  }
  if (!all.empty() &&
      (err = hipMemcpy(b->items, all.data(), sizeof(uint32_t) * all.size(),
                       hipMemcpyHostToDevice)) != hipSuccess)
    return err;
  b->nitems = (uint32_t)all.size();
  return hipSuccess;
}

template <typename T>
hipError_t build_tables(spic_batch *b, unsigned int *status) {
  const size_t nb = sizeof(spic::BinnedArgs<T>);
  std::vector<spic::BinnedArgs<T>> host(4 * (size_t)b->n);
  b->max_ns = 0;
  for (int st = 0; st < 4; ++st)
    for (int i = 0; i < b->n; ++i) {
      const int k = fill_args<T>(b->e[i], &b->field[i], b->mom[i], &b->push[i], status,
                                 st & 1, st >> 1, host[(size_t)st * b->n + i], nullptr);
      if (k > b->max_ns) b->max_ns = k;
    }
  hipError_t err;
  if (!b->tab) {
    if ((err = hipMalloc(&b->tab, 4 * nb * b->n)) != hipSuccess) return err;
  }
  if ((err = hipMemcpy(b->tab, host.data(), 4 * nb * b->n, hipMemcpyHostToDevice)) !=
      hipSuccess)
    return err;
  b->arg_bytes = nb;
  auto blocks = [](int64_t items, int per) { return (int)((items + per - 1) / per); };
  b->nodes.n = b->counts.n = b->n;
  b->nodes.start[0] = b->counts.start[0] = 0;
  for (int i = 0; i < b->n; ++i) {
    const spic_engine *e = b->e[i];
    const bool any = host[i].ns > 0;
    const int64_t nn = (int64_t)e->grd.nxn * e->grd.nyn * e->grd.nzn;
    b->nodes.start[i + 1] = b->nodes.start[i] + (any ? blocks(nn, 256) : 0);
    const int64_t ncnt = 3 * (int64_t)(e->L.ntiles + 1);
    const int64_t zeroed =
        (b->flags & SPIC_BATCH_SET_MOMENTS) && nn > ncnt ? nn : ncnt;
    b->counts.start[i + 1] = b->counts.start[i] + (any ? blocks(zeroed, 256) : 0);
    b->gen[i] = e->gen;
  }
  b->status = status;
  if ((err = build_items(b)) != hipSuccess) return err;

  b->tile_sorts = true;
  for (int i = 0; i < b->n; ++i)
    if (b->e[i]->sort_mode != SPIC_SORT_TILE) b->tile_sorts = false;
  if (!b->tile_sorts) return hipSuccess;
  constexpr int kMaxJobs = spic::kMaxInstances * spic::kMaxSpecies;
  const size_t jb = sizeof(spic::TileSortJob<T>);
  size_t words = 0;
  int nj = 0;
  b->sort_map.start[0] = 0;
  for (int i = 0; i < b->n; ++i) {
    const spic_engine *e = b->e[i];
    for (int is = 0; is < e->ns; ++is) {
      if (!e->sp[is].loaded || e->sp[is].n == 0) continue;
      words += 3 * ((size_t)e->L.ntiles + 1);
      b->sort_map.start[nj + 1] = b->sort_map.start[nj] + e->L.ntiles + spic::kTileOvfBlocks;
      ++nj;
    }
  }
  b->sort_map.n = nj;
  if (words > b->sort_ws_cap) {
    if (b->sort_ws) (void)hipFree(b->sort_ws);
    b->sort_ws = nullptr;
    b->sort_ws_cap = 0;
    if ((err = hipMalloc(&b->sort_ws, words * sizeof(unsigned int))) != hipSuccess) return err;
    b->sort_ws_cap = words;
  }
  b->sort_ws_words = words;
  if (!b->sort_tab) {
    if ((err = hipMalloc(&b->sort_tab, 4 * kMaxJobs * jb)) != hipSuccess) return err;
  }
  std::vector<spic::TileSortJob<T>> jobs(4 * (size_t)kMaxJobs);
  for (int st = 0; st < 4; ++st) {
    size_t off = 0;
    int j = 0;
    for (int i = 0; i < b->n; ++i) {
      const spic_engine *e = b->e[i];
      for (int is = 0; is < e->ns; ++is) {
        if (!e->sp[is].loaded || e->sp[is].n == 0) continue;
        jobs[(size_t)st * kMaxJobs + j] =
            tile_sort_job<T>(e, is, st & 1, st >> 1, b->sort_ws + off);
        off += 3 * ((size_t)e->L.ntiles + 1);
        ++j;
      }
    }
  }
  if ((err = hipMemcpy(b->sort_tab, jobs.data(), 4 * kMaxJobs * jb,
                       hipMemcpyHostToDevice)) != hipSuccess)
    return err;
  b->job_bytes = jb;
  return hipSuccess;
}

// A re-sort of every loaded species of the batch, from `state`, by tile
// sorts in four launches.
template <typename T>
hipError_t launch_batch_tile_sort(spic_batch *b, int state, hipStream_t s) {
  constexpr int kMaxJobs = spic::kMaxInstances * spic::kMaxSpecies;
  const spic::TileSortJob<T> *jobs =
      (const spic::TileSortJob<T> *)((char *)b->sort_tab +
                                     (size_t)state * kMaxJobs * b->job_bytes);
  const spic::BinLayout &L = b->e[0]->L;
  return spic::tile_sort_batch<T>(jobs, b->sort_map, L.tx, L.ty, L.tz, b->sort_ws,
                                  b->sort_ws_words, s);
}

// The argument tables hold only species with particles: a loaded species
// without any deposits nothing, so with SPIC_BATCH_SET_MOMENTS its moment
// set is zeroed here (fill kernels, which graph capture records as such).
hipError_t zero_empty_batch_moments(const spic_batch *b, hipStream_t s) {
  for (int i = 0; i < b->n; ++i) {
    const spic_engine *e = b->e[i];
    const size_t nn = (size_t)e->grd.nxn * e->grd.nyn * e->grd.nzn;
    for (int is = 0; is < e->ns; ++is) {
      if (!e->sp[is].loaded || e->sp[is].n != 0 || !b->mom[i][is].rhon) continue;
      const spic::KMom m = spic::kmom(&b->mom[i][is]);
      for (int c = 0; c < 10; ++c) {
        const hipError_t err = spic::zero_f64(m.m[c], nn, s);
        if (err != hipSuccess) return err;
      }
    }
  }
  return hipSuccess;
}

template <typename T>
hipError_t launch_batch(spic_batch *b, int state, hipStream_t s) {
  const spic::BinnedArgs<T> *tab =
      (const spic::BinnedArgs<T> *)((char *)b->tab + (size_t)state * b->n * b->arg_bytes);
  const spic::BinLayout &L = b->e[0]->L;
  const int set = (b->flags & SPIC_BATCH_SET_MOMENTS) ? 1 : 0;
  hipError_t err = spic::launch_batch_zero_counts<T>(tab, b->counts, b->max_ns, set, s);
  if (err == hipSuccess && set) err = zero_empty_batch_moments(b, s);
  if (err != hipSuccess) return err;
  // every instance's tiles and overflow parts in one persistent launch
  err = spic::strict_mode()
            ? spic::spic_strict::launch_binned_phase_queue<T>(
                  tab, b->items, b->nitems, b->queue_ctr, L.tx, L.ty, L.tz,
                  b->e[0]->lds_fp64, L.fixed, s)
            : spic::spic_fast::launch_binned_phase_queue<T>(
                  tab, b->items, b->nitems, b->queue_ctr, L.tx, L.ty, L.tz,
                  b->e[0]->lds_fp64, L.fixed, s);
  if (err != hipSuccess) return err;
  return spic::launch_batch_reduce<T>(tab, b->nodes, b->max_ns, L.tx, L.ty, L.tz, s);
}

}  // namespace

extern "C" {

int spic_batch_create(spic_batch **out, const spic_batch_member *members, int n,
                      int sort_every, int flags) {
  if (!out) return fail(SPIC_EINVAL, "batch pointer is NULL");
  *out = nullptr;
  if (!members) return fail(SPIC_EINVAL, "batch members are NULL");
  if (n < 1 || n > spic::kMaxInstances)
    return fail(SPIC_EINVAL, "batch instance count must be 1..%d", spic::kMaxInstances);
  if (sort_every < 0) return fail(SPIC_EINVAL, "sort_every must be >= 0");
  if (flags & ~SPIC_BATCH_SET_MOMENTS) return fail(SPIC_EINVAL, "unknown batch flags");
  int rc;
  for (int i = 0; i < n; ++i) {
    const spic_batch_member &m = members[i];
    if (!m.engine) return fail(SPIC_EINVAL, "batch member %d: engine is NULL", i);
    if (!m.field || !m.push || !m.moments)
      return fail(SPIC_EINVAL, "batch member %d: field, moments or push is NULL", i);
    if ((rc = spic::check_field(m.field)) || (rc = spic::check_push(m.push))) return rc;
    const spic_engine *e = m.engine, *e0 = members[0].engine;
    if (e->grd.prec != e0->grd.prec || e->L.tx != e0->L.tx || e->L.ty != e0->L.ty ||
        e->L.tz != e0->L.tz || e->lds_fp64 != e0->lds_fp64 ||
        e->sort_mode != e0->sort_mode || e->L.fixed != e0->L.fixed)
      return fail(SPIC_EINVAL,
                  "batch member %d: precision, tile shape, LDS partials, sort mode or "
                  "position storage differ from member 0's", i);
    for (int j = 0; j < i; ++j)
      if (members[j].engine == e)
        return fail(SPIC_EINVAL, "batch members %d and %d share an engine", j, i);
    for (int is = 0; is < e->ns; ++is)
      if (e->sp[is].loaded && (rc = spic::check_moments(&m.moments[is]))) return rc;
    if ((rc = check_subcycles(e))) return rc;
    if (e->ppt != 1)
      return fail(SPIC_EUNSUPPORTED,
                  "batch member %d: batches run one particle per thread and batch", i);
  }
  spic_batch *b = new (std::nothrow) spic_batch();
  if (!b) return fail(SPIC_ENOMEM, "batch allocation");
  b->n = n;
  b->prec = members[0].engine->grd.prec;
  b->sort_every = sort_every;
  b->flags = flags;
  for (int i = 0; i < n; ++i) {
    b->e[i] = members[i].engine;
    b->field[i] = *members[i].field;
    b->push[i] = *members[i].push;
    std::memset(b->mom[i], 0, sizeof(b->mom[i]));
    for (int is = 0; is < b->e[i]->ns; ++is)
      if (b->e[i]->sp[is].loaded) b->mom[i][is] = members[i].moments[is];
  }
  hipError_t err;
  unsigned int *st = spic::status_buffer(&err);
  if (!st) {
    delete b;
    return hip_fail(err, "status buffer");
  }
  err = b->prec == SPIC_F64 ? build_tables<double>(b, st) : build_tables<float>(b, st);
  if (err != hipSuccess) {
    spic_batch_destroy(b);
    return hip_fail(err, "batch argument tables");
  }
  *out = b;
  return SPIC_OK;
}

int spic_batch_phase(spic_batch *b, void *stream) {
  if (!b) return fail(SPIC_EINVAL, "batch is NULL");
  if (int prc = check_positions(b->e[0])) return prc;
  hipStream_t s = (hipStream_t)stream;
  hipError_t err;
  // every loaded species of the batch must be in one (layout, buffer) state
  int state = -1;
  bool stale = false;
  for (int i = 0; i < b->n; ++i) {
    const spic_engine *e = b->e[i];
    if (e->gen != b->gen[i]) stale = true;
    for (int is = 0; is < e->ns; ++is) {
      const auto &sp = e->sp[is];
      if (!sp.loaded || sp.n == 0) continue;
      if (!b->mom[i][is].rhon)
        return fail(SPIC_EINVAL, "batch member %d: species %d loaded after the batch "
                                 "was created without moments", i, is);
      const int st = sp.segi * 2 + sp.cur;
      if (state < 0) state = st;
      else if (st != state)
        return fail(SPIC_EINVAL,
                    "batch species out of step (member %d species %d): load every "
                    "species, then advance them only through the batch", i, is);
    }
  }
  const bool sort_now = b->sort_every > 0 && b->phases > 0 && b->phases % b->sort_every == 0;
  unsigned int *st;
  int rc;
  if ((rc = spic::begin_call(s, &st))) return rc;
  if (state < 0) {
    // no member holds a particle: the written moments are all zero
    err = (b->flags & SPIC_BATCH_SET_MOMENTS) ? zero_empty_batch_moments(b, s) : hipSuccess;
    if (err == hipSuccess) b->phases++;
    return spic::end_call(s, st, err);
  }
  if (stale || st != b->status) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if ((err = hipStreamIsCapturing(s, &cs)) != hipSuccess)
      return hip_fail(err, "batch capture status");
    if (cs != hipStreamCaptureStatusNone)
      return fail(SPIC_EINVAL, "batch argument tables are stale (a member was loaded, or "
                               "another thread calls); run one phase outside capture");
    if ((err = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(err, "batch phase");
    err = b->prec == SPIC_F64 ? build_tables<double>(b, st) : build_tables<float>(b, st);
    if (err != hipSuccess) return hip_fail(err, "batch argument tables");
  }
  if (sort_now && b->tile_sorts) {
    err = b->prec == SPIC_F64 ? launch_batch_tile_sort<double>(b, state, s)
                              : launch_batch_tile_sort<float>(b, state, s);
    if (err != hipSuccess) return hip_fail(err, "batch tile sort");
    for (int i = 0; i < b->n; ++i) {
      spic_engine *e = b->e[i];
      std::lock_guard<std::mutex> lock(e->mu);
      for (int is = 0; is < e->ns; ++is) {
        auto &sp = e->sp[is];
        if (!sp.loaded || sp.n == 0) continue;
        sp.cur = 1 - sp.cur;
        sp.segi = 1 - sp.segi;
        e->sorts++;
      }
    }
    state ^= 3;
  } else if (sort_now) {
    // A sort moves each species to the other buffer and the other layout;
    // members in step stay in one state.
    int after = -1;
    for (int i = 0; i < b->n; ++i) {
      spic_engine *e = b->e[i];
      std::lock_guard<std::mutex> lock(e->mu);
      for (int is = 0; is < e->ns; ++is) {
        auto &sp = e->sp[is];
        if (!sp.loaded || sp.n == 0) continue;
        if ((err = sort_any(e, is, nullptr, s)) != hipSuccess)
          return hip_fail(err, "batch sort");
        const int st2 = sp.segi * 2 + sp.cur;
        if (after < 0) after = st2;
        else if (st2 != after)
          return fail(SPIC_EINVAL,
                      "batch members out of step after a sort (member %d species %d): "
                      "their sort modes or sort histories differ", i, is);
      }
    }
    state = after;
  }
  err = b->prec == SPIC_F64 ? launch_batch<double>(b, state, s)
                            : launch_batch<float>(b, state, s);
  if (err == hipSuccess) {
    for (int i = 0; i < b->n; ++i) {
      spic_engine *e = b->e[i];
      for (int is = 0; is < e->ns; ++is)
        if (e->sp[is].loaded && e->sp[is].n > 0) e->sp[is].cur ^= 1;
      e->phases++;
    }
    b->phases++;
  }
  return spic::end_call(s, st, err);
}

int spic_batch_destroy(spic_batch *b) {
  if (!b) return SPIC_OK;
  if (b->tab || b->sort_tab || b->sort_ws || b->items) (void)hipDeviceSynchronize();
  if (b->items) (void)hipFree(b->items);
  if (b->queue_ctr) (void)hipFree(b->queue_ctr);
  if (b->tab) (void)hipFree(b->tab);
  if (b->sort_tab) (void)hipFree(b->sort_tab);
  if (b->sort_ws) (void)hipFree(b->sort_ws);
  delete b;
  return SPIC_OK;
}

}  // extern "C"
