// spic_capi.cpp -- the extern "C" entry points of libspic.so (include/spic.h).
// Validates arguments, forms the per-species push constants exactly as the
// reference does, dispatches to the fast or strict kernel set in the
// requested precision, and turns device-side domain violations and HIP
// errors into status codes.  Nothing here computes on particles.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "spic.h"
#include "spic_internal.hpp"

namespace spic {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory)
    return fail(SPIC_ENOMEM, "%s: %s", what, hipGetErrorString(e));
  return fail(SPIC_EHIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace spic

namespace {

std::atomic<int> g_async{0};
std::atomic<int> g_strict{0};
std::atomic<int> g_subcycle{SPIC_SUBCYCLE_REFUSE};
using spic::fail;
using spic::hip_fail;
using spic::kgrid;
using spic::kfield;
using spic::kpart;
using spic::kmom;
using spic::kpush;

}  // namespace

namespace spic {
// Per-thread, per-device status words: [0] mover iterates out of the grid,
// [1] deposit cells out of the grid, [2] engine particle storage full.  Thread-local so that concurrent
// callers (the reference runs interpP2G from an OpenMP team, one thread per
// species, src/sputniPIC_CPU.cpp:250-256) never clear each other's flags.
struct StatusBuf {
  int device = -1;
  unsigned int *d = nullptr;
};
thread_local std::vector<StatusBuf> t_status;

unsigned int *status_buffer(hipError_t *err) {
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess) return nullptr;
  for (auto &sb : t_status)
    if (sb.device == dev) return sb.d;
  StatusBuf sb;
  sb.device = dev;
  *err = hipMalloc((void **)&sb.d, 4 * sizeof(unsigned int));
  if (*err != hipSuccess) return nullptr;
  *err = hipMemset(sb.d, 0, 4 * sizeof(unsigned int));
  if (*err != hipSuccess) return nullptr;
  t_status.push_back(sb);
  return sb.d;
}
}  // namespace spic

namespace {
using spic::status_buffer;

// Workspace for the device sort, one per device, grown on demand.
std::mutex g_ws_mutex;
struct Workspace {
  int device = -1;
  void *p = nullptr;
  size_t bytes = 0;
};
std::vector<Workspace> g_ws;

hipError_t workspace(size_t bytes, void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  for (auto &w : g_ws)
    if (w.device == dev) {
      if (w.bytes < bytes) {
        if (w.p) (void)hipFree(w.p);
        w.p = nullptr;
        w.bytes = 0;
        if ((e = hipMalloc(&w.p, bytes)) != hipSuccess) return e;
        w.bytes = bytes;
      }
      *out = w.p;
      return hipSuccess;
    }
  Workspace w;
  w.device = dev;
  if ((e = hipMalloc(&w.p, bytes)) != hipSuccess) return e;
  w.bytes = bytes;
  g_ws.push_back(w);
  *out = w.p;
  return hipSuccess;
}

}  // namespace

namespace spic {
int check_grid(const spic_grid *g) {
  if (!g) return fail(SPIC_EINVAL, "grid is NULL");
  if (g->prec != SPIC_F64 && g->prec != SPIC_F32)
    return fail(SPIC_EINVAL, "grid.prec must be SPIC_F64 or SPIC_F32");
  if (g->nxn < 3 || g->nyn < 3 || g->nzn < 3)
    return fail(SPIC_EINVAL, "grid node counts must be >= 3");
  if (g->nxn != g->nxc + 1 || g->nyn != g->nyc + 1 || g->nzn != g->nzc + 1)
    return fail(SPIC_EINVAL, "grid node counts must be cell counts + 1");
  if ((double)g->nxn * g->nyn * g->nzn > 2147483647.0)
    return fail(SPIC_EUNSUPPORTED, "grid has more than 2^31-1 nodes");
  if (!g->xn || !g->yn || !g->zn)
    return fail(SPIC_EINVAL, "grid node-coordinate tables are NULL");
  if (!(g->Lx > 0) || !(g->Ly > 0) || !(g->Lz > 0))
    return fail(SPIC_EINVAL, "grid lengths must be positive");
  return SPIC_OK;
}

int check_species(const spic_species *sp, bool need_q, bool need_move) {
  if (!sp) return fail(SPIC_EINVAL, "species is NULL");
  if (sp->nop < 0 || sp->npmax < sp->nop)
    return fail(SPIC_EINVAL, "species needs 0 <= nop <= npmax");
  if (sp->npmax > 2147483647LL)
    return fail(SPIC_EUNSUPPORTED, "more than 2^31-1 particles in a species");
  if (sp->npmax > 0 && (!sp->x || !sp->y || !sp->z || !sp->u || !sp->v || !sp->w))
    return fail(SPIC_EINVAL, "species position/velocity array is NULL");
  if (need_q && sp->npmax > 0 && !sp->q)
    return fail(SPIC_EINVAL, "species charge array is NULL");
  if (need_move) {
    if (sp->NiterMover < 1)
      return fail(SPIC_EINVAL, "NiterMover must be >= 1");
    if (sp->n_sub_cycles < 1) return fail(SPIC_EINVAL, "n_sub_cycles must be >= 1");
    if (sp->n_sub_cycles != 1 && g_subcycle.load() == SPIC_SUBCYCLE_REFUSE)
      return fail(SPIC_EUNSUPPORTED,
                  "n_sub_cycles = %d: the reference's CPU mover pushes once with dt / "
                  "n_sub_cycles (src/Mover.cpp:11), its GPU mover loops; choose one with "
                  "spic_set_subcycling",
                  sp->n_sub_cycles);
  }
  return SPIC_OK;
}

int check_field(const spic_field *f) {
  if (!f || !f->Ex || !f->Ey || !f->Ez || !f->Bxn || !f->Byn || !f->Bzn)
    return fail(SPIC_EINVAL, "field array is NULL");
  return SPIC_OK;
}

int check_moments(const spic_moments *m) {
  if (!m || !m->rhon || !m->Jx || !m->Jy || !m->Jz || !m->pxx || !m->pxy ||
      !m->pxz || !m->pyy || !m->pyz || !m->pzz)
    return fail(SPIC_EINVAL, "moment array is NULL");
  return SPIC_OK;
}

int check_push(const spic_push *p) {
  if (!p) return fail(SPIC_EINVAL, "push parameters are NULL");
  if (!(p->c != 0)) return fail(SPIC_EINVAL, "c must be nonzero");
  return SPIC_OK;
}
}  // namespace spic

namespace {
using spic::check_grid;
using spic::check_species;
using spic::check_field;
using spic::check_moments;
using spic::check_push;

}  // namespace

namespace spic {
bool async_mode() { return g_async.load() != 0; }
bool strict_mode() { return g_strict.load() != 0; }
int subcycle_mode() { return g_subcycle.load(); }

// Prepare the status words for one call; returns the device pointer.
int begin_call(hipStream_t s, unsigned int **st) {
  hipError_t e;
  *st = status_buffer(&e);
  if (!*st) return hip_fail(e, "status buffer");
  if (!g_async.load()) {
    e = spic::zero_u32(*st, 3, s);
    if (e != hipSuccess) return hip_fail(e, "status reset");
  }
  return SPIC_OK;
}

int read_status(hipStream_t s, unsigned int *st, bool clear) {
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "kernel execution");
  unsigned int h[3] = {0, 0, 0};
  e = hipMemcpy(h, st, sizeof(h), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "status readback");
  if (clear && (h[0] || h[1] || h[2])) {
    e = hipMemset(st, 0, sizeof(h));
    if (e != hipSuccess) return hip_fail(e, "status reset");
  }
  if (h[2])
    return fail(SPIC_ENOMEM,
                "engine particle storage full: %u workgroups could not place "
                "a particle", h[2]);
  if (h[0] || h[1])
    return fail(SPIC_EDOMAIN,
                "particles left the grid by more than one cell (%u mover "
                "blocks, %u deposit blocks flagged); the reference would "
                "index outside its arrays",
                h[0], h[1]);
  return SPIC_OK;
}

int end_call(hipStream_t s, unsigned int *st, hipError_t launch) {
  if (launch != hipSuccess) return hip_fail(launch, "kernel launch");
  if (g_async.load()) return SPIC_OK;
  return read_status(s, st, false);
}

}  // namespace spic

using spic::begin_call;
using spic::end_call;
using spic::read_status;
using spic::status_buffer;

extern "C" {

const char *spic_last_error(void) { return spic::g_err; }
int spic_abi_version(void) { return SPIC_ABI_VERSION; }

int spic_set_async(int async_mode) {
  g_async.store(async_mode ? 1 : 0);
  return SPIC_OK;
}

int spic_set_strict_fp(int strict) {
  g_strict.store(strict ? 1 : 0);
  return SPIC_OK;
}

int spic_set_subcycling(int mode) {
  if (mode != SPIC_SUBCYCLE_REFUSE && mode != SPIC_SUBCYCLE_LOOP &&
      mode != SPIC_SUBCYCLE_SINGLE)
    return fail(SPIC_EINVAL, "sub-cycling mode must be SPIC_SUBCYCLE_REFUSE, _LOOP or _SINGLE");
  g_subcycle.store(mode);
  return SPIC_OK;
}

int spic_check_domain(void *stream) {
  hipError_t e;
  unsigned int *st = status_buffer(&e);
  if (!st) return hip_fail(e, "status buffer");
  return read_status((hipStream_t)stream, st, true);
}

int spic_mover_pc(const spic_grid *grd, const spic_field *field,
                  spic_species *sp, const spic_push *push, void *stream) {
  int rc;
  if ((rc = check_grid(grd)) || (rc = check_field(field)) ||
      (rc = check_species(sp, false, true)) || (rc = check_push(push)))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned int *st;
  if ((rc = begin_call(s, &st))) return rc;
  hipError_t e;
  const bool strict = g_strict.load() != 0;
  if (grd->prec == SPIC_F64) {
    auto g = kgrid<double>(grd);
    auto f = kfield<double>(field);
    auto p = kpart<double>(sp);
    auto k = kpush<double>(sp, push);
    e = strict ? spic::spic_strict::launch_mover<double>(g, f, p, k, st, s)
               : spic::spic_fast::launch_mover<double>(g, f, p, k, st, s);
  } else {
    auto g = kgrid<float>(grd);
    auto f = kfield<float>(field);
    auto p = kpart<float>(sp);
    auto k = kpush<float>(sp, push);
    e = strict ? spic::spic_strict::launch_mover<float>(g, f, p, k, st, s)
               : spic::spic_fast::launch_mover<float>(g, f, p, k, st, s);
  }
  return end_call(s, st, e);
}

int spic_interp_p2g(const spic_grid *grd, const spic_species *sp,
                    spic_moments *mom, void *stream) {
  int rc;
  if ((rc = check_grid(grd)) || (rc = check_species(sp, true, false)) ||
      (rc = check_moments(mom)))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned int *st;
  if ((rc = begin_call(s, &st))) return rc;
  hipError_t e;
  const bool strict = g_strict.load() != 0;
  auto m = kmom(mom);
  if (grd->prec == SPIC_F64) {
    auto g = kgrid<double>(grd);
    auto p = kpart<double>(sp);
    e = strict ? spic::spic_strict::launch_deposit<double>(g, p, m, st, s)
               : spic::spic_fast::launch_deposit<double>(g, p, m, st, s);
  } else {
    auto g = kgrid<float>(grd);
    auto p = kpart<float>(sp);
    e = strict ? spic::spic_strict::launch_deposit<float>(g, p, m, st, s)
               : spic::spic_fast::launch_deposit<float>(g, p, m, st, s);
  }
  return end_call(s, st, e);
}

int spic_mover_interp(const spic_grid *grd, const spic_field *field,
                      spic_species *sp, spic_moments *mom,
                      const spic_push *push, void *stream) {
  int rc;
  if ((rc = check_grid(grd)) || (rc = check_field(field)) ||
      (rc = check_species(sp, true, true)) || (rc = check_moments(mom)) ||
      (rc = check_push(push)))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned int *st;
  if ((rc = begin_call(s, &st))) return rc;
  hipError_t e;
  const bool strict = g_strict.load() != 0;
  auto m = kmom(mom);
  if (grd->prec == SPIC_F64) {
    auto g = kgrid<double>(grd);
    auto f = kfield<double>(field);
    auto p = kpart<double>(sp);
    auto k = kpush<double>(sp, push);
    e = strict ? spic::spic_strict::launch_mover_deposit<double>(g, f, p, k, m, st, s)
               : spic::spic_fast::launch_mover_deposit<double>(g, f, p, k, m, st, s);
  } else {
    auto g = kgrid<float>(grd);
    auto f = kfield<float>(field);
    auto p = kpart<float>(sp);
    auto k = kpush<float>(sp, push);
    e = strict ? spic::spic_strict::launch_mover_deposit<float>(g, f, p, k, m, st, s)
               : spic::spic_fast::launch_mover_deposit<float>(g, f, p, k, m, st, s);
  }
  return end_call(s, st, e);
}

int spic_particle_sort(const spic_grid *grd, int64_t sort_cps,
                       spic_species *sp, void *stream) {
  int rc;
  if ((rc = check_grid(grd)) || (rc = check_species(sp, true, false)))
    return rc;
  if (sort_cps < 1) return fail(SPIC_EINVAL, "sort_cps must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  size_t need = 0;
  hipError_t e =
      grd->prec == SPIC_F64
          ? spic::sort_reference_order<double>(*grd, sort_cps, *sp, nullptr, 0, &need, s)
          : spic::sort_reference_order<float>(*grd, sort_cps, *sp, nullptr, 0, &need, s);
  if (e != hipSuccess) return hip_fail(e, "sort sizing");
  void *ws = nullptr;
  if ((e = workspace(need, &ws)) != hipSuccess) return hip_fail(e, "sort workspace");
  e = grd->prec == SPIC_F64
          ? spic::sort_reference_order<double>(*grd, sort_cps, *sp, ws, need, nullptr, s)
          : spic::sort_reference_order<float>(*grd, sort_cps, *sp, ws, need, nullptr, s);
  if (e != hipSuccess) return hip_fail(e, "sort");
  if (!g_async.load()) {
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "sort execution");
  }
  return SPIC_OK;
}

int spic_zero_moments(const spic_grid *grd, spic_moments *mom, void *stream) {
  int rc;
  if ((rc = check_grid(grd)) || (rc = check_moments(mom))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const size_t nn = (size_t)grd->nxn * grd->nyn * grd->nzn;
  double *arr[10] = {mom->rhon, mom->Jx,  mom->Jy,  mom->Jz,  mom->pxx,
                     mom->pxy,  mom->pxz, mom->pyy, mom->pyz, mom->pzz};
  for (int m = 0; m < 10; ++m) {
    hipError_t e = spic::zero_f64(arr[m], nn, s);
    if (e != hipSuccess) return hip_fail(e, "zero moments");
  }
  if (!g_async.load()) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "zero moments");
  }
  return SPIC_OK;
}

int spic_device_count(int *count) {
  if (!count) return fail(SPIC_EINVAL, "count is NULL");
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return hip_fail(e, "hipGetDeviceCount");
  }
  return SPIC_OK;
}

int spic_malloc(void **dptr, size_t bytes) {
  if (!dptr) return fail(SPIC_EINVAL, "dptr is NULL");
  *dptr = nullptr;
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipMalloc");
}

int spic_free(void *dptr) {
  hipError_t e = hipFree(dptr);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipFree");
}

int spic_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipMemcpy H2D");
}

int spic_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipMemcpy D2H");
}

int spic_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipMemcpy D2D");
}

int spic_memset(void *dptr, int value, size_t bytes) {
  hipError_t e = hipMemset(dptr, value, bytes);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipMemset");
}

int spic_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipDeviceSynchronize");
}

int spic_host_alloc(void **hptr, size_t bytes) {
  if (!hptr) return fail(SPIC_EINVAL, "hptr is NULL");
  *hptr = nullptr;
  hipError_t e = hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipHostMalloc");
}

int spic_host_free(void *hptr) {
  hipError_t e = hipHostFree(hptr);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipHostFree");
}

int spic_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipMemcpyAsync D2H");
}

int spic_stream_wait_event(void *stream, void *event) {
  if (!event) return fail(SPIC_EINVAL, "event is NULL");
  hipError_t e = hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipStreamWaitEvent");
}

int spic_stream_create(void **stream) {
  if (!stream) return fail(SPIC_EINVAL, "stream is NULL");
  hipStream_t s;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  *stream = (void *)s;
  return SPIC_OK;
}

int spic_stream_destroy(void *stream) {
  hipError_t e = hipStreamDestroy((hipStream_t)stream);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipStreamDestroy");
}

int spic_stream_synchronize(void *stream) {
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipStreamSynchronize");
}

}  // extern "C"

struct spic_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int spic_graph_begin(void *stream) {
  if (!stream) return fail(SPIC_EINVAL, "graph capture needs a created stream, not NULL");
  if (!g_async.load())
    return fail(SPIC_EINVAL, "graph capture needs async mode (spic_set_async(1))");
  // the per-thread status words must exist before capture: no allocation
  // may happen while the stream is captured
  hipError_t e;
  if (!status_buffer(&e)) return hip_fail(e, "status buffer");
  e = hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipStreamBeginCapture");
}

int spic_graph_end(void *stream, spic_graph **out) {
  if (!out) return fail(SPIC_EINVAL, "graph pointer is NULL");
  *out = nullptr;
  spic_graph *g = new (std::nothrow) spic_graph();
  if (!g) return fail(SPIC_ENOMEM, "graph allocation");
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    spic_graph_destroy(g);
    return hip_fail(e, "graph capture");
  }
  *out = g;
  return SPIC_OK;
}

int spic_graph_launch(spic_graph *g, void *stream) {
  if (!g || !g->exec) return fail(SPIC_EINVAL, "graph is NULL");
  hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  if (e == hipSuccess && !g_async.load()) e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? SPIC_OK : hip_fail(e, "hipGraphLaunch");
}

int spic_graph_destroy(spic_graph *g) {
  if (!g) return SPIC_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return SPIC_OK;
}

}  // extern "C"
