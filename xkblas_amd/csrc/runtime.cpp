// runtime.cpp — the xkblas.h C API on one MI355X.
//
// What replaces the reference runtime (XKaapi DAG + DSM + HIP plugin):
//   * Tile cache (reference kaapi/kaapi_memory.c: replica bitmaps, LRU,
//     fetch/writeback).  Each host matrix known to the library (keyed by its
//     host pointer, as xkblas_find does, blas/xkblas.c:426-455) owns one
//     column-major mirror in HBM with a padded leading dimension; validity is
//     tracked per NB x NB tile (device-valid / host-valid bits).  A call
//     fetches the invalid tiles it reads with pinned-or-pageable
//     hipMemcpy2DAsync on the H2D stream; outputs become device-only until
//     xkblas_memory_coherent_async copies the dirty tiles of the requested
//     triangle back on the D2H stream (blas/xkblas.c:1953-2055).  A
//     capacity limit (XKBLAS_CACHE_LIMIT) evicts whole least-recently-used
//     mirrors after writing their dirty tiles back.
//   * Scheduling (reference kaapi/kaapi_offload_device.c, one device thread
//     per GPU with event polling): calls are enqueued in order on one compute
//     stream; copy streams are ordered against it with events; the caller
//     blocks only in xkblas_sync (blas/xkblas.c:1773-1850).
//   * Errors: illegal arguments abort like kaapi_error (blas/common.gen:
//     312-316) unless XKBLAS_ERROR_RETURN=1, HIP errors abort.
#include "blas3.h"
#include "kernels.h"
#include "split.h"
#include "runtime.h"
#include "../../include/xkblas.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <ctime>
#include <functional>
#include <thread>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace xk {

[[noreturn]] void fatal(const char* what, const char* file, int line, const char* msg)
{
  fprintf(stderr, "[xkblas_amd] fatal: %s (%s:%d): %s\n", what, file, line, msg);
  fflush(stderr);
  abort();
}

Runtime& rt()
{
  static Runtime r;
  return r;
}

static size_t env_size(const char* name, size_t dflt)
{
  const char* v = getenv(name);
  if (!v || !*v) return dflt;
  return (size_t)strtoull(v, nullptr, 10);
}

// Workspace of an allocation-free entry point (a fixed device array carved
// into per-slot regions); empty outside such a call.
static struct {
  char* base = nullptr;
  size_t off[Runtime::NSCRATCH] = {}, cap[Runtime::NSCRATCH] = {};
  int nslot = 0;
} g_arena;

void scratch_arena_set(char* base, const size_t* off, const size_t* cap, int nslot)
{
  XK_CHECK(nslot > 0 && nslot <= Runtime::NSCRATCH, "bad arena");
  g_arena.base = base;
  g_arena.nslot = nslot;
  for (int i = 0; i < nslot; ++i) {
    g_arena.off[i] = off[i];
    g_arena.cap[i] = cap[i];
  }
}

void scratch_arena_clear()
{
  g_arena.base = nullptr;
  g_arena.nslot = 0;
}

void* scratch(int slot, size_t bytes)
{
  Runtime& r = rt();
  XK_CHECK(slot >= 0 && slot < Runtime::NSCRATCH, "bad scratch slot");
  if (g_arena.base) {
    // an allocation-free call: its workspace layout covers every request
    XK_CHECK(slot < g_arena.nslot && bytes <= g_arena.cap[slot], "workspace request beyond the arena's layout");
    return g_arena.base + g_arena.off[slot];
  }
  if (bytes == 0) bytes = 256;
  if (r.scratch_size[slot] < bytes) {
    size_t sz = std::max(bytes + bytes / 4, (size_t)16 << 20);
    if (r.scratch_ptr[slot]) {
      // kernels already enqueued may still use the old buffer: it is freed
      // at the next synchronisation point
      r.retired_scratch.emplace_back(r.scratch_ptr[slot], r.scratch_size[slot]);
      r.scratch_ptr[slot] = nullptr;
      r.scratch_size[slot] = 0;
    }
    // scratch counts against the cache limit: mirrors outside the running
    // call make room first (best effort: the call's own operands stay)
    r.evict_for(sz, r.call_keep, false);
    r.scratch_ptr[slot] = r.device_alloc(sz, r.call_keep);
    r.scratch_size[slot] = sz;
    r.cache_used += sz;
  }
  return r.scratch_ptr[slot];
}

// ---------------------------------------------------------------- tile cache
static long pad_ld(long m, size_t elt)
{
  const long q = 256 / (long)elt;  // 256-byte aligned columns
  return std::max(q, (m + q - 1) / q * q);
}

std::vector<Replica*>& Runtime::begin_call()
{
  call_keep.clear();
  return call_keep;
}

// Evict least-recently-used mirrors outside `keep` (dirty tiles written back
// first) until `need` more bytes fit the limit.  strict: an operand set that
// cannot fit is fatal; otherwise returns false.
// A mirror whose every tile is also valid on the host: dropping it moves no
// data.
static bool mirror_clean(const Replica* r)
{
  for (size_t t = 0; t < r->state.size(); ++t)
    if (!(r->state[t] & HOST_VALID)) return false;
  return true;
}

// Victims: clean mirrors first (no write-back, no wait), least recently used
// first among them; then the least recently used dirty one.
bool Runtime::evict_for(size_t need, const std::vector<Replica*>& keep, bool strict)
{
  while (cache_used + need > cache_limit) {
    Replica* victim = nullptr;
    bool vclean = false;
    for (auto& kv : map) {
      Replica* r = kv.second;
      if (std::find(keep.begin(), keep.end(), r) != keep.end()) continue;
      const bool clean = mirror_clean(r);
      if (!victim || (clean && !vclean) || (clean == vclean && r->lru < victim->lru)) {
        victim = r;
        vclean = clean;
      }
    }
    if (!victim) {
      // scratch counts against the limit, but only the operands themselves
      // beyond it are fatal: scratch may overshoot the limit
      size_t scratch_bytes = 0;
      for (int i = 0; i < NSCRATCH; ++i) scratch_bytes += scratch_size[i];
      for (const auto& b : retired_scratch) scratch_bytes += b.second;
      if (!strict || cache_used - std::min(cache_used, scratch_bytes) + need <= cache_limit) return false;
      char msg[256];
      snprintf(msg, sizeof msg, "operands need %zu bytes of HBM beyond the cache limit of %zu bytes", need,
               cache_limit);
      fatal("tile cache", __FILE__, __LINE__, msg);
    }
    if (!vclean) {
      writeback(victim, 0, 0, victim->M, victim->N, 0);
      XK_HIP(hipStreamSynchronize(s_d2h));
      drain_stages();
    }
    drop(victim);
    stats.evictions++;
  }
  return true;
}

// hipMalloc; when the device is out of memory, every mirror outside `keep`
// is evicted and the allocation tried once more.
void* Runtime::device_alloc(size_t bytes, const std::vector<Replica*>& keep)
{
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    sync_all();
    // every stream has drained: outgrown scratch buffers can go first
    for (auto& b : retired_scratch) {
      XK_HIP(hipFree(b.first));
      cache_used -= b.second;
    }
    retired_scratch.clear();
    e = hipMalloc(&p, bytes);
  }
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    std::vector<Replica*> victims;
    for (auto& kv : map)
      if (std::find(keep.begin(), keep.end(), kv.second) == keep.end()) victims.push_back(kv.second);
    for (Replica* v : victims) {
      writeback(v, 0, 0, v->M, v->N, 0);
      XK_HIP(hipStreamSynchronize(s_d2h));
      drain_stages();
      drop(v);
      stats.evictions++;
    }
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    char msg[160];
    snprintf(msg, sizeof msg, "hipMalloc of %zu bytes: %s", bytes, hipGetErrorString(e));
    fatal("device memory", __FILE__, __LINE__, msg);
  }
  return p;
}

void Runtime::drop(Replica* r)
{
  map.erase(r->host);
  clear_panels(r);
  if (r->dev) {
    XK_HIP(hipDeviceSynchronize());
    XK_HIP(hipFree(r->dev));
  }
  cache_used -= r->bytes;
  if (r->info_slot >= 0) free_info.push_back(r->info_slot);
  delete r;
}

// Retired mirrors, at a public synchronisation point (no call is running and
// every stream has drained).
void Runtime::free_retired()
{
  for (auto& b : retired_scratch) {
    XK_HIP(hipFree(b.first));
    cache_used -= b.second;
  }
  retired_scratch.clear();
  for (Replica* r : retired) {
    clear_panels(r);
    if (r->dev) XK_HIP(hipFree(r->dev));
    cache_used -= r->bytes;
    if (r->info_slot >= 0) free_info.push_back(r->info_slot);
    delete r;
  }
  retired.clear();
}

int Runtime::take_info_slot()
{
  if (!free_info.empty()) {
    const int slot = free_info.back();
    free_info.pop_back();
    return slot;
  }
  if (next_info == info_slots) {
    // every slot is taken: grow both arrays (pending copies into the old host
    // array and kernels writing the old device array finish first)
    sync_all();
    const int n2 = info_slots * 2;
    int *d2 = nullptr, *h2 = nullptr;
    XK_HIP(hipMalloc(&d2, sizeof(int) * n2));
    XK_HIP(hipHostMalloc(&h2, sizeof(int) * n2, hipHostMallocDefault));
    for (int i = 0; i < n2; ++i) h2[i] = i < info_slots ? h_info[i] : INT_MAX;
    XK_HIP(hipFree(d_info));
    XK_HIP(hipHostFree(h_info));
    d_info = d2;
    h_info = h2;
    info_slots = n2;
  }
  return next_info++;
}

Replica* Runtime::replica(const void* host, int M, int N, long ld, int elt, int nb, std::vector<Replica*>& keep)
{
  auto it = map.find(host);
  if (it != map.end()) {
    Replica* r = it->second;
    if (r->M == M && r->N == N && r->ld == ld && r->elt == elt) {
      r->lru = ++clock;
      keep.push_back(r);
      return r;
    }
    // The reference keeps the stale descriptor (blas/xkblas.c:416-421); here
    // a pointer reused with new dimensions gets a fresh mirror, read from
    // host memory once the old one's dirty tiles are back there.
    writeback(r, 0, 0, r->M, r->N, 0);
    XK_HIP(hipStreamSynchronize(s_d2h));
    drain_stages();
    if (std::find(keep.begin(), keep.end(), r) != keep.end()) {
      // an operand of this call with other dimensions (e.g. dgemm with B
      // aliasing A): the old mirror stays alive, unkeyed, until the call's
      // work has drained
      map.erase(host);
      retired.push_back(r);
    } else {
      drop(r);
    }
  }
  Replica* r = new Replica();
  r->owner = std::this_thread::get_id();
  r->host = host;
  r->M = M; r->N = N; r->ld = ld; r->elt = elt;
  r->nb = nb;
  r->mt = (M + nb - 1) / nb;
  r->nt = (N + nb - 1) / nb;
  r->ldd = pad_ld(M, elt);
  r->bytes = (size_t)r->ldd * (size_t)std::max(N, 1) * elt;
  r->state.assign((size_t)r->mt * r->nt, HOST_VALID);
  evict_for(r->bytes, keep);
  r->dev = (char*)device_alloc(r->bytes, keep);
  cache_used += r->bytes;
  r->lru = ++clock;
  map[host] = r;
  keep.push_back(r);
  return r;
}

// Fetch the host-valid-only tiles of r intersecting rows [r0,r0+m) x cols
// [c0,c0+n) to the device (H2D stream).
void Runtime::fetch(Replica* r, int r0, int c0, int m, int n)
{
  if (m <= 0 || n <= 0) return;
  const int ti0 = r0 / r->nb, ti1 = (r0 + m - 1) / r->nb;
  const int tj0 = c0 / r->nb, tj1 = (c0 + n - 1) / r->nb;
  for (int tj = tj0; tj <= tj1; ++tj) {
    // coalesce vertical runs of invalid tiles into one 2-D copy
    int ti = ti0;
    while (ti <= ti1) {
      if (r->state[tj * r->mt + ti] & DEV_VALID) { ++ti; continue; }
      int te = ti;
      while (te + 1 <= ti1 && !(r->state[tj * r->mt + te + 1] & DEV_VALID)) ++te;
      const int row0 = ti * r->nb, row1 = std::min(r->M, (te + 1) * r->nb);
      const int col0 = tj * r->nb, col1 = std::min(r->N, (tj + 1) * r->nb);
      const char* src = (const char*)r->host + ((size_t)row0 + (size_t)col0 * r->ld) * r->elt;
      char* dst = r->dev + ((size_t)row0 + (size_t)col0 * r->ldd) * r->elt;
      copy_h2d(dst, r->ldd * r->elt, src, r->ld * r->elt, (size_t)(row1 - row0) * r->elt, (size_t)(col1 - col0));
      stats.h2d_bytes += (double)(row1 - row0) * (col1 - col0) * r->elt;
      for (int t = ti; t <= te; ++t) r->state[tj * r->mt + t] |= DEV_VALID;
      pending_h2d = true;
      ti = te + 1;
    }
  }
}

bool Runtime::all_valid(const Replica* r, int r0, int c0, int m, int n) const
{
  if (m <= 0 || n <= 0) return true;
  const int ti0 = r0 / r->nb, ti1 = (r0 + m - 1) / r->nb;
  const int tj0 = c0 / r->nb, tj1 = (c0 + n - 1) / r->nb;
  for (int tj = tj0; tj <= tj1; ++tj)
    for (int ti = ti0; ti <= ti1; ++ti)
      if (!(r->state[tj * r->mt + ti] & DEV_VALID)) return false;
  return true;
}

void Runtime::clear_panels(Replica* r)
{
  for (auto& pe : r->panels) XK_HIP(hipEventDestroy(pe.ev));  // released once complete
  r->panels.clear();
}

// Mark tiles as written on the device (host copy stale).  Any later write
// orders after the whole previous computation, so per-panel events lapse.
void Runtime::mark_written(Replica* r, int r0, int c0, int m, int n, int uplo)
{
  if (m <= 0 || n <= 0) return;
  clear_panels(r);
  const int ti0 = r0 / r->nb, ti1 = (r0 + m - 1) / r->nb;
  const int tj0 = c0 / r->nb, tj1 = (c0 + n - 1) / r->nb;
  for (int tj = tj0; tj <= tj1; ++tj)
    for (int ti = ti0; ti <= ti1; ++ti) {
      if (uplo == Lower && ti < tj) continue;
      if (uplo == Upper && ti > tj) continue;
      r->state[tj * r->mt + ti] = DEV_VALID;
    }
}

// Enqueue device-to-host copies of the dirty tiles in the uplo part
// (0 = whole matrix) of rows [r0,..) x cols [c0,..).
void Runtime::writeback(Replica* r, int r0, int c0, int m, int n, int uplo)
{
  if (m <= 0 || n <= 0) return;
  if (!r->panels.empty()) {
    // a product pipelined by row panels: each panel's rows go back once
    // that panel is computed, under the computation of the next ones
    int lo = 0;
    const std::vector<Replica::PanelEvent> pv = r->panels;
    for (const auto& pe : pv) {
      const int a = std::max(lo, r0), b = std::min(pe.end, r0 + m);
      if (a < b) {
        XK_HIP(hipStreamWaitEvent(s_d2h, pe.ev, 0));
        writeback_rows(r, a, c0, b - a, n, uplo);
      }
      lo = pe.end;
    }
    if (lo < r0 + m) {
      XK_HIP(hipEventRecord(ev_comp, s_comp));
      XK_HIP(hipStreamWaitEvent(s_d2h, ev_comp, 0));
      writeback_rows(r, std::max(lo, r0), c0, r0 + m - std::max(lo, r0), n, uplo);
    }
    return;
  }
  XK_HIP(hipEventRecord(ev_comp, s_comp));
  XK_HIP(hipStreamWaitEvent(s_d2h, ev_comp, 0));
  writeback_rows(r, r0, c0, m, n, uplo);
}

// The D2H copies of writeback (the stream's ordering is set by the caller).
void Runtime::writeback_rows(Replica* r, int r0, int c0, int m, int n, int uplo)
{
  if (m <= 0 || n <= 0) return;
  const int ti0 = r0 / r->nb, ti1 = (r0 + m - 1) / r->nb;
  const int tj0 = c0 / r->nb, tj1 = (c0 + n - 1) / r->nb;
  bool any = false;
  for (int tj = tj0; tj <= tj1; ++tj) {
    int ti = ti0;
    auto wanted = [&](int t) {
      if (uplo == Lower && t < tj) return false;
      if (uplo == Upper && t > tj) return false;
      return !(r->state[tj * r->mt + t] & HOST_VALID);
    };
    while (ti <= ti1) {
      if (!wanted(ti)) { ++ti; continue; }
      int te = ti;
      while (te + 1 <= ti1 && wanted(te + 1)) ++te;
      const int row0 = ti * r->nb, row1 = std::min(r->M, (te + 1) * r->nb);
      const int col0 = tj * r->nb, col1 = std::min(r->N, (tj + 1) * r->nb);
      char* dst = (char*)r->host + ((size_t)row0 + (size_t)col0 * r->ld) * r->elt;
      const char* src = r->dev + ((size_t)row0 + (size_t)col0 * r->ldd) * r->elt;
      copy_d2h(dst, r->ld * r->elt, src, r->ldd * r->elt, (size_t)(row1 - row0) * r->elt, (size_t)(col1 - col0));
      stats.d2h_bytes += (double)(row1 - row0) * (col1 - col0) * r->elt;
      for (int t = ti; t <= te; ++t) r->state[tj * r->mt + t] |= HOST_VALID;
      any = true;
      ti = te + 1;
    }
  }
  if (any) {
    XK_HIP(hipEventRecord(ev_d2h, s_d2h));
    pending_d2h = true;
  }
}

// ---------------------------------------------------------------- staging
// Copy threads for packing/unpacking staging buffers (the calling thread
// takes a share too).  Jobs are ranges of columns.
namespace {
class CopyPool {
 public:
  explicit CopyPool(int n) {
    for (int i = 0; i < n; ++i) threads_.emplace_back([this, i] { loop(i + 1); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  // fn(lo, hi) over [0, count), split into (threads + 1) parts; returns when all are done
  void run(size_t count, const std::function<void(size_t, size_t)>& fn) {
    const size_t parts = threads_.size() + 1;
    {
      std::lock_guard<std::mutex> g(m_);
      fn_ = &fn;
      count_ = count;
      parts_ = parts;
      left_ = threads_.size();
      ++gen_;
    }
    cv_.notify_all();
    slice(0, fn);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return left_ == 0; });
    fn_ = nullptr;
  }

 private:
  void slice(size_t part, const std::function<void(size_t, size_t)>& fn) const {
    const size_t lo = count_ * part / parts_, hi = count_ * (part + 1) / parts_;
    if (lo < hi) fn(lo, hi);
  }
  void loop(size_t part) {
    unsigned long long seen = 0;
    for (;;) {
      const std::function<void(size_t, size_t)>* fn;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        fn = fn_;
      }
      slice(part, *fn);
      {
        std::lock_guard<std::mutex> g(m_);
        if (--left_ == 0) done_.notify_one();
      }
    }
  }
  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t, size_t)>* fn_ = nullptr;
  size_t count_ = 0, parts_ = 1, left_ = 0;
  unsigned long long gen_ = 0;
  bool stop_ = false;
};

CopyPool& copy_pool()
{
  static CopyPool pool([] {
    const char* e = getenv("XKBLAS_COPY_THREADS");
    const int hw = (int)std::thread::hardware_concurrency();
    const int n = e && *e ? atoi(e) : std::min(8, std::max(1, hw));
    return std::max(0, n - 1);
  }());
  return pool;
}

// columns j in [0, height): dst + j dpitch <- src + j spitch, width bytes each
void copy_columns(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t height)
{
  copy_pool().run(height, [&](size_t lo, size_t hi) {
    for (size_t j = lo; j < hi; ++j) memcpy(dst + j * dpitch, src + j * spitch, width);
  });
}
}  // namespace

bool Runtime::host_pinned(const void* p, size_t bytes) const
{
  const uintptr_t a = (uintptr_t)p;
  for (const auto& r : pinned)
    if (a >= r.first && a + bytes <= r.first + r.second) return true;
  return false;
}

// The next staging buffer, with its previous copy finished (and unpacked).
int Runtime::take_stage()
{
  const int b = stage_next;
  stage_next = (b + 1) % NSTAGE;
  if (!stage[b]) XK_HIP(hipHostMalloc((void**)&stage[b], STAGE_BYTES, hipHostMallocDefault));
  finish_stage(b);
  return b;
}

void Runtime::finish_stage(int b)
{
  XK_HIP(hipEventSynchronize(stage_ev[b]));
  Unpack& u = unpack[b];
  if (u.live) {
    copy_columns(u.dst, u.dpitch, stage[b], u.width, u.width, u.height);
    u.live = false;
  }
}

void Runtime::drain_stages()
{
  for (int i = 0; i < NSTAGE; ++i) finish_stage((stage_next + i) % NSTAGE);
}

void Runtime::copy_h2d(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t height)
{
  const size_t span = (height - 1) * spitch + width;
  TraceCopy tc{nullptr, nullptr, 0, width * height, 1, 2};
  if (tracing) trace_copy_begin(s_h2d, tc);
  if (!staging || width > STAGE_BYTES || host_pinned(src, span)) {
    XK_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyHostToDevice, s_h2d));
    if (tracing) trace_copy_end(s_h2d, tc);
    return;
  }
  drain_stages();  // host data written back through the ring is in place first
  const size_t per = STAGE_BYTES / width;
  for (size_t c = 0; c < height; c += per) {
    const size_t cc = std::min(per, height - c);
    const int b = take_stage();
    copy_columns(stage[b], width, src + c * spitch, spitch, width, cc);
    XK_HIP(hipMemcpy2DAsync(dst + c * dpitch, dpitch, stage[b], width, width, cc, hipMemcpyHostToDevice, s_h2d));
    XK_HIP(hipEventRecord(stage_ev[b], s_h2d));
  }
  if (tracing) trace_copy_end(s_h2d, tc);
}

void Runtime::copy_d2h(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t height)
{
  const size_t span = (height - 1) * dpitch + width;
  TraceCopy tc{nullptr, nullptr, 0, width * height, 2, 3};
  if (tracing) trace_copy_begin(s_d2h, tc);
  if (!staging || width > STAGE_BYTES || host_pinned(dst, span)) {
    XK_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost, s_d2h));
    if (tracing) trace_copy_end(s_d2h, tc);
    return;
  }
  const size_t per = STAGE_BYTES / width;
  for (size_t c = 0; c < height; c += per) {
    const size_t cc = std::min(per, height - c);
    const int b = take_stage();
    XK_HIP(hipMemcpy2DAsync(stage[b], width, src + c * spitch, spitch, width, cc, hipMemcpyDeviceToHost, s_d2h));
    XK_HIP(hipEventRecord(stage_ev[b], s_d2h));
    unpack[b] = Unpack{dst + c * dpitch, dpitch, width, cc, true};
  }
  if (tracing) trace_copy_end(s_d2h, tc);
}

// Order the compute stream after the pending copies.
void Runtime::before_compute()
{
  if (pending_h2d) {
    XK_HIP(hipEventRecord(ev_h2d, s_h2d));
    XK_HIP(hipStreamWaitEvent(s_comp, ev_h2d, 0));
    pending_h2d = false;
  }
  if (pending_d2h) {
    // device tiles being copied back must not be overwritten meanwhile
    XK_HIP(hipStreamWaitEvent(s_comp, ev_d2h, 0));
    pending_d2h = false;
  }
}

Exec Runtime::exec() const
{
  Exec x;
  x.s = s_comp;
  x.s_diag = lookahead ? s_diag : nullptr;
  x.mode = mode;
  x.nsplit = nsplit;
  x.ev0 = ev_pot0;
  x.ev1 = ev_pot1;
  return x;
}

void Runtime::init_device()
{
  int ndev = 0;
  XK_HIP(hipGetDeviceCount(&ndev));
  XK_CHECK(ndev > 0, "no HIP device visible");
  XK_HIP(hipSetDevice(0));
  XK_HIP(hipStreamCreateWithFlags(&s_comp, hipStreamNonBlocking));
  XK_HIP(hipStreamCreateWithFlags(&s_diag, hipStreamNonBlocking));
  XK_HIP(hipStreamCreateWithFlags(&s_h2d, hipStreamNonBlocking));
  XK_HIP(hipStreamCreateWithFlags(&s_d2h, hipStreamNonBlocking));
  XK_HIP(hipEventCreateWithFlags(&ev_comp, hipEventDisableTiming));
  XK_HIP(hipEventCreateWithFlags(&ev_h2d, hipEventDisableTiming));
  XK_HIP(hipEventCreateWithFlags(&ev_d2h, hipEventDisableTiming));
  XK_HIP(hipEventCreateWithFlags(&ev_pot0, hipEventDisableTiming));
  XK_HIP(hipEventCreateWithFlags(&ev_pot1, hipEventDisableTiming));
  for (int b = 0; b < NSTAGE; ++b) XK_HIP(hipEventCreateWithFlags(&stage_ev[b], hipEventDisableTiming));
  staging = !(getenv("XKBLAS_STAGING") && atoi(getenv("XKBLAS_STAGING")) == 0);
  info_slots = INFO_SLOTS;
  XK_HIP(hipMalloc(&d_info, sizeof(int) * info_slots));
  XK_HIP(hipHostMalloc(&h_info, sizeof(int) * info_slots, hipHostMallocDefault));
  for (int i = 0; i < info_slots; ++i) h_info[i] = INT_MAX;
  size_t freeb = 0, total = 0;
  XK_HIP(hipMemGetInfo(&freeb, &total));
  // reference default: 98 % of free memory minus a reserve
  // (kaapi/kaapi_impl.c:208, kaapi/kaapi_plugin_hip.c:2365-2366)
  const char* lim = getenv("XKBLAS_CACHE_LIMIT");
  size_t reserve = (size_t)4 << 30;
  size_t dflt = freeb > reserve ? (size_t)((double)(freeb - reserve) * 0.9) : freeb / 2;
  if (lim && *lim) {
    double v = atof(lim);
    cache_limit = v <= 1.0 ? (size_t)(v * (double)freeb) : (size_t)v;
  } else {
    cache_limit = dflt;
  }
  device_ready = true;
  if (const char* td = getenv("XKBLAS_TRACE_DIR"))
    if (*td && !trace_start(td)) fprintf(stderr, "[xkblas_amd] cannot write the trace into '%s'\n", td);
}

void Runtime::sync_all()
{
  if (!device_ready) return;
  XK_HIP(hipStreamSynchronize(s_h2d));
  XK_HIP(hipStreamSynchronize(s_comp));
  XK_HIP(hipStreamSynchronize(s_diag));
  XK_HIP(hipStreamSynchronize(s_d2h));
  drain_stages();
  prof_collect();
}

// ------------------------------------------------------------- profiling
static uint64_t now_ns()
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

void Runtime::prof_begin(const char* cls, double flops)
{
  if (!profiling && !tracing) return;
  ProfEntry e;
  e.cls = cls;
  e.flops = flops;
  e.push_ns = now_ns();
  XK_HIP(hipEventCreate(&e.b));
  XK_HIP(hipEventCreate(&e.e));
  XK_HIP(hipEventRecord(e.b, s_comp));
  prof_open.push_back(e);
}

void Runtime::prof_end()
{
  if (prof_open.empty()) return;
  ProfEntry e = prof_open.back();
  prof_open.pop_back();
  XK_HIP(hipEventRecord(e.e, s_comp));
  prof_pending.push_back(e);
}

void Runtime::prof_collect()
{
  for (auto& e : prof_pending) {
    float ms = 0.0f;
    XK_HIP(hipEventSynchronize(e.e));
    XK_HIP(hipEventElapsedTime(&ms, e.b, e.e));
    if (profiling) {
      ProfStat& st = prof[e.cls];
      st.count += 1;
      st.ms += ms;
      st.flops += e.flops;
    }
    if (tracing) {
      // "Resource,Type,Stream,Push,Start,End,Duration,Name,TaskId" and
      // "Resource,Type,Push,Start,End,Duration,Name,TaskId"
      const double push = (double)(e.push_ns - trace_t0) * 1e-9, t0 = trace_sec(e.b), t1 = trace_sec(e.e);
      const long id = ++trace_id;
      fprintf(f_kern, "1,HIP,1,%.15f,%.15f,%.15f,%.15f,%s,%ld\n", push, t0, t1, t1 - t0, e.cls, id);
      fprintf(f_task, "1,HIP,%.15f,%.15f,%.15f,%.15f,%s,%ld\n", push, t0, t1, t1 - t0, e.cls, id);
    }
    XK_HIP(hipEventDestroy(e.b));
    XK_HIP(hipEventDestroy(e.e));
  }
  prof_pending.clear();
  if (tracing) {
    for (auto& c : trace_copies) {
      XK_HIP(hipEventSynchronize(c.e));
      const double push = (double)(c.push_ns - trace_t0) * 1e-9, t0 = trace_sec(c.b), t1 = trace_sec(c.e);
      // "Resource,Type,Stream,Push,Start,End,Duration,Size,Dir"
      fprintf(f_cpy, "1,HIP,%d,%.15f,%.15f,%.15f,%.15f,%zu,%d\n", c.stream, push, t0, t1, t1 - t0, c.bytes, c.dir);
      XK_HIP(hipEventDestroy(c.b));
      XK_HIP(hipEventDestroy(c.e));
    }
    trace_copies.clear();
    fflush(f_kern);
    fflush(f_task);
    fflush(f_cpy);
  }
}

// seconds from the trace's base event to `ev` (both complete)
double Runtime::trace_sec(hipEvent_t ev)
{
  float ms = 0.0f;
  XK_HIP(hipEventElapsedTime(&ms, trace_base, ev));
  return (double)ms * 1e-3;
}

bool Runtime::trace_start(const char* dir)
{
  if (tracing) trace_stop();
  std::string d = dir && *dir ? dir : ".";
  auto open = [&](const char* name, const char* header) {
    FILE* f = fopen((d + "/" + name).c_str(), "w");
    if (f) fprintf(f, "%s\n", header);
    return f;
  };
  f_kern = open("kern.csv", "Resource,Type,Stream,Push,Start,End,Duration,Name,TaskId");
  f_task = open("task.csv", "Resource,Type,Push,Start,End,Duration,Name,TaskId");
  f_cpy = open("cpy.csv", "Resource,Type,Stream,Push,Start,End,Duration,Size,Dir");
  f_par = open("parallel.csv", "Resource,Type,Start,End,Duration,Name");
  if (!f_kern || !f_task || !f_cpy || !f_par) {
    for (FILE* f : {f_kern, f_task, f_cpy, f_par})
      if (f) fclose(f);
    f_kern = f_task = f_cpy = f_par = nullptr;
    return false;
  }
  sync_all();
  XK_HIP(hipEventCreate(&trace_base));
  XK_HIP(hipEventRecord(trace_base, s_comp));
  XK_HIP(hipEventSynchronize(trace_base));
  trace_t0 = now_ns();
  trace_id = 0;
  tracing = true;
  return true;
}

void Runtime::trace_stop()
{
  if (!tracing) return;
  sync_all();
  const double t1 = (double)(now_ns() - trace_t0) * 1e-9;
  fprintf(f_par, "0,CPU,%.15f,%.15f,%.15f,xkblas\n", 0.0, t1, t1);
  for (FILE* f : {f_kern, f_task, f_cpy, f_par}) fclose(f);
  f_kern = f_task = f_cpy = f_par = nullptr;
  XK_HIP(hipEventDestroy(trace_base));
  trace_base = nullptr;
  tracing = false;
}

void Runtime::trace_copy_begin(hipStream_t s, TraceCopy& c)
{
  c.push_ns = now_ns();
  XK_HIP(hipEventCreate(&c.b));
  XK_HIP(hipEventCreate(&c.e));
  XK_HIP(hipEventRecord(c.b, s));
}

void Runtime::trace_copy_end(hipStream_t s, TraceCopy& c)
{
  XK_HIP(hipEventRecord(c.e, s));
  trace_copies.push_back(c);
}

}  // namespace xk

using namespace xk;

// ======================================================================
// C API
// ======================================================================
std::mutex& xk::api_mutex()
{
  static std::mutex m;
  return m;
}
#define g_mu xk::api_mutex()
typedef std::lock_guard<std::mutex> Lock;

static bool err_return() { return getenv("XKBLAS_ERROR_RETURN") && atoi(getenv("XKBLAS_ERROR_RETURN")) != 0; }

// kaapi_error semantics (blas/common.gen:312-316): print, then abort unless
// XKBLAS_ERROR_RETURN asks for the code.
static int arg_error(const char* fn, const char* msg, int code)
{
  fprintf(stderr, "[xkblas_amd] %s: %s\n", fn, msg);
  fflush(stderr);
  if (!err_return()) abort();
  return code;
}

static void ensure_init()
{
  Runtime& r = rt();
  if (!r.device_ready) {
    r.init_device();
    batch_noalloc_init();
    const char* ts = getenv("XKBLAS_TILE_SIZE");
    if (ts && *ts) r.nb = (size_t)atol(ts);
    const char* dm = getenv("XKBLAS_DEFAULT_MATH");
    if (dm && *dm) r.mode = atoi(dm);
    const char* ns = getenv("XKBLAS_SPLIT_COUNT");
    if (ns && *ns) r.nsplit = atoi(ns);
    const char* la = getenv("XKBLAS_LOOKAHEAD");
    if (la && *la) r.lookahead = atoi(la) != 0;
    const char* pr = getenv("XKBLAS_PROFILE");
    if (pr && *pr) r.profiling = atoi(pr) != 0;
  }
}

extern "C" {

int xkblas_init(void)
{
  Lock l(g_mu);
  ensure_init();
  rt().refcount++;
  return 0;
}

int xkblas_finalize(void)
{
  Lock l(g_mu);
  Runtime& r = rt();
  if (r.refcount > 0) r.refcount--;
  if (r.refcount == 0 && r.device_ready) {
    r.sync_all();
    r.free_retired();
    r.trace_stop();
    if (getenv("XKBLAS_VERBOSE") && atoi(getenv("XKBLAS_VERBOSE")) > 0) {
      fprintf(stderr, "[xkblas_amd] h2d %.3f GB  d2h %.3f GB  evictions %ld  mirrors %zu (%.3f GB)\n",
              r.stats.h2d_bytes / 1e9, r.stats.d2h_bytes / 1e9, r.stats.evictions, r.map.size(),
              r.cache_used / 1e9);
      for (auto& kv : r.prof)
        fprintf(stderr, "[xkblas_amd]   %-14s launches %6ld  time %10.3f ms  %8.2f TFLOP/s\n", kv.first.c_str(),
                kv.second.count, kv.second.ms, kv.second.ms > 0 ? kv.second.flops / kv.second.ms / 1e9 : 0.0);
    }
  }
  return 0;
}

int xkblas_sync(void)
{
  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  r.sync_all();
  r.free_retired();
  return 0;
}

void xkblas_set_param(size_t nb, size_t p)
{
  (void)p;
  Lock l(g_mu);
  rt().nb = nb;
}

size_t xkblas_get_param(void)
{
  Lock l(g_mu);
  return rt().nb;
}

int xkblas_get_ngpus(void) { return 1; }
int xkblas_set_ngpus(int ngpus) { return ngpus == 1 ? 0 : -1; }

int xkblas_get_devicecount(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

int xkblas_get_device_count(int* count)
{
  if (!count) return 22; /* EINVAL */
  *count = xkblas_get_devicecount();
  return 0;
}

void xkblas_set_modemath(xkblas_mode_math_t mode)
{
  Lock l(g_mu);
  rt().mode = (int)mode;
}

const char* xkblas_get_configuration(void)
{
  return "xkblas_amd: gfx950 MFMA (FP64/FP32), HBM tile cache, split-BF16/FP8 CRT modes, 1 GPU";
}

uint64_t xkblas_elapsedns(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

double xkblas_elapsedtime(void) { return (double)xkblas_elapsedns() * 1e-9; }

// Tile-size choice (blas/xkblas.c:2239-2530): the set NB wins; otherwise the
// reference's per-kernel rule for one GPU (its default-mapping branch,
// blas/xkblas.c:2437-2521, with ngpus = 1).  POTRF has no kernel id there and
// takes the default branch.
size_t xkblas_auto_tilesize(void* ctx, xkblas_kernel_t kernel, size_t M, size_t N, size_t K)
{
  (void)ctx;
  (void)K;
  size_t nb = rt().nb;
  if (nb) return nb;
  if (N > M) M = N;
  const size_t a64 = ~(size_t)63;
  switch (kernel) {
    case KERN_SYRK:
    case KERN_HERK: {
      // M / 2 rounded down to a multiple of 2048 when that large; M itself
      // when M / 2 is at most 1024 (the rule repeated with factor 1)
      size_t fact = 2;
      for (;;) {
        nb = M / fact;
        if (nb >= 2048) nb &= ~(size_t)2047;
        if (nb <= 1024 && fact == 2) {
          fact = 1;
          continue;
        }
        break;
      }
      if (nb > 4096) nb = M / (2 * fact);
      if (nb < 512) nb = 512;
      break;
    }
    case KERN_SYR2K:
    case KERN_HER2K: {
      // M / 4, M / 3, M / 2 (64-aligned): the first factor whose block is
      // above 2048, down to factor 2
      size_t fact = 4;
      for (;;) {
        nb = (M / fact + 63) & a64;
        if (nb <= 2048 && fact > 2) {
          --fact;
          continue;
        }
        break;
      }
      if (nb > 4096) nb = M / (2 * fact);
      if (nb < 512) nb = 512;
      break;
    }
    case KERN_SYMM:
    case KERN_HEMM:
      nb = M & ~(size_t)1023;
      if (nb < 1024) nb = 1024;
      if (nb < 4096) nb *= 2;
      break;
    case KERN_TRSM:
    case KERN_TRMM:
      nb = (M / 2 + 63) & a64;
      if (nb < 2048) nb = 2048;
      break;
    case KERN_GEMM:
    case KERN_GEMMT:
      nb = M;
      if (M > 1 && M > 16384 && nb < 2048) nb = 2048;
      if (nb > 4096) nb = M / 2;
      if (nb < 512) nb = 512;
      nb = (nb + 63) & a64;
      break;
    default:  // SWAP, COPYSCALE and this build's POTRF
      nb = M;
      if (nb >= 4096) nb = M / 2;
      if (nb < 1024) nb = 1024;
      nb = (nb + 63) & a64;
      break;
  }
  return nb;
}

// Device properties (blas/xkblas.h:221-256, kaapi/kaapi_offload.c:575-611):
// the reference's HIP plugin knows MI250/MI300A by name and reports any other
// device as XKBLAS_ARCH_UNKNOWN; this build names the MI355X (gfx950)
// XKBLAS_ARCH_MI355X.  No unified memory is used (has_unified 0), as in a
// reference build without KAAPI_UNIFIED.
int xkblas_get_device_properties(xkblas_dev_prop_t* prop)
{
  if (!prop) return 22; /* EINVAL */
  prop->arch = XKBLAS_ARCH_UNKNOWN;
  prop->has_unified = 0;
  prop->device_count = xkblas_get_devicecount();
  if (prop->device_count > 0) {
    hipDeviceProp_t dp;
    if (hipGetDeviceProperties(&dp, 0) == hipSuccess && strncmp(dp.gcnArchName, "gfx950", 6) == 0)
      prop->arch = XKBLAS_ARCH_MI355X;
  }
  return 0;
}

// Unified memory is not used by this build: the reference returns EINVAL
// when the devices cannot use it (kaapi/kaapi_offload.c:631-640).
int xkblas_force_set_unified(void) { return 22; /* EINVAL */ }

// Host registration (blas/xkblas.c:668-774).  The reference pins existing
// host memory through the GPU driver; here copies from pageable memory go
// through the runtime's ring of pinned staging buffers at pinned-copy rates,
// so registration has nothing to do: requests complete at once, with status
// 0, and handles are distinct non-zero numbers.
static std::atomic<uint64_t> g_reg_handle{0};
uint64_t xkblas_register_memory_async(void* ptr, size_t sz)
{
  (void)ptr;
  (void)sz;
  return ++g_reg_handle;
}
uint64_t xkblas_unregister_memory_async(void* ptr, size_t sz)
{
  (void)ptr;
  (void)sz;
  return ++g_reg_handle;
}
int xkblas_register_memory_test(uint64_t handle)
{
  (void)handle;
  return 1;
}
int xkblas_register_memory_wait(uint64_t handle)
{
  (void)handle;
  return 0;
}
int xkblas_register_memory_waitall(void) { return 0; }
int xkblas_register_memory(void* ptr, size_t sz)
{
  return xkblas_register_memory_wait(xkblas_register_memory_async(ptr, sz));
}
int xkblas_unregister_memory(void* ptr, size_t sz)
{
  xkblas_unregister_memory_async(ptr, sz);
  return xkblas_register_memory_waitall();
}

void* xkblas_malloc(size_t sz)
{
  void* p = nullptr;
  if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) return nullptr;
  Lock l(g_mu);
  rt().pinned.emplace_back((uintptr_t)p, sz);
  return p;
}

void xkblas_free(void* ptr, size_t sz)
{
  (void)sz;
  if (!ptr) return;
  {
    Lock l(g_mu);
    auto& v = rt().pinned;
    v.erase(std::remove_if(v.begin(), v.end(), [&](const std::pair<uintptr_t, size_t>& r) {
              return r.first == (uintptr_t)ptr;
            }),
            v.end());
  }
  (void)hipHostFree(ptr);
}

int xkblas_memory_coherent_async(int uplo, int memflag, size_t M, size_t N, void* A, size_t ld, size_t elsize)
{
  (void)ld;
  (void)elsize;
  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  auto it = r.map.find(A);
  if (it == r.map.end()) return 0;  // unknown matrix: nothing to do (reference returns 0)
  Replica* rep = it->second;
  if (uplo != 0 && uplo != Lower && uplo != Upper) {
    fprintf(stderr, "[xkblas_amd] xkblas_memory_coherent_async: invalid argument uplo\n");
    abort();
  }
  r.writeback(rep, 0, 0, (int)std::min<size_t>(M, rep->M), (int)std::min<size_t>(N, rep->N), uplo);
  if (memflag == 1 && uplo != 0) {
    // strict mode: device copies of the other triangle become useless
    for (int tj = 0; tj < rep->nt; ++tj)
      for (int ti = 0; ti < rep->mt; ++ti) {
        if ((uplo == Lower && ti < tj) || (uplo == Upper && ti > tj))
          rep->state[tj * rep->mt + ti] &= (uint8_t)~DEV_VALID;
      }
  }
  return 0;
}

int xkblas_memory_syncall(void)
{
  {
    Lock l(g_mu);
    ensure_init();
    Runtime& r = rt();
    std::vector<Replica*> all;
    for (auto& kv : r.map) all.push_back(kv.second);
    for (Replica* rep : all) r.writeback(rep, 0, 0, rep->M, rep->N, 0);
  }
  return xkblas_sync();
}

int xkblas_memory_invalidate(const void* A)
{
  Lock l(g_mu);
  Runtime& r = rt();
  auto it = r.map.find(A);
  if (it == r.map.end()) return 0;
  r.sync_all();
  r.free_retired();
  r.drop(it->second);
  return 0;
}

int xkblas_memory_invalidate_caches(void)
{
  Lock l(g_mu);
  Runtime& r = rt();
  if (!r.device_ready) return 0;
  r.sync_all();
  r.free_retired();
  // the calling thread's mirrors: another thread's results may still wait
  // for its coherent_async
  const std::thread::id me = std::this_thread::get_id();
  std::vector<Replica*> mine;
  for (auto& kv : r.map)
    if (kv.second->owner == me) mine.push_back(kv.second);
  for (Replica* rep : mine) r.drop(rep);
  return 0;
}

int xkblas_memory_free(void) { return xkblas_memory_invalidate_caches(); }

void xkblas_set_split_count(int nsplit)
{
  Lock l(g_mu);
  rt().nsplit = nsplit;
}

int xkblas_get_split_count(void)
{
  Lock l(g_mu);
  return rt().nsplit;
}

int xkblas_potrf_info(const void* A)
{
  Lock l(g_mu);
  Runtime& r = rt();
  auto it = r.map.find(A);
  if (it == r.map.end() || it->second->info_slot < 0) return -1;
  const int v = r.h_info[it->second->info_slot];
  return v == INT_MAX ? 0 : v;
}

}  // extern "C"

// ---------------------------------------------------------------- compute
namespace {

// Row-panel pipelining of products whose operands are still in host memory
// (XKBLAS_PIPELINE=0 turns it off; xkblas_set_pipeline).
int g_pipeline = -1;
bool pipeline_enabled()
{
  if (g_pipeline < 0) {
    const char* e = getenv("XKBLAS_PIPELINE");
    g_pipeline = (e && *e) ? (atoi(e) != 0) : 1;
  }
  return g_pipeline != 0;
}

template <typename T>
int gemm_async(const char* fn, int transA, int transB, int M, int N, int K, const T* alpha, const T* A, int LDA,
               const T* B, int LDB, const T* beta, T* C, int LDC)
{
  // argument checks and codes of blas/zgemm.c:131-180
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -1);
  if (transB < NoTrans || transB > ConjTrans) return arg_error(fn, "illegal value of transB", -2);
  const int Am = transA == NoTrans ? M : K, An = transA == NoTrans ? K : M;
  const int Bm = transB == NoTrans ? K : N, Bn = transB == NoTrans ? N : K;
  if (M < 0) return arg_error(fn, "illegal value of M", -3);
  if (N < 0) return arg_error(fn, "illegal value of N", -4);
  if (K < 0) return arg_error(fn, "illegal value of K", -5);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, Bm)) return arg_error(fn, "illegal value of LDB", -10);
  if (LDC < std::max(1, M)) return arg_error(fn, "illegal value of LDC", -13);
  if (M == 0 || N == 0 || ((*alpha == T(0) || K == 0) && *beta == T(1))) return 0;  // quick return

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_GEMM, M, N, K);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_ab = K > 0 && *alpha != T(0);
  Replica* rA = use_ab ? r.replica(A, Am, An, LDA, sizeof(T), nb, keep) : nullptr;
  Replica* rB = use_ab ? r.replica(B, Bm, Bn, LDB, sizeof(T), nb, keep) : nullptr;
  Replica* rC = r.replica(C, M, N, LDC, sizeof(T), nb, keep);
  const bool readC = *beta != T(0);  // C is write-only when beta == 0 (ztask_gemm.c:133-134)
  Exec x = r.exec();
  const double flops = 2.0 * M * N * (double)K;
  // Operands still in host memory: the product runs by row panels of C (one
  // tile row each), and panel p waits only for B and its own rows of op(A)
  // and C, so the copies of panel p+1 run under the product of panel p and
  // each panel's rows go back (xkblas_memory_coherent_async) as soon as they
  // are final.  Every element sees the same operands and inner order as in
  // one launch (in the split modes the row scalings span the whole inner
  // dimension and B's residue planes are computed once), so the result is
  // the same bit for bit.
  const int panels = (M + nb - 1) / nb;
  const bool rowsA = transA == NoTrans;
  const bool pipe = use_ab && rA != rC && rB != rC && panels > 1 && pipeline_enabled() &&
                    (!r.all_valid(rA, 0, 0, Am, An) || (readC && !r.all_valid(rC, 0, 0, M, N)));
  if (pipe) {
    r.fetch(rB, 0, 0, Bm, Bn);
    const bool split = sizeof(T) == 8 && split_enabled<T>(x);
    if (split) split_cache_register((const double*)rB->dev, rB->ldd, N, K, transB != NoTrans);
    std::vector<Replica::PanelEvent> pev;
    r.prof_begin(sizeof(T) == 8 ? "gemm_f64" : "gemm_f32", flops);
    for (int p = 0; p < panels; ++p) {
      const int r0 = p * nb, rows = std::min(nb, M - r0);
      if (rowsA) r.fetch(rA, r0, 0, rows, K);
      else r.fetch(rA, 0, r0, K, rows);
      if (readC) r.fetch(rC, r0, 0, rows, N);
      r.before_compute();
      const T* Ap = (const T*)rA->dev + (rowsA ? (long)r0 : (long)r0 * rA->ldd);
      dev_gemm<T>(x, transA, transB, rows, N, K, *alpha, Ap, rA->ldd, (const T*)rB->dev, rB->ldd, *beta,
                  (T*)rC->dev + r0, rC->ldd);
      hipEvent_t ev;
      XK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      XK_HIP(hipEventRecord(ev, x.s));
      pev.push_back(Replica::PanelEvent{r0 + rows, ev});
    }
    r.prof_end();
    if (split) split_cache_clear();
    r.mark_written(rC, 0, 0, M, N, 0);
    rC->panels = pev;
    return 0;
  }
  if (rA) r.fetch(rA, 0, 0, Am, An);
  if (rB) r.fetch(rB, 0, 0, Bm, Bn);
  if (readC) r.fetch(rC, 0, 0, M, N);
  r.before_compute();
  r.prof_begin(sizeof(T) == 8 ? "gemm_f64" : "gemm_f32", flops);
  dev_gemm<T>(x, transA, transB, M, N, use_ab ? K : 0, use_ab ? *alpha : T(0), rA ? (const T*)rA->dev : (const T*)rC->dev,
              rA ? rA->ldd : rC->ldd, rB ? (const T*)rB->dev : (const T*)rC->dev, rB ? rB->ldd : rC->ldd, *beta,
              (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, M, N, 0);
  return 0;
}

template <typename T>
int syrk_async(const char* fn, int uplo, int trans, int N, int K, const T* alpha, const T* A, int LDA,
               const T* beta, T* C, int LDC)
{
  // checks of blas/zsyrk.c:111-160 (same order and codes)
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -1);
  if (trans != NoTrans && trans != Trans && trans != ConjTrans) return arg_error(fn, "illegal value of trans", -2);
  const int Am = trans == NoTrans ? N : K, An = trans == NoTrans ? K : N;
  if (N < 0) return arg_error(fn, "illegal value of N", -3);
  if (K < 0) return arg_error(fn, "illegal value of K", -4);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -7);
  if (LDC < std::max(1, N)) return arg_error(fn, "illegal value of LDC", -10);
  if (N == 0 || ((*alpha == T(0) || K == 0) && *beta == T(1))) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_SYRK, N, N, K);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_a = K > 0 && *alpha != T(0);
  Replica* rA = use_a ? r.replica(A, Am, An, LDA, sizeof(T), nb, keep) : nullptr;
  Replica* rC = r.replica(C, N, N, LDC, sizeof(T), nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, An);
  r.fetch(rC, 0, 0, N, N);  // the other triangle must stay intact on the device
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin(sizeof(T) == 8 ? "syrk_f64" : "syrk_f32", (double)K * N * (N + 1.0));
  dev_syrk<T>(x, uplo, trans, N, use_a ? K : 0, use_a ? *alpha : T(0), rA ? (const T*)rA->dev : (const T*)rC->dev,
              rA ? rA->ldd : rC->ldd, *beta, (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, N, N, uplo);
  return 0;
}

template <typename T>
int trsm_async(const char* fn, int side, int uplo, int transA, int diag, int N, int NRHS, const T* alpha, const T* A,
               int LDA, T* B, int LDB)
{
  // checks of blas/ztrsm.c:117-167; N = rows of B, NRHS = columns
  const int NA = side == Left ? N : NRHS;
  if (side != Left && side != Right) return arg_error(fn, "illegal value of side", -1);
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -2);
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -3);
  if (diag != Unit && diag != NonUnit) return arg_error(fn, "illegal value of diag", -4);
  if (N < 0) return arg_error(fn, "illegal value of N", -5);
  if (NRHS < 0) return arg_error(fn, "illegal value of NRHS", -6);
  if (LDA < std::max(1, NA)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, N)) return arg_error(fn, "illegal value of LDB", -10);
  if (std::min(N, NRHS) == 0) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_TRSM, N, NRHS, NA);
  std::vector<Replica*>& keep = r.begin_call();
  Replica* rA = r.replica(A, NA, NA, LDA, sizeof(T), nb, keep);
  Replica* rB = r.replica(B, N, NRHS, LDB, sizeof(T), nb, keep);
  r.fetch(rA, 0, 0, NA, NA);
  r.fetch(rB, 0, 0, N, NRHS);
  r.before_compute();
  Exec x = r.exec();
  const double flops = side == Left ? (double)N * N * NRHS : (double)N * NRHS * NRHS;
  r.prof_begin(sizeof(T) == 8 ? "trsm_f64" : "trsm_f32", flops);
  // inner diagonal block order: the tile size, capped for the inverse GEMMs
  const int dnb = std::min(nb, 512);
  dev_trsm<T>(x, side, uplo, transA, diag, N, NRHS, *alpha, (const T*)rA->dev, rA->ldd, (T*)rB->dev, rB->ldd, dnb);
  r.prof_end();
  r.mark_written(rB, 0, 0, N, NRHS, 0);
  return 0;
}

template <typename T>
int potrf_async(const char* fn, int uplo, int N, T* A, int LDA)
{
  // checks of blas/zpotrf.c:57-68
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -2);
  if (N < 0) return arg_error(fn, "illegal value of N", -5);
  if (LDA < std::max(1, N)) return arg_error(fn, "illegal value of LDA", -8);
  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  std::vector<Replica*>& keep = r.begin_call();
  if (N == 0) return 0;
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_POTRF, N, N, N);
  Replica* rA = r.replica(A, N, N, LDA, sizeof(T), nb, keep);
  if (rA->info_slot < 0) rA->info_slot = r.take_info_slot();
  r.fetch(rA, 0, 0, N, N);
  r.before_compute();
  Exec x = r.exec();
  int* dinfo = r.d_info + rA->info_slot;
  fill_int(x.s, dinfo, 1, INT_MAX);
  r.prof_begin(sizeof(T) == 8 ? "potrf_f64" : "potrf_f32", (double)N * N * N / 3.0 + (double)N * N / 2.0 + N / 6.0);
  dev_potrf<T>(x, uplo, N, (T*)rA->dev, rA->ldd, nb, dinfo);
  r.prof_end();
  XK_HIP(hipMemcpyAsync(r.h_info + rA->info_slot, dinfo, sizeof(int), hipMemcpyDeviceToHost, x.s));
  r.mark_written(rA, 0, 0, N, N, uplo);
  return 0;
}

template <typename T>
int gemmt_async(const char* fn, int uplo, int transA, int transB, int N, int K, const T* alpha, const T* A, int LDA,
                const T* B, int LDB, const T* beta, T* C, int LDC)
{
  // checks and codes of blas/zgemmt.c:123-165 (uplo 0 = all of C, as there;
  // any other value than Upper/Lower/0 is refused here)
  if (uplo != 0 && uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -1);
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -1);
  if (transB < NoTrans || transB > ConjTrans) return arg_error(fn, "illegal value of transB", -2);
  const int Am = transA == NoTrans ? N : K, An = transA == NoTrans ? K : N;
  const int Bm = transB == NoTrans ? K : N, Bn = transB == NoTrans ? N : K;
  if (N < 0) return arg_error(fn, "illegal value of N", -4);
  if (K < 0) return arg_error(fn, "illegal value of K", -5);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, Bm)) return arg_error(fn, "illegal value of LDB", -10);
  if (LDC < std::max(1, N)) return arg_error(fn, "illegal value of LDC", -13);
  if (N == 0 || ((*alpha == T(0) || K == 0) && *beta == T(1))) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_GEMMT, N, N, K);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_ab = K > 0 && *alpha != T(0);
  Replica* rA = use_ab ? r.replica(A, Am, An, LDA, sizeof(T), nb, keep) : nullptr;
  Replica* rB = use_ab ? r.replica(B, Bm, Bn, LDB, sizeof(T), nb, keep) : nullptr;
  Replica* rC = r.replica(C, N, N, LDC, sizeof(T), nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, An);
  if (rB) r.fetch(rB, 0, 0, Bm, Bn);
  r.fetch(rC, 0, 0, N, N);  // the other triangle stays intact on the device
  r.before_compute();
  Exec x = r.exec();
  const bool tri = uplo == Upper || uplo == Lower;
  r.prof_begin(sizeof(T) == 8 ? "gemmt_f64" : "gemmt_f32", tri ? (double)K * N * (N + 1.0) : 2.0 * N * N * (double)K);
  dev_gemmt<T>(x, uplo, transA, transB, N, use_ab ? K : 0, use_ab ? *alpha : T(0),
               rA ? (const T*)rA->dev : (const T*)rC->dev, rA ? rA->ldd : rC->ldd,
               rB ? (const T*)rB->dev : (const T*)rC->dev, rB ? rB->ldd : rC->ldd, *beta, (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, N, N, tri ? uplo : 0);
  return 0;
}

template <typename T>
int syr2k_async(const char* fn, int uplo, int trans, int N, int K, const T* alpha, const T* A, int LDA, const T* B,
                int LDB, const T* beta, T* C, int LDC)
{
  // checks and codes of blas/zsyr2k.c:119-156
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -1);
  if (trans != NoTrans && trans != Trans) return arg_error(fn, "illegal value of trans", -2);
  const int Am = trans == NoTrans ? N : K, An = trans == NoTrans ? K : N;
  if (N < 0) return arg_error(fn, "illegal value of N", -3);
  if (K < 0) return arg_error(fn, "illegal value of K", -4);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -7);
  if (LDB < std::max(1, Am)) return arg_error(fn, "illegal value of LDB", -9);
  if (LDC < std::max(1, N)) return arg_error(fn, "illegal value of LDC", -12);
  if (N == 0 || ((*alpha == T(0) || K == 0) && *beta == T(1))) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_SYR2K, N, N, K);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_ab = K > 0 && *alpha != T(0);
  Replica* rA = use_ab ? r.replica(A, Am, An, LDA, sizeof(T), nb, keep) : nullptr;
  Replica* rB = use_ab ? r.replica(B, Am, An, LDB, sizeof(T), nb, keep) : nullptr;
  Replica* rC = r.replica(C, N, N, LDC, sizeof(T), nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, An);
  if (rB) r.fetch(rB, 0, 0, Am, An);
  r.fetch(rC, 0, 0, N, N);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin(sizeof(T) == 8 ? "syr2k_f64" : "syr2k_f32", 2.0 * K * N * (N + 1.0));
  dev_syr2k<T>(x, uplo, trans, N, use_ab ? K : 0, use_ab ? *alpha : T(0), rA ? (const T*)rA->dev : (const T*)rC->dev,
               rA ? rA->ldd : rC->ldd, rB ? (const T*)rB->dev : (const T*)rC->dev, rB ? rB->ldd : rC->ldd, *beta,
               (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, N, N, uplo);
  return 0;
}

template <typename T>
int symm_async(const char* fn, int side, int uplo, int M, int N, const T* alpha, const T* A, int LDA, const T* B,
               int LDB, const T* beta, T* C, int LDC)
{
  // checks and codes of blas/zsymm.c:115-148
  if (side != Left && side != Right) return arg_error(fn, "illegal value of side", -1);
  if (uplo != Lower && uplo != Upper) return arg_error(fn, "illegal value of uplo", -2);
  const int Am = side == Left ? M : N;
  if (M < 0) return arg_error(fn, "illegal value of M", -3);
  if (N < 0) return arg_error(fn, "illegal value of N", -4);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -7);
  if (LDB < std::max(1, M)) return arg_error(fn, "illegal value of LDB", -9);
  if (LDC < std::max(1, M)) return arg_error(fn, "illegal value of LDC", -12);
  if (M == 0 || N == 0 || (*alpha == T(0) && *beta == T(1))) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_SYMM, M, N, Am);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_ab = *alpha != T(0);
  Replica* rA = use_ab ? r.replica(A, Am, Am, LDA, sizeof(T), nb, keep) : nullptr;
  Replica* rB = use_ab ? r.replica(B, M, N, LDB, sizeof(T), nb, keep) : nullptr;
  Replica* rC = r.replica(C, M, N, LDC, sizeof(T), nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, Am);
  if (rB) r.fetch(rB, 0, 0, M, N);
  if (*beta != T(0)) r.fetch(rC, 0, 0, M, N);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin(sizeof(T) == 8 ? "symm_f64" : "symm_f32", 2.0 * M * N * (double)Am);
  dev_symm<T>(x, side, uplo, M, N, use_ab ? *alpha : T(0), rA ? (const T*)rA->dev : (const T*)rC->dev,
              rA ? rA->ldd : rC->ldd, rB ? (const T*)rB->dev : (const T*)rC->dev, rB ? rB->ldd : rC->ldd, *beta,
              (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, M, N, 0);
  return 0;
}

template <typename T>
int trmm_async(const char* fn, int side, int uplo, int transA, int diag, int N, int NRHS, const T* alpha, const T* A,
               int LDA, T* B, int LDB)
{
  // checks of blas/ztrmm.c:116-153; N = rows of B, NRHS = columns
  const int NA = side == Left ? N : NRHS;
  if (side != Left && side != Right) return arg_error(fn, "illegal value of side", -1);
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -2);
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -3);
  if (diag != Unit && diag != NonUnit) return arg_error(fn, "illegal value of diag", -4);
  if (N < 0) return arg_error(fn, "illegal value of N", -5);
  if (NRHS < 0) return arg_error(fn, "illegal value of NRHS", -6);
  if (LDA < std::max(1, NA)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, N)) return arg_error(fn, "illegal value of LDB", -10);
  if (std::min(N, NRHS) == 0) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_TRMM, N, NRHS, NA);
  std::vector<Replica*>& keep = r.begin_call();
  Replica* rA = r.replica(A, NA, NA, LDA, sizeof(T), nb, keep);
  Replica* rB = r.replica(B, N, NRHS, LDB, sizeof(T), nb, keep);
  r.fetch(rA, 0, 0, NA, NA);
  r.fetch(rB, 0, 0, N, NRHS);
  r.before_compute();
  Exec x = r.exec();
  const double flops = side == Left ? (double)N * N * NRHS : (double)N * NRHS * NRHS;
  r.prof_begin(sizeof(T) == 8 ? "trmm_f64" : "trmm_f32", flops);
  dev_trmm<T>(x, side, uplo, transA, diag, N, NRHS, *alpha, (const T*)rA->dev, rA->ldd, (T*)rB->dev, rB->ldd);
  r.prof_end();
  r.mark_written(rB, 0, 0, N, NRHS, 0);
  return 0;
}

// Complex GEMM: T is the real type; alpha/beta point at (re, im) pairs.
template <typename T>
int cgemm_async(const char* fn, int transA, int transB, int M, int N, int K, const T* alpha, const T* A, int LDA,
                const T* B, int LDB, const T* beta, T* C, int LDC)
{
  // argument checks and codes of blas/zgemm.c:131-180
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -1);
  if (transB < NoTrans || transB > ConjTrans) return arg_error(fn, "illegal value of transB", -2);
  const int Am = transA == NoTrans ? M : K, An = transA == NoTrans ? K : M;
  const int Bm = transB == NoTrans ? K : N, Bn = transB == NoTrans ? N : K;
  if (M < 0) return arg_error(fn, "illegal value of M", -3);
  if (N < 0) return arg_error(fn, "illegal value of N", -4);
  if (K < 0) return arg_error(fn, "illegal value of K", -5);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, Bm)) return arg_error(fn, "illegal value of LDB", -10);
  if (LDC < std::max(1, M)) return arg_error(fn, "illegal value of LDC", -13);
  const bool a0 = alpha[0] == T(0) && alpha[1] == T(0), b1 = beta[0] == T(1) && beta[1] == T(0);
  if (M == 0 || N == 0 || ((a0 || K == 0) && b1)) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_GEMM, M, N, K);
  const int elt = 2 * sizeof(T);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_ab = K > 0 && !a0;
  Replica* rA = use_ab ? r.replica(A, Am, An, LDA, elt, nb, keep) : nullptr;
  Replica* rB = use_ab ? r.replica(B, Bm, Bn, LDB, elt, nb, keep) : nullptr;
  Replica* rC = r.replica(C, M, N, LDC, elt, nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, An);
  if (rB) r.fetch(rB, 0, 0, Bm, Bn);
  if (beta[0] != T(0) || beta[1] != T(0)) r.fetch(rC, 0, 0, M, N);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin(sizeof(T) == 8 ? "gemm_c128" : "gemm_c64", 8.0 * M * N * (double)K);
  const T zero[2] = {T(0), T(0)};
  dev_cgemm<T>(x, transA, transB, M, N, use_ab ? K : 0, use_ab ? alpha : zero, rA ? (const T*)rA->dev : nullptr,
               rA ? rA->ldd : 1, rB ? (const T*)rB->dev : nullptr, rB ? rB->ldd : 1, beta, (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, M, N, 0);
  return 0;
}

template <typename T>
int cherk_async(const char* fn, int uplo, int trans, int N, int K, const T* alpha, const T* A, int LDA, const T* beta,
                T* C, int LDC)
{
  // checks and codes of blas/zherk.c:104-137 (alpha, beta real)
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -1);
  if (trans != NoTrans && trans != ConjTrans) return arg_error(fn, "illegal value of trans", -2);
  const int Am = trans == NoTrans ? N : K, An = trans == NoTrans ? K : N;
  if (N < 0) return arg_error(fn, "illegal value of N", -3);
  if (K < 0) return arg_error(fn, "illegal value of K", -4);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -7);
  if (LDC < std::max(1, N)) return arg_error(fn, "illegal value of LDC", -10);
  if (N == 0 || ((*alpha == T(0) || K == 0) && *beta == T(1))) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_HERK, N, N, K);
  const int elt = 2 * sizeof(T);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_a = K > 0 && *alpha != T(0);
  Replica* rA = use_a ? r.replica(A, Am, An, LDA, elt, nb, keep) : nullptr;
  Replica* rC = r.replica(C, N, N, LDC, elt, nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, An);
  r.fetch(rC, 0, 0, N, N);  // the other triangle stays intact on the device
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin(sizeof(T) == 8 ? "herk_c128" : "herk_c64", 4.0 * K * N * (N + 1.0));
  dev_cherk<T>(x, uplo, trans, N, use_a ? K : 0, use_a ? *alpha : T(0), rA ? (const T*)rA->dev : nullptr,
               rA ? rA->ldd : 1, *beta, (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, N, N, uplo);
  return 0;
}

template <typename T>
int chemm_async(const char* fn, int herm, int side, int uplo, int M, int N, const T* alpha, const T* A, int LDA,
                const T* B, int LDB, const T* beta, T* C, int LDC)
{
  // checks and codes of blas/zhemm.c:113-147 (zsymm.c:115-148 the same)
  if (side != Left && side != Right) return arg_error(fn, "illegal value of side", -1);
  if (uplo != Lower && uplo != Upper) return arg_error(fn, "illegal value of uplo", -2);
  const int Am = side == Left ? M : N;
  if (M < 0) return arg_error(fn, "illegal value of M", -3);
  if (N < 0) return arg_error(fn, "illegal value of N", -4);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -7);
  if (LDB < std::max(1, M)) return arg_error(fn, "illegal value of LDB", -9);
  if (LDC < std::max(1, M)) return arg_error(fn, "illegal value of LDC", -12);
  const bool a0 = alpha[0] == T(0) && alpha[1] == T(0), b1 = beta[0] == T(1) && beta[1] == T(0);
  if (M == 0 || N == 0 || (a0 && b1)) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_HEMM, M, N, Am);
  const int elt = 2 * sizeof(T);
  std::vector<Replica*>& keep = r.begin_call();
  Replica* rA = !a0 ? r.replica(A, Am, Am, LDA, elt, nb, keep) : nullptr;
  Replica* rB = !a0 ? r.replica(B, M, N, LDB, elt, nb, keep) : nullptr;
  Replica* rC = r.replica(C, M, N, LDC, elt, nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, Am);
  if (rB) r.fetch(rB, 0, 0, M, N);
  if (beta[0] != T(0) || beta[1] != T(0)) r.fetch(rC, 0, 0, M, N);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin(herm ? (sizeof(T) == 8 ? "hemm_c128" : "hemm_c64") : (sizeof(T) == 8 ? "symm_c128" : "symm_c64"),
               8.0 * M * N * (double)Am);
  dev_chemm<T>(x, herm, side, uplo, M, N, alpha, rA ? (const T*)rA->dev : nullptr, rA ? rA->ldd : 1,
               rB ? (const T*)rB->dev : nullptr, rB ? rB->ldd : 1, beta, (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, M, N, 0);
  return 0;
}

// complex SYRK / SYR2K / HER2K / GEMMT / TRMM: the real templates' checks
// with complex scalars (HER2K's beta real, passed here as (beta, 0))
template <typename T>
int csyr2k_async(const char* fn, int herm, int two, int uplo, int trans, int N, int K, const T* alpha, const T* A,
                 int LDA, const T* B, int LDB, const T* beta, T* C, int LDC)
{
  // blas/zsyrk.c:104-133, zsyr2k.c:119-156, zher2k.c:115-150
  const int t = herm ? ConjTrans : Trans;
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -1);
  if (trans != NoTrans && trans != t) return arg_error(fn, "illegal value of trans", -2);
  const int Am = trans == NoTrans ? N : K, An = trans == NoTrans ? K : N;
  if (N < 0) return arg_error(fn, "illegal value of N", -3);
  if (K < 0) return arg_error(fn, "illegal value of K", -4);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -7);
  if (two && LDB < std::max(1, Am)) return arg_error(fn, "illegal value of LDB", -9);
  if (LDC < std::max(1, N)) return arg_error(fn, "illegal value of LDC", two ? -12 : -10);
  const bool a0 = alpha[0] == T(0) && alpha[1] == T(0), b1 = beta[0] == T(1) && beta[1] == T(0);
  if (N == 0 || ((a0 || K == 0) && b1)) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, two ? KERN_SYR2K : KERN_SYRK, N, N, K);
  const int elt = 2 * sizeof(T);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_ab = K > 0 && !a0;
  Replica* rA = use_ab ? r.replica(A, Am, An, LDA, elt, nb, keep) : nullptr;
  Replica* rB = use_ab && two ? r.replica(B, Am, An, LDB, elt, nb, keep) : nullptr;
  Replica* rC = r.replica(C, N, N, LDC, elt, nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, An);
  if (rB) r.fetch(rB, 0, 0, Am, An);
  r.fetch(rC, 0, 0, N, N);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin(two ? (herm ? "her2k_c" : "syr2k_c") : "syrk_c", (two ? 8.0 : 4.0) * K * N * (N + 1.0));
  const T zero[2] = {T(0), T(0)};
  const T* dA = rA ? (const T*)rA->dev : nullptr;
  const T* dB = two ? (rB ? (const T*)rB->dev : nullptr) : dA;
  const long ldb = two ? (rB ? rB->ldd : 1) : (rA ? rA->ldd : 1);
  if (two) {
    dev_csyr2k<T>(x, herm, uplo, trans, N, use_ab ? K : 0, use_ab ? alpha : zero, dA, rA ? rA->ldd : 1, dB, ldb, beta,
                  (T*)rC->dev, rC->ldd);
  } else {
    // SYRK: one product op(A) op(A)^T
    const int ta = trans == NoTrans ? NoTrans : Trans, tb = trans == NoTrans ? Trans : NoTrans;
    dev_cgemmt<T>(x, uplo, ta, tb, N, use_ab ? K : 0, use_ab ? alpha : zero, dA, rA ? rA->ldd : 1, dB, ldb, beta,
                  (T*)rC->dev, rC->ldd);
  }
  r.prof_end();
  r.mark_written(rC, 0, 0, N, N, uplo);
  return 0;
}

template <typename T>
int cgemmt_async(const char* fn, int uplo, int transA, int transB, int N, int K, const T* alpha, const T* A, int LDA,
                 const T* B, int LDB, const T* beta, T* C, int LDC)
{
  // blas/zgemmt.c:123-165
  if (uplo != 0 && uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -1);
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -1);
  if (transB < NoTrans || transB > ConjTrans) return arg_error(fn, "illegal value of transB", -2);
  const int Am = transA == NoTrans ? N : K, An = transA == NoTrans ? K : N;
  const int Bm = transB == NoTrans ? K : N, Bn = transB == NoTrans ? N : K;
  if (N < 0) return arg_error(fn, "illegal value of N", -4);
  if (K < 0) return arg_error(fn, "illegal value of K", -5);
  if (LDA < std::max(1, Am)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, Bm)) return arg_error(fn, "illegal value of LDB", -10);
  if (LDC < std::max(1, N)) return arg_error(fn, "illegal value of LDC", -13);
  const bool a0 = alpha[0] == T(0) && alpha[1] == T(0), b1 = beta[0] == T(1) && beta[1] == T(0);
  if (N == 0 || ((a0 || K == 0) && b1)) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_GEMMT, N, N, K);
  const int elt = 2 * sizeof(T);
  std::vector<Replica*>& keep = r.begin_call();
  const bool use_ab = K > 0 && !a0;
  Replica* rA = use_ab ? r.replica(A, Am, An, LDA, elt, nb, keep) : nullptr;
  Replica* rB = use_ab ? r.replica(B, Bm, Bn, LDB, elt, nb, keep) : nullptr;
  Replica* rC = r.replica(C, N, N, LDC, elt, nb, keep);
  if (rA) r.fetch(rA, 0, 0, Am, An);
  if (rB) r.fetch(rB, 0, 0, Bm, Bn);
  r.fetch(rC, 0, 0, N, N);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin("gemmt_c", 8.0 * N * N * (double)K);
  const T zero[2] = {T(0), T(0)};
  dev_cgemmt<T>(x, uplo, transA, transB, N, use_ab ? K : 0, use_ab ? alpha : zero, rA ? (const T*)rA->dev : nullptr,
                rA ? rA->ldd : 1, rB ? (const T*)rB->dev : nullptr, rB ? rB->ldd : 1, beta, (T*)rC->dev, rC->ldd);
  r.prof_end();
  r.mark_written(rC, 0, 0, N, N, uplo == Upper || uplo == Lower ? uplo : 0);
  return 0;
}

template <typename T>
int ctrmm_async(const char* fn, int side, int uplo, int transA, int diag, int N, int NRHS, const T* alpha, const T* A,
                int LDA, T* B, int LDB)
{
  // blas/ztrmm.c:116-153
  const int NA = side == Left ? N : NRHS;
  if (side != Left && side != Right) return arg_error(fn, "illegal value of side", -1);
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -2);
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -3);
  if (diag != Unit && diag != NonUnit) return arg_error(fn, "illegal value of diag", -4);
  if (N < 0) return arg_error(fn, "illegal value of N", -5);
  if (NRHS < 0) return arg_error(fn, "illegal value of NRHS", -6);
  if (LDA < std::max(1, NA)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, N)) return arg_error(fn, "illegal value of LDB", -10);
  if (std::min(N, NRHS) == 0) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_TRMM, N, NRHS, NA);
  const int elt = 2 * sizeof(T);
  std::vector<Replica*>& keep = r.begin_call();
  Replica* rA = r.replica(A, NA, NA, LDA, elt, nb, keep);
  Replica* rB = r.replica(B, N, NRHS, LDB, elt, nb, keep);
  r.fetch(rA, 0, 0, NA, NA);
  r.fetch(rB, 0, 0, N, NRHS);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin("trmm_c", 4.0 * (side == Left ? (double)N * N * NRHS : (double)N * NRHS * NRHS));
  dev_ctrmm<T>(x, side, uplo, transA, diag, N, NRHS, alpha, (const T*)rA->dev, rA->ldd, (T*)rB->dev, rB->ldd);
  r.prof_end();
  r.mark_written(rB, 0, 0, N, NRHS, 0);
  return 0;
}

template <typename T>
int ctrsm_async(const char* fn, int side, int uplo, int transA, int diag, int N, int NRHS, const T* alpha, const T* A,
                int LDA, T* B, int LDB)
{
  // checks of blas/ztrsm.c:117-167
  const int NA = side == Left ? N : NRHS;
  if (side != Left && side != Right) return arg_error(fn, "illegal value of side", -1);
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -2);
  if (transA < NoTrans || transA > ConjTrans) return arg_error(fn, "illegal value of transA", -3);
  if (diag != Unit && diag != NonUnit) return arg_error(fn, "illegal value of diag", -4);
  if (N < 0) return arg_error(fn, "illegal value of N", -5);
  if (NRHS < 0) return arg_error(fn, "illegal value of NRHS", -6);
  if (LDA < std::max(1, NA)) return arg_error(fn, "illegal value of LDA", -8);
  if (LDB < std::max(1, N)) return arg_error(fn, "illegal value of LDB", -10);
  if (std::min(N, NRHS) == 0) return 0;

  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_TRSM, N, NRHS, NA);
  const int elt = 2 * sizeof(T);
  std::vector<Replica*>& keep = r.begin_call();
  Replica* rA = r.replica(A, NA, NA, LDA, elt, nb, keep);
  Replica* rB = r.replica(B, N, NRHS, LDB, elt, nb, keep);
  r.fetch(rA, 0, 0, NA, NA);
  r.fetch(rB, 0, 0, N, NRHS);
  r.before_compute();
  Exec x = r.exec();
  r.prof_begin("trsm_c", 4.0 * (side == Left ? (double)N * N * NRHS : (double)N * NRHS * NRHS));
  dev_ctrsm<T>(x, side, uplo, transA, diag, N, NRHS, alpha, (const T*)rA->dev, rA->ldd, (T*)rB->dev, rB->ldd,
               std::min(nb, 256));
  r.prof_end();
  r.mark_written(rB, 0, 0, N, NRHS, 0);
  return 0;
}

template <typename T>
int cpotrf_async(const char* fn, int uplo, int N, T* A, int LDA)
{
  // checks of blas/zpotrf.c:57-68
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -2);
  if (N < 0) return arg_error(fn, "illegal value of N", -5);
  if (LDA < std::max(1, N)) return arg_error(fn, "illegal value of LDA", -8);
  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  std::vector<Replica*>& keep = r.begin_call();
  if (N == 0) return 0;
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_POTRF, N, N, N);
  Replica* rA = r.replica(A, N, N, LDA, 2 * sizeof(T), nb, keep);
  if (rA->info_slot < 0) rA->info_slot = r.take_info_slot();
  r.fetch(rA, 0, 0, N, N);
  r.before_compute();
  Exec x = r.exec();
  int* dinfo = r.d_info + rA->info_slot;
  fill_int(x.s, dinfo, 1, INT_MAX);
  r.prof_begin("potrf_c", 4.0 * ((double)N * N * N / 3.0 + (double)N * N / 2.0));
  dev_cpotrf<T>(x, uplo, N, (T*)rA->dev, rA->ldd, std::min(nb, 512), dinfo);
  r.prof_end();
  XK_HIP(hipMemcpyAsync(r.h_info + rA->info_slot, dinfo, sizeof(int), hipMemcpyDeviceToHost, x.s));
  r.mark_written(rA, 0, 0, N, N, uplo);
  return 0;
}

}  // namespace

extern "C" {

int xkblas_zpotrf_async(int uplo, int N, Complex64_t* A, int LDA)
{
  return cpotrf_async<double>("xkblas_zpotrf_async", uplo, N, (double*)A, LDA);
}

int xkblas_cpotrf_async(int uplo, int N, Complex32_t* A, int LDA)
{
  return cpotrf_async<float>("xkblas_cpotrf_async", uplo, N, (float*)A, LDA);
}

int xkblas_ztrsm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex64_t* alpha,
                       const Complex64_t* A, int LDA, Complex64_t* B, int LDB)
{
  return ctrsm_async<double>("xkblas_ztrsm_async", side, uplo, transA, diag, N, NRHS, (const double*)alpha,
                             (const double*)A, LDA, (double*)B, LDB);
}

int xkblas_ctrsm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const Complex32_t* alpha,
                       const Complex32_t* A, int LDA, Complex32_t* B, int LDB)
{
  return ctrsm_async<float>("xkblas_ctrsm_async", side, uplo, transA, diag, N, NRHS, (const float*)alpha,
                            (const float*)A, LDA, (float*)B, LDB);
}

#define XK_CPLX_OPS(P, C, R)                                                                                      \
  int xkblas_##P##syrk_async(int uplo, int trans, int N, int K, const C* alpha, const C* A, int LDA, const C* beta, \
                             C* Cm, int LDC)                                                                      \
  {                                                                                                               \
    return csyr2k_async<R>("xkblas_" #P "syrk_async", 0, 0, uplo, trans, N, K, (const R*)alpha, (const R*)A, LDA, \
                           nullptr, 1, (const R*)beta, (R*)Cm, LDC);                                              \
  }                                                                                                               \
  int xkblas_##P##syr2k_async(int uplo, int trans, int N, int K, const C* alpha, const C* A, int LDA, const C* B, \
                              int LDB, const C* beta, C* Cm, int LDC)                                             \
  {                                                                                                               \
    return csyr2k_async<R>("xkblas_" #P "syr2k_async", 0, 1, uplo, trans, N, K, (const R*)alpha, (const R*)A, LDA,\
                           (const R*)B, LDB, (const R*)beta, (R*)Cm, LDC);                                        \
  }                                                                                                               \
  int xkblas_##P##her2k_async(int uplo, int trans, int N, int K, const C* alpha, const C* A, int LDA, const C* B, \
                              int LDB, const R* beta, C* Cm, int LDC)                                             \
  {                                                                                                               \
    const R b2[2] = {*beta, R(0)};                                                                                \
    return csyr2k_async<R>("xkblas_" #P "her2k_async", 1, 1, uplo, trans, N, K, (const R*)alpha, (const R*)A, LDA,\
                           (const R*)B, LDB, b2, (R*)Cm, LDC);                                                    \
  }                                                                                                               \
  int xkblas_##P##gemmt_async(int uplo, int transA, int transB, int N, int K, const C* alpha, const C* A, int LDA, \
                              const C* B, int LDB, const C* beta, C* Cm, int LDC)                                 \
  {                                                                                                               \
    return cgemmt_async<R>("xkblas_" #P "gemmt_async", uplo, transA, transB, N, K, (const R*)alpha, (const R*)A,   \
                           LDA, (const R*)B, LDB, (const R*)beta, (R*)Cm, LDC);                                   \
  }                                                                                                               \
  int xkblas_##P##symm_async(int side, int uplo, int M, int N, const C* alpha, const C* A, int LDA, const C* B,    \
                             int LDB, const C* beta, C* Cm, int LDC)                                              \
  {                                                                                                               \
    return chemm_async<R>("xkblas_" #P "symm_async", 0, side, uplo, M, N, (const R*)alpha, (const R*)A, LDA,      \
                          (const R*)B, LDB, (const R*)beta, (R*)Cm, LDC);                                         \
  }                                                                                                               \
  int xkblas_##P##trmm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const C* alpha,           \
                             const C* A, int LDA, C* B, int LDB)                                                  \
  {                                                                                                               \
    return ctrmm_async<R>("xkblas_" #P "trmm_async", side, uplo, transA, diag, N, NRHS, (const R*)alpha,          \
                          (const R*)A, LDA, (R*)B, LDB);                                                          \
  }
XK_CPLX_OPS(z, Complex64_t, double)
XK_CPLX_OPS(c, Complex32_t, float)
#undef XK_CPLX_OPS

int xkblas_zherk_async(int uplo, int trans, int N, int K, const CFloat64_t* alpha, const Complex64_t* A, int LDA,
                       const CFloat64_t* beta, Complex64_t* C, int LDC)
{
  return cherk_async<double>("xkblas_zherk_async", uplo, trans, N, K, alpha, (const double*)A, LDA, beta,
                             (double*)C, LDC);
}

int xkblas_cherk_async(int uplo, int trans, int N, int K, const float* alpha, const Complex32_t* A, int LDA,
                       const float* beta, Complex32_t* C, int LDC)
{
  return cherk_async<float>("xkblas_cherk_async", uplo, trans, N, K, alpha, (const float*)A, LDA, beta, (float*)C,
                            LDC);
}

int xkblas_zhemm_async(int side, int uplo, int M, int N, const Complex64_t* alpha, const Complex64_t* A, int LDA,
                       const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC)
{
  return chemm_async<double>("xkblas_zhemm_async", 1, side, uplo, M, N, (const double*)alpha, (const double*)A, LDA,
                             (const double*)B, LDB, (const double*)beta, (double*)C, LDC);
}

int xkblas_chemm_async(int side, int uplo, int M, int N, const Complex32_t* alpha, const Complex32_t* A, int LDA,
                       const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC)
{
  return chemm_async<float>("xkblas_chemm_async", 1, side, uplo, M, N, (const float*)alpha, (const float*)A, LDA,
                            (const float*)B, LDB, (const float*)beta, (float*)C, LDC);
}

int xkblas_zgemm_async(int transA, int transB, int M, int N, int K, const Complex64_t* alpha, const Complex64_t* A,
                       int LDA, const Complex64_t* B, int LDB, const Complex64_t* beta, Complex64_t* C, int LDC)
{
  return cgemm_async<double>("xkblas_zgemm_async", transA, transB, M, N, K, (const double*)alpha,
                             (const double*)A, LDA, (const double*)B, LDB, (const double*)beta, (double*)C, LDC);
}

int xkblas_cgemm_async(int transA, int transB, int M, int N, int K, const Complex32_t* alpha, const Complex32_t* A,
                       int LDA, const Complex32_t* B, int LDB, const Complex32_t* beta, Complex32_t* C, int LDC)
{
  return cgemm_async<float>("xkblas_cgemm_async", transA, transB, M, N, K, (const float*)alpha, (const float*)A,
                            LDA, (const float*)B, LDB, (const float*)beta, (float*)C, LDC);
}

#define XK_NEXT_OPS(P, T)                                                                                         \
  int xkblas_##P##gemmt_async(int uplo, int transA, int transB, int N, int K, const T* alpha, const T* A, int LDA, \
                              const T* B, int LDB, const T* beta, T* C, int LDC)                                  \
  {                                                                                                               \
    return gemmt_async<T>("xkblas_" #P "gemmt_async", uplo, transA, transB, N, K, alpha, A, LDA, B, LDB, beta, C, \
                          LDC);                                                                                   \
  }                                                                                                               \
  int xkblas_##P##syr2k_async(int uplo, int trans, int N, int K, const T* alpha, const T* A, int LDA, const T* B,  \
                              int LDB, const T* beta, T* C, int LDC)                                              \
  {                                                                                                               \
    return syr2k_async<T>("xkblas_" #P "syr2k_async", uplo, trans, N, K, alpha, A, LDA, B, LDB, beta, C, LDC);    \
  }                                                                                                               \
  int xkblas_##P##symm_async(int side, int uplo, int M, int N, const T* alpha, const T* A, int LDA, const T* B,    \
                             int LDB, const T* beta, T* C, int LDC)                                               \
  {                                                                                                               \
    return symm_async<T>("xkblas_" #P "symm_async", side, uplo, M, N, alpha, A, LDA, B, LDB, beta, C, LDC);       \
  }                                                                                                               \
  int xkblas_##P##trmm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const T* alpha,           \
                             const T* A, int LDA, T* B, int LDB)                                                  \
  {                                                                                                               \
    return trmm_async<T>("xkblas_" #P "trmm_async", side, uplo, transA, diag, N, NRHS, alpha, A, LDA, B, LDB);    \
  }
XK_NEXT_OPS(d, double)
XK_NEXT_OPS(s, float)
#undef XK_NEXT_OPS

int xkblas_dgemm_async(int transA, int transB, int M, int N, int K, const double* alpha, const double* A, int LDA,
                       const double* B, int LDB, const double* beta, double* C, int LDC)
{
  return gemm_async<double>("xkblas_dgemm_async", transA, transB, M, N, K, alpha, A, LDA, B, LDB, beta, C, LDC);
}

int xkblas_sgemm_async(int transA, int transB, int M, int N, int K, const float* alpha, const float* A, int LDA,
                       const float* B, int LDB, const float* beta, float* C, int LDC)
{
  return gemm_async<float>("xkblas_sgemm_async", transA, transB, M, N, K, alpha, A, LDA, B, LDB, beta, C, LDC);
}

int xkblas_dsyrk_async(int uplo, int trans, int N, int K, const double* alpha, const double* A, int LDA,
                       const double* beta, double* C, int LDC)
{
  return syrk_async<double>("xkblas_dsyrk_async", uplo, trans, N, K, alpha, A, LDA, beta, C, LDC);
}

int xkblas_ssyrk_async(int uplo, int trans, int N, int K, const float* alpha, const float* A, int LDA,
                       const float* beta, float* C, int LDC)
{
  return syrk_async<float>("xkblas_ssyrk_async", uplo, trans, N, K, alpha, A, LDA, beta, C, LDC);
}

int xkblas_dtrsm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const double* alpha,
                       const double* A, int LDA, double* B, int LDB)
{
  return trsm_async<double>("xkblas_dtrsm_async", side, uplo, transA, diag, N, NRHS, alpha, A, LDA, B, LDB);
}

int xkblas_strsm_async(int side, int uplo, int transA, int diag, int N, int NRHS, const float* alpha,
                       const float* A, int LDA, float* B, int LDB)
{
  return trsm_async<float>("xkblas_strsm_async", side, uplo, transA, diag, N, NRHS, alpha, A, LDA, B, LDB);
}

int xkblas_dpotrf_async(int uplo, int N, double* A, int LDA)
{
  return potrf_async<double>("xkblas_dpotrf_async", uplo, N, A, LDA);
}

int xkblas_spotrf_async(int uplo, int N, float* A, int LDA)
{
  return potrf_async<float>("xkblas_spotrf_async", uplo, N, A, LDA);
}

// ---- mixed-precision SPD solve with FP64 iterative refinement (no
// counterpart in the reference; LAPACK dsposv's argument order and codes)
static int g_ir_moduli = 7, g_ir_itermax = 30;

void xkblas_set_posv_ir(int moduli, int itermax)
{
  Lock l(g_mu);
  g_ir_moduli = moduli > 0 ? std::min(std::max(moduli, 4), 13) : 7;
  g_ir_itermax = itermax >= 0 ? itermax : 30;
}

int xkblas_dposv_ir(int uplo, int N, int NRHS, const double* A, int LDA, const double* B, int LDB, double* X,
                    int LDX, int* iter)
{
  const char* fn = "xkblas_dposv_ir";
  if (uplo != Upper && uplo != Lower) return arg_error(fn, "illegal value of uplo", -1);
  if (N < 0) return arg_error(fn, "illegal value of N", -2);
  if (NRHS < 0) return arg_error(fn, "illegal value of NRHS", -3);
  if (LDA < std::max(1, N)) return arg_error(fn, "illegal value of LDA", -5);
  if (LDB < std::max(1, N)) return arg_error(fn, "illegal value of LDB", -7);
  if (LDX < std::max(1, N)) return arg_error(fn, "illegal value of LDX", -9);
  if (!iter) return arg_error(fn, "iter is NULL", -10);
  if ((const void*)X == (const void*)B) return arg_error(fn, "X must not be B", -8);
  *iter = 0;
  if (N == 0 || NRHS == 0) return 0;
  Lock l(g_mu);
  ensure_init();
  Runtime& r = rt();
  const int nb = (int)xkblas_auto_tilesize(nullptr, KERN_POTRF, N, N, N);
  std::vector<Replica*>& keep = r.begin_call();
  Replica* rA = r.replica(A, N, N, LDA, sizeof(double), nb, keep);
  Replica* rB = r.replica(B, N, NRHS, LDB, sizeof(double), nb, keep);
  Replica* rX = r.replica(X, N, NRHS, LDX, sizeof(double), nb, keep);
  r.fetch(rA, 0, 0, N, N);
  r.fetch(rB, 0, 0, N, NRHS);
  r.before_compute();
  Exec xf = r.exec();              // the accurate factor: the current mode
  Exec x = xf;                     // residuals, updates and solves: FP64
  x.mode = XKBLAS_DEFAULT_MATH;
  Exec xl = xf;                    // the reduced factor: split-BF16, fewer moduli
  xl.mode = XKBLAS_SPLIT_BF16_MATH;
  xl.nsplit = g_ir_moduli;
  r.prof_begin("posv_ir_f64", (double)N * N * N / 3.0 + 2.0 * N * N * (double)NRHS);
  const int info = dev_posv_ir(x, xl, xf, uplo, N, NRHS, (const double*)rA->dev, rA->ldd, (const double*)rB->dev,
                               rB->ldd, (double*)rX->dev, rX->ldd, nb, g_ir_itermax, iter);
  r.prof_end();
  r.mark_written(rX, 0, 0, N, NRHS, 0);
  return info;
}

// ---- trace in the reference trace reader's CSV schema (bin/ukilli.cpp:848-1113)
int xkblas_trace_start(const char* dir)
{
  Lock l(g_mu);
  ensure_init();
  return rt().trace_start(dir) ? 0 : -1;
}

int xkblas_trace_stop(void)
{
  Lock l(g_mu);
  rt().trace_stop();
  return 0;
}

// ---- profiling extension
void xkblas_profile_enable(int on)
{
  Lock l(g_mu);
  rt().profiling = on != 0;
}

void xkblas_profile_reset(void)
{
  Lock l(g_mu);
  rt().sync_all();
  rt().prof.clear();
}

int xkblas_profile_get(const char* cls, long* launches, double* total_ms, double* flops)
{
  Lock l(g_mu);
  Runtime& r = rt();
  auto it = r.prof.find(cls);
  if (it == r.prof.end()) {
    if (launches) *launches = 0;
    if (total_ms) *total_ms = 0;
    if (flops) *flops = 0;
    return -1;
  }
  if (launches) *launches = it->second.count;
  if (total_ms) *total_ms = it->second.ms;
  if (flops) *flops = it->second.flops;
  return 0;
}

void xkblas_set_pipeline(int on)
{
  Lock l(g_mu);
  g_pipeline = on != 0;
}

void xkblas_set_lookahead(int on)
{
  Lock l(g_mu);
  rt().lookahead = on != 0;
}

long xkblas_cache_bytes(void)
{
  Lock l(g_mu);
  return (long)rt().cache_used;
}

long xkblas_cache_evictions(void)
{
  Lock l(g_mu);
  return rt().stats.evictions;
}

void xkblas_cache_traffic(double* h2d_bytes, double* d2h_bytes)
{
  Lock l(g_mu);
  if (h2d_bytes) *h2d_bytes = rt().stats.h2d_bytes;
  if (d2h_bytes) *d2h_bytes = rt().stats.d2h_bytes;
}

void xkblas_set_cache_limit(size_t bytes)
{
  Lock l(g_mu);
  ensure_init();
  rt().cache_limit = bytes;
}

}  // extern "C"
