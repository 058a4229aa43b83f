// runtime.h — state of the single-GPU runtime behind xkblas.h.
#pragma once
#include "blas3.h"
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <thread>
#include <vector>

namespace xk {

enum : uint8_t { DEV_VALID = 1, HOST_VALID = 2 };

// Device mirror of one host matrix (the reference's matrix descriptor plus
// the DSM replica state of its tiles, blas/common.gen:167-187,
// kaapi/kaapi_memory.h:151-163).
struct Replica {
  const void* host = nullptr;
  int M = 0, N = 0;
  long ld = 0;
  int elt = 8;
  int nb = 512, mt = 0, nt = 0;
  char* dev = nullptr;
  long ldd = 0;
  size_t bytes = 0;
  std::vector<uint8_t> state;  // per tile (column-major tile order)
  uint64_t lru = 0;
  int info_slot = -1;
  // the thread whose call created the mirror: xkblas_memory_invalidate_caches
  // drops the mirrors of the calling thread only (each thread's descriptor
  // list in the reference, blas/xkblas.c:1855-1908)
  std::thread::id owner;
  // Rows [.., end) of the mirror are final once `ev` completes on the
  // compute stream (a product pipelined by row panels): the copy back of a
  // row panel waits only for its own event.  Empty otherwise.
  struct PanelEvent {
    int end;
    hipEvent_t ev;
  };
  std::vector<PanelEvent> panels;
};

struct ProfEntry {
  const char* cls;
  double flops;
  hipEvent_t b, e;
  uint64_t push_ns;  // host time of the enqueue (trace)
};

// One host<->device copy seen by the trace: events around it on its stream.
struct TraceCopy {
  hipEvent_t b, e;
  uint64_t push_ns;
  size_t bytes;
  int dir;     // 1 H2D, 2 D2H (the reference's trace codes)
  int stream;  // 1 compute, 2 H2D, 3 D2H
};
struct ProfStat {
  long count = 0;
  double ms = 0.0, flops = 0.0;
};

struct Runtime {
  static constexpr int NSCRATCH = 34;
  static constexpr int INFO_SLOTS = 1024;

  bool device_ready = false;
  int refcount = 0;
  size_t nb = 0;  // 0 = heuristic per call
  int mode = 0;
  int nsplit = 0;
  bool lookahead = true;
  bool profiling = false;

  hipStream_t s_comp = nullptr, s_diag = nullptr, s_h2d = nullptr, s_d2h = nullptr;
  hipEvent_t ev_comp = nullptr, ev_h2d = nullptr, ev_d2h = nullptr;
  hipEvent_t ev_pot0 = nullptr, ev_pot1 = nullptr;  // POTRF's lookahead ordering
  bool pending_h2d = false, pending_d2h = false;

  std::unordered_map<const void*, Replica*> map;
  size_t cache_used = 0, cache_limit = 0;
  uint64_t clock = 0;

  void* scratch_ptr[NSCRATCH] = {};
  size_t scratch_size[NSCRATCH] = {};

  // POTRF info words, one slot per factored mirror (freed with it; the
  // arrays grow when every slot is taken)
  int* d_info = nullptr;
  int* h_info = nullptr;
  int info_slots = 0, next_info = 0;
  std::vector<int> free_info;

  // The replicas of the current call (never evicted or dropped while it
  // runs): an API entry clears it, replica() appends.
  std::vector<Replica*> call_keep;
  // Mirrors that an operand of the running call still uses but that lost
  // their pointer key (an aliased operand of other dimensions): freed once
  // every stream has drained (sync_all).
  std::vector<Replica*> retired;
  // Scratch buffers outgrown during a call: kernels already enqueued may
  // still use them, so they are freed at the next point where every stream
  // has drained instead of synchronising the device in the middle of a call
  // (their bytes stay counted against the cache limit until then).
  std::vector<std::pair<void*, size_t>> retired_scratch;

  // Pageable host memory goes through a ring of pinned staging buffers: a
  // host-to-device chunk is packed by the copy threads, then sent; a
  // device-to-host chunk is unpacked when its buffer is next taken or at
  // sync.  Memory from xkblas_malloc (pinned) is copied directly.
  static constexpr int NSTAGE = 4;
  static constexpr size_t STAGE_BYTES = (size_t)64 << 20;
  bool staging = true;
  char* stage[NSTAGE] = {};
  hipEvent_t stage_ev[NSTAGE] = {};
  struct Unpack {
    char* dst = nullptr;
    size_t dpitch = 0, width = 0, height = 0;
    bool live = false;
  } unpack[NSTAGE];
  int stage_next = 0;
  std::vector<std::pair<uintptr_t, size_t>> pinned;  // xkblas_malloc ranges

  struct {
    double h2d_bytes = 0, d2h_bytes = 0;
    long evictions = 0;
  } stats;

  std::vector<ProfEntry> prof_open, prof_pending;
  std::map<std::string, ProfStat> prof;

  // Trace in the CSV schema of the reference's trace reader (bin/ukilli.cpp):
  // kern.csv / task.csv (one row per call's device work), cpy.csv (copies),
  // parallel.csv (the traced region).  Times in seconds from trace start.
  bool tracing = false;
  hipEvent_t trace_base = nullptr;
  uint64_t trace_t0 = 0;
  long trace_id = 0;
  FILE *f_kern = nullptr, *f_task = nullptr, *f_cpy = nullptr, *f_par = nullptr;
  std::vector<TraceCopy> trace_copies;

  void init_device();
  void sync_all();
  Replica* replica(const void* host, int M, int N, long ld, int elt, int nb, std::vector<Replica*>& keep);
  void fetch(Replica* r, int r0, int c0, int m, int n);
  void mark_written(Replica* r, int r0, int c0, int m, int n, int uplo);
  void writeback(Replica* r, int r0, int c0, int m, int n, int uplo);
  void writeback_rows(Replica* r, int r0, int c0, int m, int n, int uplo);
  bool all_valid(const Replica* r, int r0, int c0, int m, int n) const;
  void clear_panels(Replica* r);
  std::vector<Replica*>& begin_call();
  bool evict_for(size_t need, const std::vector<Replica*>& keep, bool strict = true);
  void* device_alloc(size_t bytes, const std::vector<Replica*>& keep);
  void drop(Replica* r);
  void free_retired();
  int take_info_slot();
  void before_compute();
  bool host_pinned(const void* p, size_t bytes) const;
  void copy_h2d(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t height);
  void copy_d2h(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t height);
  int take_stage();
  void finish_stage(int b);
  void drain_stages();
  Exec exec() const;
  void prof_begin(const char* cls, double flops);
  void prof_end();
  void prof_collect();
  bool trace_start(const char* dir);
  void trace_stop();
  double trace_sec(hipEvent_t ev);
  void trace_copy_begin(hipStream_t s, TraceCopy& c);
  void trace_copy_end(hipStream_t s, TraceCopy& c);
};

Runtime& rt();

// serialises every C API entry point
std::mutex& api_mutex();

}  // namespace xk
