// Host runtime of libcuspatial_amd: per-thread error text, the caching device
// pool, and output allocation through a caller resource.
//
// The pool plays the role RMM's pool resource plays for the reference
// (cpp/benchmarks/fixture/rmm_pool_raii.hpp): hot-path temporaries are carved
// from cached device blocks instead of hipMalloc/hipFree per call. Blocks are
// binned by size; a cached block is reused for a request of at least half its
// size.
//
// The pool is stream-ordered (as RMM's stream-ordered resources are). A block is
// freed when its owner leaves scope, usually before the kernels queued on the
// call's stream have finished with it. Each free block carries the stream it was
// last used on and, once the call that used it has ended, that call's completion
// event:
//   * a request on the same stream takes it with no wait (the stream orders the
//     old work before the new);
//   * a request on another stream takes it behind hipStreamWaitEvent on the
//     completion event (no host synchronisation);
//   * a block whose call is still running (no event yet) goes only to its own
//     stream.
// The completion event is recorded once per outermost StreamScope (one per C
// entry point call) and shared by every block the call allocated or freed.
// Blocks freed outside any scope (qpip_free of an output) keep the completion
// event of the call that produced them: the caller's own uses must be finished,
// as with hipFree.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "qpip_internal.hpp"

namespace qpip {

namespace {
thread_local std::string g_last_error;

// A recorded completion event, shared by the blocks it orders; returned to its
// device's event cache when the last block drops it.
struct CompletionEvent;
using EventPtr = std::shared_ptr<CompletionEvent>;

struct BlockInfo {
  size_t size         = 0;
  hipStream_t stream  = nullptr;  // stream of the last use
  uint64_t scope      = 0;        // scope still running on `stream` (0: none)
  EventPtr event;                 // completion of the last use (null: not yet recorded)
  bool idle           = false;    // no work can touch the block any more
  bool bound          = false;    // `stream` is known (allocated or freed inside a scope)
};

struct DevicePool {
  std::mutex mu;
  std::unordered_map<void*, BlockInfo> live;
  std::unordered_map<void*, BlockInfo> free_info;
  std::multimap<size_t, void*> free_by_size;
  std::mutex ev_mu;  // guards spare_events (events are dropped while mu is held)
  std::vector<hipEvent_t> spare_events;
  size_t live_bytes   = 0;
  size_t cached_bytes = 0;
};

std::mutex g_pools_mu;
std::vector<DevicePool*> g_pools;  // never freed: events return to them at exit
size_t g_limit = 0;                // 0: no cap

struct CompletionEvent {
  hipEvent_t ev;
  DevicePool* pool;
  ~CompletionEvent()
  {
    std::lock_guard<std::mutex> g(pool->ev_mu);
    pool->spare_events.push_back(ev);
  }
};

DevicePool& pool_for_device(int dev)
{
  std::lock_guard<std::mutex> g(g_pools_mu);
  if (static_cast<size_t>(dev) >= g_pools.size()) g_pools.resize(dev + 1, nullptr);
  if (!g_pools[dev]) g_pools[dev] = new DevicePool();  // lives until process exit
  return *g_pools[dev];
}

int current_device()
{
  int dev = 0;
  QPIP_HIP(hipGetDevice(&dev));
  return dev;
}

// For pool_free, which runs in destructors and must not throw: after a device
// error the runtime may report it here too; the block is then looked up on
// device 0's pool (a single-device process) and the error is left for the
// caller's next checked call.
int current_device_nothrow()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  return dev;
}

// Open scopes of this thread, innermost last. A scope on the stream of the
// enclosing one joins it (one event for the outermost).
struct ScopeState {
  hipStream_t stream;
  uint64_t id;
  int depth;
  bool completed;
  DevicePool* pool;
  std::vector<void*> touched;  // blocks allocated or freed in the scope
};
thread_local std::vector<ScopeState> t_scopes;
std::atomic<uint64_t> g_scope_ids{1};

ScopeState* current_scope() { return t_scopes.empty() ? nullptr : &t_scopes.back(); }

size_t round_size(size_t bytes)
{
  if (bytes <= (size_t{1} << 20)) return (bytes + 511) & ~size_t{511};
  size_t const gran = size_t{2} << 20;
  return (bytes + gran - 1) / gran * gran;
}

bool stream_capturing(hipStream_t s)
{
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return true;  // unknown: act as if capturing (no cross-stream handoff)
  }
  return st != hipStreamCaptureStatusNone;
}

// Takes a cached block of [size, 2 size] usable on `stream` (caller holds p.mu).
bool take(DevicePool& p, size_t size, hipStream_t stream, void*& blk, BlockInfo& info)
{
  bool may_wait = true;  // decided on first need: a capturing stream waits on nothing
  bool checked  = false;
  for (auto it = p.free_by_size.lower_bound(size);
       it != p.free_by_size.end() && it->first <= 2 * size; ++it) {
    auto fi = p.free_info.find(it->second);
    BlockInfo const& b = fi->second;
    // a block used outside any scope (unknown stream, no event) waits for
    // release_cached, which synchronises the device
    bool ok = b.idle || (b.bound && b.stream == stream);
    if (!ok && b.event) {
      if (!checked) {
        may_wait = !stream_capturing(stream);
        checked  = true;
      }
      if (may_wait) {
        QPIP_HIP(hipStreamWaitEvent(stream, b.event->ev, 0));
        ok = true;
      }
    }
    if (!ok) continue;
    blk  = it->second;
    info = std::move(fi->second);
    p.free_info.erase(fi);
    p.free_by_size.erase(it);
    return true;
  }
  return false;
}

// Frees every cached block (caller holds p.mu). Kernels of other threads may
// still use some of them: the device is synchronised first.
void release_cached(DevicePool& p)
{
  if (p.free_by_size.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : p.free_by_size) {
    (void)hipFree(kv.second);
    p.cached_bytes -= kv.first;
  }
  p.free_by_size.clear();
  p.free_info.clear();
}
}  // namespace

void set_last_error(std::string const& msg) { g_last_error = msg; }
std::string const& last_error() { return g_last_error; }

void* pool_alloc(size_t bytes)
{
  if (bytes == 0) return nullptr;
  int const dev      = current_device();
  DevicePool& p      = pool_for_device(dev);
  ScopeState* sc     = current_scope();
  hipStream_t stream = sc ? sc->stream : nullptr;
  size_t const size  = round_size(bytes);
  std::lock_guard<std::mutex> g(p.mu);
  if (g_limit && p.live_bytes + size > g_limit)
    throw oom_error("out of memory: library memory limit of " + std::to_string(g_limit) +
                    " bytes reached (requested " + std::to_string(size) + ")");
  void* blk = nullptr;
  BlockInfo info;
  if (take(p, size, stream, blk, info)) {
    p.cached_bytes -= info.size;
  } else {
    hipError_t err = hipMalloc(&blk, size);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      release_cached(p);
      err = hipMalloc(&blk, size);
    }
    if (err != hipSuccess) {
      (void)hipGetLastError();
      throw oom_error("out of memory: hipMalloc of " + std::to_string(size) + " bytes failed");
    }
    info.size = size;
  }
  info.stream = stream;
  info.scope  = sc ? sc->id : 0;
  info.event.reset();
  info.idle  = false;
  info.bound = sc != nullptr;
  p.live_bytes += info.size;
  p.live[blk] = std::move(info);
  if (sc) sc->touched.push_back(blk);
  return blk;
}

void pool_free(void* blk)
{
  if (!blk) return;
  int const dev  = current_device_nothrow();
  DevicePool& p  = pool_for_device(dev);
  ScopeState* sc = current_scope();
  std::lock_guard<std::mutex> g(p.mu);
  auto it = p.live.find(blk);
  if (it == p.live.end()) return;  // not ours
  BlockInfo info = std::move(it->second);
  p.live.erase(it);
  p.live_bytes -= info.size;
  if (sc && sc->completed) {
    info.idle = true;
    info.event.reset();
  } else if (sc) {
    // last used by this call: ordered on its stream, event at the scope's end
    info.stream = sc->stream;
    info.scope  = sc->id;
    info.bound  = true;
    info.event.reset();
    sc->touched.push_back(blk);
  }
  // else: a caller's qpip_free; the block keeps its producing call's event
  p.cached_bytes += info.size;
  p.free_by_size.emplace(info.size, blk);
  p.free_info[blk] = std::move(info);
}

StreamScope::StreamScope(hipStream_t stream)
{
  ScopeState* cur = current_scope();
  if (cur && cur->stream == stream && !cur->completed) {
    ++cur->depth;
    return;
  }
  t_scopes.push_back(ScopeState{stream, g_scope_ids.fetch_add(1), 1, false,
                                &pool_for_device(current_device_nothrow()), {}});
}

StreamScope::~StreamScope()
{
  ScopeState& sc = t_scopes.back();
  if (--sc.depth > 0) return;
  DevicePool& p = *sc.pool;
  EventPtr ev;
  if (!sc.touched.empty() && !stream_capturing(sc.stream)) {
    hipEvent_t e = nullptr;
    {
      std::lock_guard<std::mutex> g(p.ev_mu);
      if (!p.spare_events.empty()) {
        e = p.spare_events.back();
        p.spare_events.pop_back();
      }
    }
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
    if (e && hipEventRecord(e, sc.stream) == hipSuccess) {
      ev = EventPtr(new CompletionEvent{e, &p});
    } else if (e) {
      (void)hipGetLastError();
      std::lock_guard<std::mutex> g(p.ev_mu);
      p.spare_events.push_back(e);
    }
  }
  if (ev) {
    std::lock_guard<std::mutex> g(p.mu);
    for (void* b : sc.touched) {
      auto lv = p.live.find(b);
      if (lv != p.live.end() && lv->second.scope == sc.id) {
        lv->second.event = ev;
        lv->second.scope = 0;
        continue;
      }
      auto fv = p.free_info.find(b);
      if (fv != p.free_info.end() && fv->second.scope == sc.id) {
        fv->second.event = ev;
        fv->second.scope = 0;
      }
    }
  }
  // without an event (none recorded, e.g. under capture) the blocks stay bound
  // to their stream
  t_scopes.pop_back();
}

CompletedScope::CompletedScope()
{
  t_scopes.push_back(ScopeState{nullptr, g_scope_ids.fetch_add(1), 1, true,
                                &pool_for_device(current_device_nothrow()), {}});
}

CompletedScope::~CompletedScope() { t_scopes.pop_back(); }

void pool_set_limit(size_t bytes) { g_limit = bytes; }

void pool_trim()
{
  DevicePool& p = pool_for_device(current_device());
  std::lock_guard<std::mutex> g(p.mu);
  release_cached(p);
}

void pool_stats(size_t* live, size_t* cached)
{
  DevicePool& p = pool_for_device(current_device());
  std::lock_guard<std::mutex> g(p.mu);
  *live   = p.live_bytes;
  *cached = p.cached_bytes;
}

size_t pool_bytes_in_use()
{
  DevicePool& p = pool_for_device(current_device());
  std::lock_guard<std::mutex> g(p.mu);
  return p.live_bytes;
}

// --- kernel timing ------------------------------------------------------------
namespace {
struct ProfRecord {
  std::string name;
  hipEvent_t start;
  hipEvent_t stop;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRecord*> g_prof_pending;
std::vector<hipEvent_t> g_prof_free_events;
std::map<std::string, std::pair<uint64_t, double>> g_prof_totals;  // name -> (count, ms)

hipEvent_t prof_event()
{
  if (!g_prof_free_events.empty()) {
    hipEvent_t e = g_prof_free_events.back();
    g_prof_free_events.pop_back();
    return e;
  }
  // timing only: no system-scope fence (no L2 writeback and invalidate) at each
  // record, which would slow the timed work itself; the report reads the events
  // after a stream synchronisation
  hipEvent_t e;
  QPIP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  return e;
}

void prof_drain_locked()
{
  for (ProfRecord* r : g_prof_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(r->stop) == hipSuccess &&
        hipEventElapsedTime(&ms, r->start, r->stop) == hipSuccess) {
      auto& t = g_prof_totals[r->name];
      t.first += 1;
      t.second += ms;
    }
    g_prof_free_events.push_back(r->start);
    g_prof_free_events.push_back(r->stop);
    delete r;
  }
  g_prof_pending.clear();
}
}  // namespace

// Launches enqueued while this thread captures a launch sequence are not timed
// (an event pair per launch would become part of the captured graph).
thread_local bool t_capturing = false;

bool profiling_enabled() { return g_prof_on && !t_capturing; }

void set_capturing(bool on) { t_capturing = on; }

void profile_begin(char const* name, hipStream_t stream, void** token)
{
  std::lock_guard<std::mutex> g(g_prof_mu);
  auto* r  = new ProfRecord{name, prof_event(), prof_event()};
  QPIP_HIP(hipEventRecord(r->start, stream));
  *token = r;
}

void profile_end(void* token, hipStream_t stream)
{
  std::lock_guard<std::mutex> g(g_prof_mu);
  auto* r = static_cast<ProfRecord*>(token);
  (void)hipEventRecord(r->stop, stream);
  g_prof_pending.push_back(r);
  if (g_prof_pending.size() > 4096) prof_drain_locked();
}

void* OutputAllocator::allocate(size_t bytes)
{
  if (bytes == 0) return nullptr;
  owned_.reserve(owned_.size() + 1);  // the push below cannot throw after the allocation
  void* p = nullptr;
  if (!mr_) {
    p = pool_alloc(bytes);
  } else {
    p = mr_->allocate(mr_->ctx, bytes, stream_);
    if (!p) throw oom_error("out of memory: caller resource returned NULL for " +
                            std::to_string(bytes) + " bytes");
  }
  owned_.push_back({p, bytes});
  return p;
}

void OutputAllocator::deallocate(void* p, size_t bytes)
{
  if (!p) return;
  for (size_t i = 0; i < owned_.size(); ++i)
    if (owned_[i].p == p) {
      owned_.erase(owned_.begin() + static_cast<std::ptrdiff_t>(i));
      break;
    }
  if (!mr_)
    pool_free(p);
  else
    mr_->deallocate(mr_->ctx, p, bytes, stream_);
}

OutputAllocator::~OutputAllocator()
{
  for (auto it = owned_.rbegin(); it != owned_.rend(); ++it) {
    if (!mr_)
      pool_free(it->p);
    else
      mr_->deallocate(mr_->ctx, it->p, it->bytes, stream_);
  }
}

// --- tuning parameters ---------------------------------------------------------
namespace {
struct TuningParam {
  char const* name;
  std::atomic<int64_t> value;
  int64_t lo, hi;
};
TuningParam g_tuning[kNumTuning] = {
  {"pip_min_pairs_per_wave", {2}, 1, int64_t{1} << 30},
  {"pip_pairs_shift", {24}, 0, 40},
  {"pip_waves_per_block", {2}, 1, 16},  // PIP test workgroup: 2, 4 or 8 waves (rounded up)
  {"pip_stage_edges", {64}, 1, 128},    // PIP ring staging depth: 64 or 128 edges
  {"pip_xcd_run", {16}, 1, 256},        // PIP test workgroups per XCD run (1: dispatch order)
  {"pip_box_filter", {1}, 0, 1},        // PIP test skips points outside the polygon's box
  {"pip_polygon_group", {0}, 0, 4096},  // > 0: polygon-major PIP test, pairs per wave
};
}  // namespace

int64_t tuning(Tuning t) { return g_tuning[static_cast<int>(t)].value.load(std::memory_order_relaxed); }

}  // namespace qpip

extern "C" {

int qpip_tuning_set(const char* name, int64_t value)
{
  for (auto& p : qpip::g_tuning) {
    if (name && strcmp(name, p.name) == 0) {
      if (value < p.lo || value > p.hi) {
        qpip::set_last_error(std::string("tuning value out of range for ") + p.name);
        return QPIP_ERROR_LOGIC;
      }
      p.value.store(value, std::memory_order_relaxed);
      return QPIP_SUCCESS;
    }
  }
  qpip::set_last_error(std::string("unknown tuning parameter: ") + (name ? name : "(null)"));
  return QPIP_ERROR_LOGIC;
}

int qpip_tuning_get(const char* name, int64_t* value)
{
  for (auto& p : qpip::g_tuning) {
    if (name && strcmp(name, p.name) == 0) {
      if (value) *value = p.value.load(std::memory_order_relaxed);
      return QPIP_SUCCESS;
    }
  }
  qpip::set_last_error(std::string("unknown tuning parameter: ") + (name ? name : "(null)"));
  return QPIP_ERROR_LOGIC;
}

const char* qpip_last_error_message(void) { return qpip::last_error().c_str(); }

const char* qpip_version(void) { return "cuspatial_amd 0.1.0 (gfx950)"; }

void qpip_free(void* ptr) { qpip::pool_free(ptr); }

void qpip_set_memory_limit(size_t bytes) { qpip::pool_set_limit(bytes); }

void qpip_profile_enable(int on)
{
  std::lock_guard<std::mutex> g(qpip::g_prof_mu);
  qpip::g_prof_on = on != 0;
}

void qpip_profile_reset(void)
{
  std::lock_guard<std::mutex> g(qpip::g_prof_mu);
  qpip::prof_drain_locked();
  qpip::g_prof_totals.clear();
}

int64_t qpip_profile_report(char* buf, size_t len)
{
  std::lock_guard<std::mutex> g(qpip::g_prof_mu);
  try {
    qpip::prof_drain_locked();
  } catch (...) {
  }
  std::string out = "{";
  bool first      = true;
  for (auto const& kv : qpip::g_prof_totals) {
    char item[256];
    snprintf(item, sizeof(item), "%s\"%s\": [%llu, %.6f]", first ? "" : ", ", kv.first.c_str(),
             static_cast<unsigned long long>(kv.second.first), kv.second.second);
    out += item;
    first = false;
  }
  out += "}";
  if (buf && len > 0) {
    size_t const n = out.size() < len - 1 ? out.size() : len - 1;
    memcpy(buf, out.data(), n);
    buf[n] = 0;
  }
  return static_cast<int64_t>(out.size());
}

void qpip_memory_info(size_t* live_bytes, size_t* cached_bytes)
{
  size_t live = 0, cached = 0;
  try {
    qpip::pool_stats(&live, &cached);
  } catch (...) {
  }
  if (live_bytes) *live_bytes = live;
  if (cached_bytes) *cached_bytes = cached;
}

void qpip_trim(void)
{
  try {
    qpip::pool_trim();
  } catch (...) {
  }
}

}  // extern "C"
