// Host runtime of libcuspatial_amd: per-thread error text, the caching device
// pool, and output allocation through a caller resource.
//
// The pool plays the role RMM's pool resource plays for the reference
// (cpp/benchmarks/fixture/rmm_pool_raii.hpp): hot-path temporaries are carved
// from cached device blocks instead of hipMalloc/hipFree per call. Blocks are
// binned by size; a cached block is reused for a request of at least half its
// size.
//
// A temporary is handed back when it leaves scope, which is often before the
// kernels queued on the call's stream have finished with it. Reusing it in the
// same call is safe (the same stream orders the work), and so is reusing it in a
// later call of the same host thread (every entry point synchronizes its stream
// before returning). Another thread, on another stream, must not get it: its
// kernels could overwrite the block while the first stream still reads it. So
// free blocks are kept per host thread; when a thread exits, its blocks are idle
// and move to a shared list that any thread may draw from.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "qpip_internal.hpp"

namespace qpip {

namespace {
thread_local std::string g_last_error;

using FreeList = std::multimap<size_t, void*>;  // size -> block

struct DevicePool {
  std::mutex mu;
  std::unordered_map<void*, size_t> live;  // block -> size
  FreeList idle;                           // blocks of exited threads
  size_t live_bytes   = 0;
  size_t cached_bytes = 0;
};

std::mutex g_pools_mu;
std::vector<DevicePool*> g_pools;  // never freed: used by thread exit handlers
size_t g_limit = 0;                // 0: no cap

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

// This thread's free blocks, per device. Touched only by its own thread; the
// byte counters it shares are updated under the device pool's mutex.
struct ThreadCache {
  std::vector<std::pair<DevicePool*, FreeList>> lists;  // index: device
  FreeList& list(int dev, DevicePool* p)
  {
    if (static_cast<size_t>(dev) >= lists.size()) lists.resize(dev + 1, {nullptr, FreeList()});
    lists[dev].first = p;
    return lists[dev].second;
  }
  ~ThreadCache()
  {
    for (auto& [p, fl] : lists) {
      if (!p || fl.empty()) continue;
      std::lock_guard<std::mutex> g(p->mu);
      p->idle.insert(fl.begin(), fl.end());
    }
  }
};
thread_local ThreadCache t_cache;

size_t round_size(size_t bytes)
{
  if (bytes <= (size_t{1} << 20)) return (bytes + 511) & ~size_t{511};
  size_t const gran = size_t{2} << 20;
  return (bytes + gran - 1) / gran * gran;
}

bool take(FreeList& fl, size_t size, void*& blk, size_t& bsz)
{
  auto it = fl.lower_bound(size);
  if (it == fl.end() || it->first > 2 * size) return false;
  blk = it->second;
  bsz = it->first;
  fl.erase(it);
  return true;
}

// Frees this thread's cached blocks and the shared idle ones (caller holds p.mu).
void release_cached(DevicePool& p, FreeList& mine)
{
  for (auto& kv : mine) {
    (void)hipFree(kv.second);
    p.cached_bytes -= kv.first;
  }
  mine.clear();
  for (auto& kv : p.idle) {
    (void)hipFree(kv.second);
    p.cached_bytes -= kv.first;
  }
  p.idle.clear();
}
}  // namespace

void set_last_error(std::string const& msg) { g_last_error = msg; }
std::string const& last_error() { return g_last_error; }

void* pool_alloc(size_t bytes)
{
  if (bytes == 0) return nullptr;
  int const dev     = current_device();
  DevicePool& p     = pool_for_device(dev);
  FreeList& mine    = t_cache.list(dev, &p);
  size_t const size = round_size(bytes);
  std::lock_guard<std::mutex> g(p.mu);
  if (g_limit && p.live_bytes + size > g_limit)
    throw oom_error("out of memory: library memory limit of " + std::to_string(g_limit) +
                    " bytes reached (requested " + std::to_string(size) + ")");
  void* blk  = nullptr;
  size_t bsz = 0;
  if (take(mine, size, blk, bsz) || take(p.idle, size, blk, bsz)) {
    p.cached_bytes -= bsz;
    p.live[blk] = bsz;
    p.live_bytes += bsz;
    return blk;
  }
  hipError_t err = hipMalloc(&blk, size);
  if (err != hipSuccess) {
    (void)hipGetLastError();
    release_cached(p, mine);
    err = hipMalloc(&blk, size);
  }
  if (err != hipSuccess) {
    (void)hipGetLastError();
    throw oom_error("out of memory: hipMalloc of " + std::to_string(size) + " bytes failed");
  }
  p.live[blk] = size;
  p.live_bytes += size;
  return blk;
}

void pool_free(void* blk)
{
  if (!blk) return;
  int const dev  = current_device_nothrow();
  DevicePool& p  = pool_for_device(dev);
  FreeList& mine = t_cache.list(dev, &p);
  std::lock_guard<std::mutex> g(p.mu);
  auto it = p.live.find(blk);
  if (it == p.live.end()) return;  // not ours
  size_t const sz = it->second;
  p.live.erase(it);
  p.live_bytes -= sz;
  mine.emplace(sz, blk);
  p.cached_bytes += sz;
}

void pool_set_limit(size_t bytes) { g_limit = bytes; }

void pool_trim()
{
  int const dev  = current_device();
  DevicePool& p  = pool_for_device(dev);
  FreeList& mine = t_cache.list(dev, &p);
  std::lock_guard<std::mutex> g(p.mu);
  release_cached(p, mine);
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
  if (!mr_) return pool_alloc(bytes);
  void* p = mr_->allocate(mr_->ctx, bytes, stream_);
  if (!p) throw oom_error("out of memory: caller resource returned NULL for " +
                          std::to_string(bytes) + " bytes");
  return p;
}

void OutputAllocator::deallocate(void* p, size_t bytes)
{
  if (!p) return;
  if (!mr_)
    pool_free(p);
  else
    mr_->deallocate(mr_->ctx, p, bytes, stream_);
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

void qpip_trim(void)
{
  try {
    qpip::pool_trim();
  } catch (...) {
  }
}

}  // extern "C"
