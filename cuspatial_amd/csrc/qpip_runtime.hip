// Host runtime of libcuspatial_amd: per-thread error text, the caching device
// pool, and output allocation through a caller resource.
//
// The pool plays the role RMM's pool resource plays for the reference
// (cpp/benchmarks/fixture/rmm_pool_raii.hpp): hot-path temporaries are carved
// from cached device blocks instead of hipMalloc/hipFree per call. Blocks are
// binned by size; a cached block is reused for a request of at least half its
// size. Every entry point synchronizes its stream before returning, so a block
// handed back to the pool is idle and may be reused by any stream.
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

struct DevicePool {
  std::mutex mu;
  std::multimap<size_t, void*> free_blocks;  // size -> block
  std::unordered_map<void*, size_t> live;    // block -> size
  size_t live_bytes   = 0;
  size_t cached_bytes = 0;
};

std::mutex g_pools_mu;
std::vector<DevicePool*> g_pools;
size_t g_limit = 0;  // 0: no cap

DevicePool& pool_for_current_device()
{
  int dev = 0;
  QPIP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_pools_mu);
  if (static_cast<size_t>(dev) >= g_pools.size()) g_pools.resize(dev + 1, nullptr);
  if (!g_pools[dev]) g_pools[dev] = new DevicePool();  // lives until process exit
  return *g_pools[dev];
}

size_t round_size(size_t bytes)
{
  if (bytes <= (size_t{1} << 20)) return (bytes + 511) & ~size_t{511};
  size_t const gran = size_t{2} << 20;
  return (bytes + gran - 1) / gran * gran;
}

void release_cached(DevicePool& p)
{
  for (auto& kv : p.free_blocks) (void)hipFree(kv.second);
  p.free_blocks.clear();
  p.cached_bytes = 0;
}
}  // namespace

void set_last_error(std::string const& msg) { g_last_error = msg; }
std::string const& last_error() { return g_last_error; }

void* pool_alloc(size_t bytes)
{
  if (bytes == 0) return nullptr;
  DevicePool& p     = pool_for_current_device();
  size_t const size = round_size(bytes);
  std::lock_guard<std::mutex> g(p.mu);
  if (g_limit && p.live_bytes + size > g_limit)
    throw oom_error("out of memory: library memory limit of " + std::to_string(g_limit) +
                    " bytes reached (requested " + std::to_string(size) + ")");
  auto it = p.free_blocks.lower_bound(size);
  if (it != p.free_blocks.end() && it->first <= 2 * size) {
    void* blk        = it->second;
    size_t const bsz = it->first;
    p.free_blocks.erase(it);
    p.cached_bytes -= bsz;
    p.live[blk] = bsz;
    p.live_bytes += bsz;
    return blk;
  }
  void* blk      = nullptr;
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
  p.live[blk] = size;
  p.live_bytes += size;
  return blk;
}

void pool_free(void* blk)
{
  if (!blk) return;
  DevicePool& p = pool_for_current_device();
  std::lock_guard<std::mutex> g(p.mu);
  auto it = p.live.find(blk);
  if (it == p.live.end()) return;  // not ours
  size_t const sz = it->second;
  p.live.erase(it);
  p.live_bytes -= sz;
  p.free_blocks.emplace(sz, blk);
  p.cached_bytes += sz;
}

void pool_set_limit(size_t bytes) { g_limit = bytes; }

void pool_trim()
{
  DevicePool& p = pool_for_current_device();
  std::lock_guard<std::mutex> g(p.mu);
  release_cached(p);
}

size_t pool_bytes_in_use()
{
  DevicePool& p = pool_for_current_device();
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
  hipEvent_t e;
  QPIP_HIP(hipEventCreate(&e));
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

bool profiling_enabled() { return g_prof_on; }

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

}  // namespace qpip

extern "C" {

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
