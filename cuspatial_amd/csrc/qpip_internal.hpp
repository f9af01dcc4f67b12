// Internal declarations shared by the HIP translation units of libcuspatial_amd.
// Host side: error types, the caching device pool, output allocation through a
// caller-supplied resource. Device side: wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cuspatial_amd.h"

namespace qpip {

// ----------------------------------------------------------------------------
// Errors. Mirrors the reference's two exception types (cpp/include/cuspatial/
// error.hpp:30-50): logic_error for violated preconditions, cuda_error for
// runtime failures. They are translated to status codes at the C boundary.
// ----------------------------------------------------------------------------
struct logic_error : std::logic_error {
  using std::logic_error::logic_error;
};
struct device_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct oom_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// A launch produced more rows than a caller-sized capacity holds.
struct capacity_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define QPIP_STR2(x) #x
#define QPIP_STR(x)  QPIP_STR2(x)

// Same reason strings as CUSPATIAL_EXPECTS; the location names this library.
#define QPIP_EXPECTS(cond, reason)                                                              \
  do {                                                                                          \
    if (!(cond))                                                                                \
      throw ::qpip::logic_error("cuSpatial failure at: " __FILE_NAME__ ":" QPIP_STR(__LINE__) ": " \
                                reason);                                                        \
  } while (0)

#define QPIP_HIP(call)                                                                        \
  do {                                                                                        \
    hipError_t qpip_e_ = (call);                                                              \
    if (qpip_e_ != hipSuccess) {                                                              \
      if (qpip_e_ == hipErrorOutOfMemory)                                                     \
        throw ::qpip::oom_error(std::string("HIP out of memory at " __FILE__ ":" QPIP_STR(    \
                                  __LINE__) ": ") +                                           \
                                hipGetErrorString(qpip_e_));                                  \
      throw ::qpip::device_error(std::string("HIP error at " __FILE__ ":" QPIP_STR(__LINE__) \
                                             ": ") +                                          \
                                 hipGetErrorName(qpip_e_) + " " + hipGetErrorString(qpip_e_)); \
    }                                                                                         \
  } while (0)

#define QPIP_LAUNCH_CHECK() QPIP_HIP(hipGetLastError())

void set_last_error(std::string const& msg);

// Records the message of the exception in flight and returns its status code.
// Call only inside a catch block: every C entry point ends with
// `catch (...) { return qpip::error_status(); }`.
inline int error_status()
{
  try {
    throw;
  } catch (logic_error const& e) {
    set_last_error(e.what());
    return QPIP_ERROR_LOGIC;
  } catch (oom_error const& e) {
    set_last_error(e.what());
    return QPIP_ERROR_OOM;
  } catch (capacity_error const& e) {
    set_last_error(e.what());
    return QPIP_ERROR_CAPACITY;
  } catch (std::exception const& e) {
    set_last_error(e.what());
    return QPIP_ERROR_DEVICE;
  } catch (...) {
    set_last_error("cuSpatial failure: unknown exception");
    return QPIP_ERROR_DEVICE;
  }
}

// Runs f and returns QPIP_SUCCESS, or the status code of what it threw.
template <class F>
int guarded(F&& f)
{
  try {
    f();
    return QPIP_SUCCESS;
  } catch (...) {
    return error_status();
  }
}

// Runtime tuning parameters (qpip_tuning_set / qpip_tuning_get), read at launch.
enum class Tuning : int {
  PipMinPairsPerWave = 0,
  PipPairsShift      = 1,
  PipWavesPerBlock   = 2,
  PipStageEdges      = 3,
  PipXcdRun          = 4,
  PipBoxFilter       = 5,
  PipPolygonGroup    = 6
};
constexpr int kNumTuning = 7;
int64_t tuning(Tuning t);

// ----------------------------------------------------------------------------
// Caching device pool (library-owned temporaries and default-resource outputs).
//
// The pool is stream-ordered: a block freed while kernels queued on a stream
// may still use it goes back to the pool at once, and is handed out again
//   * on the same stream with no wait (the stream orders the work), or
//   * on another stream behind a hipStreamWaitEvent on the completion event of
//     the call that last used it.
// Each C entry point opens a StreamScope for the stream its work runs on; when
// the outermost scope ends, one event is recorded on that stream and stamped on
// every block the call allocated or freed. So entry points need not synchronise
// their stream before returning, and temporaries of one thread's call can serve
// another thread's stream safely.
// ----------------------------------------------------------------------------
void* pool_alloc(size_t bytes);
void pool_free(void* p);
void pool_set_limit(size_t bytes);
void pool_trim();
size_t pool_bytes_in_use();
void pool_stats(size_t* live, size_t* cached);

class StreamScope {
 public:
  explicit StreamScope(hipStream_t stream);
  ~StreamScope();
  StreamScope(StreamScope const&)            = delete;
  StreamScope& operator=(StreamScope const&) = delete;
};

// Inside a CompletedScope, freed blocks are known idle (the caller has
// synchronised every stream that used them): any stream may reuse them at once.
class CompletedScope {
 public:
  CompletedScope();
  ~CompletedScope();
  CompletedScope(CompletedScope const&)            = delete;
  CompletedScope& operator=(CompletedScope const&) = delete;
};

// RAII temporary buffer from the pool.
template <class T>
class dbuf {
 public:
  dbuf() = default;
  explicit dbuf(size_t n) { reset(n); }
  ~dbuf() { release(); }
  dbuf(dbuf const&)            = delete;
  dbuf& operator=(dbuf const&) = delete;
  dbuf(dbuf&& o) noexcept : p_(o.p_), n_(o.n_)
  {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  dbuf& operator=(dbuf&& o) noexcept
  {
    if (this != &o) {
      release();
      p_   = o.p_;
      n_   = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void reset(size_t n)
  {
    release();
    n_ = n;
    p_ = n ? static_cast<T*>(pool_alloc(n * sizeof(T))) : nullptr;
  }
  void release()
  {
    if (p_) pool_free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* release_ownership()
  {
    T* p = p_;
    p_   = nullptr;
    n_   = 0;
    return p;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_     = nullptr;
  size_t n_ = 0;
};

// Output buffers: from the caller's resource when one is given, else from the
// pool (then freed with qpip_free). Outputs are owned here until keep() hands
// them to the caller: an entry point that throws releases what it allocated.
class OutputAllocator {
 public:
  OutputAllocator(qpip_resource const* mr, hipStream_t stream) : mr_(mr), stream_(stream) {}
  ~OutputAllocator();
  OutputAllocator(OutputAllocator const&)            = delete;
  OutputAllocator& operator=(OutputAllocator const&) = delete;
  void* allocate(size_t bytes);
  void deallocate(void* p, size_t bytes);
  // The caller owns every output allocated so far.
  void keep() { owned_.clear(); }
  // Move a pool buffer into an output: no copy for the default resource.
  template <class T>
  T* adopt(dbuf<T>& tmp, size_t n)
  {
    if (n == 0) return nullptr;
    if (!mr_) {
      T* p = tmp.release_ownership();
      owned_.push_back({p, n * sizeof(T)});
      return p;
    }
    T* out = static_cast<T*>(allocate(n * sizeof(T)));
    QPIP_HIP(hipMemcpyAsync(out, tmp.get(), n * sizeof(T), hipMemcpyDeviceToDevice, stream_));
    return out;
  }
  template <class T>
  T* make(size_t n)
  {
    return n ? static_cast<T*>(allocate(n * sizeof(T))) : nullptr;
  }

 private:
  struct Owned {
    void* p;
    size_t bytes;
  };
  qpip_resource const* mr_;
  hipStream_t stream_;
  std::vector<Owned> owned_;
};

// ----------------------------------------------------------------------------
// Kernel timing (qpip_profile_*): when enabled, each launch site records a pair
// of HIP events on its stream; the report sums elapsed time per kernel name.
// ----------------------------------------------------------------------------
bool profiling_enabled();
void set_capturing(bool on);  // this thread is capturing a launch sequence
void profile_begin(char const* name, hipStream_t stream, void** token);
void profile_end(void* token, hipStream_t stream);

class KernelTimer {
 public:
  KernelTimer(char const* name, hipStream_t stream) : stream_(stream)
  {
    if (profiling_enabled()) profile_begin(name, stream, &token_);
  }
  ~KernelTimer()
  {
    if (token_) profile_end(token_, stream_);
  }
  KernelTimer(KernelTimer const&)            = delete;
  KernelTimer& operator=(KernelTimer const&) = delete;

 private:
  hipStream_t stream_;
  void* token_ = nullptr;
};

// ----------------------------------------------------------------------------
// Device helpers (wave64).
// ----------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }

__device__ __forceinline__ uint64_t lanemask_lt()
{
  int const l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
  int const l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v)
{
  int const l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint64_t t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Block-wide exclusive scan of one value per thread (blockDim.x == NT, NT % 64 == 0).
// `lds` needs NT / 64 + 1 entries. Returns the exclusive prefix; *total gets the sum.
template <int NT, class T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T* total)
{
  constexpr int NW = NT / kWave;
  int const w      = threadIdx.x / kWave;
  int const l      = lane_id();
  T incl           = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T t = __shfl_up(incl, o, kWave);
    if (l >= o) incl += t;
  }
  if (l == kWave - 1) lds[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    T acc = 0;
    for (int i = 0; i < NW; ++i) {
      T const c = lds[i];
      lds[i]    = acc;
      acc += c;
    }
    lds[NW] = acc;
  }
  __syncthreads();
  T const r = lds[w] + incl - v;
  if (total) *total = lds[NW];
  __syncthreads();
  return r;
}

// ----------------------------------------------------------------------------
// Radix sort (qpip_sort.hip).
// Stable LSD sort of (key, value) pairs on bits [0, bits) of the key, 8-bit digits.
// keys_a holds the input keys and, with keys_b, is overwritten; values_in == nullptr
// means the values are the identity 0..n-1; the sorted values always end in
// values_out (values_tmp is scratch).
//
// Every pass works on tiles of kSortTile consecutive pairs (one workgroup each).
// The first pass's digit counts per tile (tiles x 256) may come from the producer
// of the keys (`first_counts`, e.g. the keygen kernel; they are overwritten); they
// are counted here otherwise.
// ----------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortItems   = 16;
constexpr int kSortTile    = kSortThreads * kSortItems;
constexpr int kRadixBits   = 8;
constexpr int kRadix       = 1 << kRadixBits;

inline uint64_t sort_tiles(uint64_t n) { return (n + kSortTile - 1) / kSortTile; }

struct SortResult {
  uint32_t* keys;
  uint32_t* values;
};
int radix_passes(int bits);
void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream);
SortResult radix_sort_pairs(uint32_t* keys_a,
                            uint32_t* keys_b,
                            uint32_t const* values_in,
                            uint32_t* values_out,
                            uint32_t* values_tmp,
                            uint64_t n,
                            int bits,
                            hipStream_t stream,
                            uint32_t* first_counts       = nullptr,
                            char const* tag              = "sort",
                            uint32_t* counts_ws          = nullptr,
                            uint32_t* sums_ws            = nullptr);
// Segmented variant: the pairs form segments (consecutive ranges) sorted each on
// its own, with tiles that never straddle two segments. Tile t covers positions
// [start, start + count) of segment `segment`; segment s covers positions
// [begin, ...) and tiles [tile_begin, tile_end). The values are the positions (the
// identity); the first pass's digit counts per tile (tiles x 256) come from the
// producer of the keys; the last pass writes each key OR-ed with its segment's
// `key_or` (the job number above the Morton key, for the batched forest) and,
// when `values_local` is given, each value minus its segment's `begin` there
// (with `values_job_local`, values_out itself gets those); `key_samples`
// (ceil(n / 32) entries, or null) receives every 32nd sorted key.
struct SortTile {
  uint32_t start, count, segment, pad;
};
struct SortSegment {
  uint32_t begin, tile_begin, tile_end, key_or;
};
SortResult segmented_radix_sort_pairs(uint32_t* keys_a,
                                      uint32_t* keys_b,
                                      uint32_t* values_out,
                                      uint32_t* values_tmp,
                                      uint64_t n,
                                      int bits,
                                      SortTile const* tiles,
                                      uint64_t num_tiles,
                                      SortSegment const* segments,
                                      uint64_t num_segments,
                                      uint32_t* first_counts,
                                      uint32_t* values_local,
                                      uint32_t* key_samples,
                                      bool values_job_local,
                                      hipStream_t stream,
                                      char const* tag);
// Tiles of segments [begin[s], begin[s + 1]) of kSortTile positions (the last of a
// segment shorter; empty segments get none).
void make_sort_tiles(std::vector<uint32_t> const& begin, std::vector<SortTile>& tiles,
                     std::vector<SortSegment>& segments, std::vector<uint32_t> const& key_or);

// Workspace of radix_sort_pairs for n pairs when the caller provides it (no
// allocation inside the call, as a captured launch sequence needs): counts_ws
// (unused when first_counts is given) and sums_ws.
inline uint64_t sort_counts_elems(uint64_t n) { return sort_tiles(n) * kRadix; }
inline uint64_t sort_sums_elems(uint64_t n) { return (sort_tiles(n) + 63) / 64 * kRadix; }

inline unsigned grid_for(uint64_t n, int threads, unsigned cap = 1u << 20)
{
  uint64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

}  // namespace qpip
