// Shared host/device definitions of the HIP join / groupby / hash-partition library.
//
// Host side: error plumbing behind the C ABI (status codes + thread-local message), the
// allocator wrapper (caller callbacks or stream-ordered hipMallocAsync), and the column
// descriptors handed to kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ldf_hip.h"

namespace ldf {

// ------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define LDF_HIP_TRY(expr)                                                                      \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess) {                                                                    \
      throw ::ldf::Error(LDF_E_HIP, std::string("HIP error ") + hipGetErrorString(_e) + " at " + \
                                      __FILE__ + ":" + std::to_string(__LINE__));              \
    }                                                                                          \
  } while (0)

#define LDF_EXPECTS(cond, msg)                                                  \
  do {                                                                          \
    if (!(cond)) { throw ::ldf::Error(LDF_E_INVALID, std::string(msg)); }       \
  } while (0)

// Runs `f` and converts exceptions into a status code + ldf_last_error() message.
template <typename F>
int guarded(F&& f)
{
  try {
    f();
    return LDF_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc& e) {
    set_last_error(std::string("host allocation failed: ") + e.what());
    return LDF_E_OOM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return LDF_E_INTERNAL;
  }
}

// ------------------------------------------------------------------------------------------
// allocation (all device memory of a call goes through here)
// ------------------------------------------------------------------------------------------
// Keeps freed memory in the current device's default pool (release threshold = max) so that
// stream-ordered allocations after a synchronisation do not go back to the driver.

// Records HIP events on `s` around a kernel launch when ldf_profile_enable(1) is set.
class ProfScope {
 public:
  ProfScope(const char* name, hipStream_t s);
  ~ProfScope();
  ProfScope(const ProfScope&)            = delete;
  ProfScope& operator=(const ProfScope&) = delete;

 private:
  const char* name_;
  hipStream_t s_;
  hipEvent_t a_ = nullptr;
};

// Caching device allocator behind the default allocation path (no ldf_allocator given): blocks
// are rounded to size classes (<= 12.5% slack) and kept per (device, stream) after they are
// freed, so a later allocation on the same stream reuses them in stream order without a
// hipMalloc / hipFree round trip.  On allocation failure the cache of the device is released
// and the allocation retried once.  Throws Error(LDF_E_OOM).
void* pool_allocate(size_t bytes, hipStream_t s);
void pool_deallocate(void* p, hipStream_t s);
void pool_release_all();
void pool_stats(int64_t* live_blocks, int64_t* live_bytes, int64_t* cached_blocks, int64_t* cached_bytes);

class Alloc {
 public:
  Alloc(const ldf_allocator* a, hipStream_t s) : a_(a), s_(s) {}
  hipStream_t stream() const { return s_; }
  const ldf_allocator* raw() const { return a_; }

  void* allocate(size_t bytes)
  {
    if (bytes == 0) bytes = 1;
    bytes = (bytes + 255) & ~size_t(255);
    void* p = nullptr;
    if (a_ != nullptr && a_->alloc != nullptr) {
      p = a_->alloc(a_->ctx, bytes, (void*)s_);
      if (p == nullptr) throw Error(LDF_E_OOM, "allocator returned NULL for " + std::to_string(bytes) + " bytes");
      if ((reinterpret_cast<uintptr_t>(p) & 255) != 0) {  // the kernels store 16-byte vectors
        if (a_->free != nullptr) a_->free(a_->ctx, p, bytes, (void*)s_);
        throw Error(LDF_E_INVALID, "allocator returned a pointer not aligned to 256 bytes");
      }
    } else {
      p = pool_allocate(bytes, s_);
    }
    return p;
  }
  void deallocate(void* p, size_t bytes)
  {
    if (p == nullptr) return;
    if (bytes == 0) bytes = 1;
    bytes = (bytes + 255) & ~size_t(255);
    if (a_ != nullptr && a_->free != nullptr) {
      a_->free(a_->ctx, p, bytes, (void*)s_);
    } else if (a_ == nullptr || a_->alloc == nullptr) {
      pool_deallocate(p, s_);
    }
  }

 private:
  const ldf_allocator* a_;
  hipStream_t s_;
};

// Scoped temporary device buffer.
template <typename T>
class Buf {
 public:
  Buf() = default;
  Buf(Alloc& a, size_t n) : a_(&a), n_(n) { p_ = static_cast<T*>(a.allocate(n * sizeof(T))); }
  Buf(const Buf&)            = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept { *this = std::move(o); }
  Buf& operator=(Buf&& o) noexcept
  {
    reset();
    a_ = o.a_;
    p_ = o.p_;
    n_ = o.n_;
    o.p_ = nullptr;
    o.n_ = 0;
    return *this;
  }
  ~Buf() { reset(); }
  void reset()
  {
    if (p_ != nullptr && a_ != nullptr) a_->deallocate(p_, n_ * sizeof(T));
    p_ = nullptr;
    n_ = 0;
  }
  // hands the buffer to the caller (ownership leaves this object)
  T* release()
  {
    T* p = p_;
    p_   = nullptr;
    n_   = 0;
    return p;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  Alloc* a_ = nullptr;
  T* p_     = nullptr;
  size_t n_ = 0;
};

// ------------------------------------------------------------------------------------------
// column descriptors
// ------------------------------------------------------------------------------------------
inline int type_width(int32_t t)
{
  switch (t) {
    case LDF_TYPE_BOOL:
    case LDF_TYPE_UINT8:
    case LDF_TYPE_INT8: return 1;
    case LDF_TYPE_UINT16:
    case LDF_TYPE_INT16: return 2;
    case LDF_TYPE_UINT32:
    case LDF_TYPE_INT32:
    case LDF_TYPE_FLOAT:
    case LDF_TYPE_DATE32: return 4;
    case LDF_TYPE_UINT64:
    case LDF_TYPE_INT64:
    case LDF_TYPE_DOUBLE:
    case LDF_TYPE_TIMESTAMP: return 8;
    case LDF_TYPE_LARGE_STRING: return 0;
    default: return -1;
  }
}
inline bool is_string(int32_t t) { return t == LDF_TYPE_LARGE_STRING; }
inline bool is_float(int32_t t) { return t == LDF_TYPE_FLOAT || t == LDF_TYPE_DOUBLE; }
inline bool is_signed_int(int32_t t)
{
  return t == LDF_TYPE_INT8 || t == LDF_TYPE_INT16 || t == LDF_TYPE_INT32 || t == LDF_TYPE_INT64 ||
         t == LDF_TYPE_DATE32 || t == LDF_TYPE_TIMESTAMP;
}
inline bool is_unsigned_int(int32_t t)
{
  return t == LDF_TYPE_BOOL || t == LDF_TYPE_UINT8 || t == LDF_TYPE_UINT16 || t == LDF_TYPE_UINT32 ||
         t == LDF_TYPE_UINT64;
}

// Kernel-side view of one column (passed by value).
struct DCol {
  const void* data;
  const uint8_t* valid;
  const int64_t* offsets;
  int32_t type;
  int32_t width;  // 0 for strings
};

inline DCol to_dcol(const ldf_column_view& c)
{
  DCol d;
  d.data    = c.data;
  d.valid   = c.validity;
  d.offsets = c.offsets;
  d.type    = c.type_id;
  d.width   = type_width(c.type_id);
  return d;
}

constexpr int kMaxKeyCols = 8;
struct DKeys {
  DCol c[kMaxKeyCols];
  int32_t n;
};

void validate_view(const ldf_column_view& c, int64_t nrows, const char* what);

// ------------------------------------------------------------------------------------------
// launch helpers
// ------------------------------------------------------------------------------------------
inline unsigned grid_for(int64_t n, int block, int64_t max_blocks = 256 * 64)
{
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (unsigned)g;
}
inline uint64_t next_pow2(uint64_t x)
{
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// ------------------------------------------------------------------------------------------
// host entry points of the shared kernels (defined in the .hip files)
// ------------------------------------------------------------------------------------------
// out[i] = sum_{j<i} in[i] for i in [0, n]; out has n+1 entries; returns nothing (device only)
void exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, Alloc& a);
void exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, int64_t n, Alloc& a);
// Variants that take their scratch from `ws` (scan_workspace_elems(n) words) instead of
// allocating: usable inside a captured launch sequence.
int64_t scan_workspace_elems(int64_t n);
void exclusive_scan_u64_ws(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* ws, hipStream_t s);
void exclusive_scan_u32_to_u64_ws(const uint32_t* in, uint64_t* out, int64_t n, uint64_t* ws, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s);
void fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
// n fills (4- or 8-byte words, at most kMaxFillJobs) in one launch
struct FillJob {
  void* p;
  uint64_t v;
  int64_t n;
  int32_t w;
};
constexpr int kMaxFillJobs = 8;
struct FillJobs {
  FillJob j[kMaxFillJobs];
  int32_t n;
};
void fill_many(const FillJob* jobs, int n, hipStream_t s);
void fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t s);
// Small device-to-host readbacks go through a per-thread pinned buffer (grown on demand and kept
// for the thread's lifetime): copies into pinned memory skip the runtime's staging of pageable
// memory, so several can be queued before one synchronisation.  The buffer holds one readback
// at a time: consume it before the next call that uses it.
void* pinned_stage(size_t bytes);
void copy_to_host(void* dst, const void* dev, size_t bytes, hipStream_t s);  // then synchronises s
uint64_t read_u64(const uint64_t* dev, hipStream_t s);

// Gather `src` rows by `map` (int32 row ids; a negative id produces a null row).
// `force_nullable` allocates a validity buffer even when the source has none.
void gather_column(const ldf_column_view& src, const int32_t* map, int64_t n_out, bool force_nullable,
                   Alloc& a, ldf_column* out);
// The same for `ncols` columns sharing one map (read once per row for up to kMaxGather columns).
// `random_map` says the map visits source rows in no particular order (a build side's rows, a
// permutation): the source rows may then be packed into records read once per output row.
// `null_counts` (device, ncols entries, zeroed by the caller): the output columns' null counts
// are added there and not read back (out[i].null_count is left for the caller to set), so a
// caller gathering several tables reads them all with one synchronisation.
// On error every column this call allocated is freed.
constexpr int kMaxGather = 8;
void gather_columns(const ldf_column_view* const* src, int ncols, const int32_t* map, int64_t n_out,
                    bool force_nullable, Alloc& a, ldf_column* out, bool random_map = false,
                    unsigned long long* null_counts = nullptr);
void alloc_column(int32_t type, int64_t n, bool nullable, int64_t string_bytes, Alloc& a,
                  ldf_column* out);
void free_column(ldf_column* c, Alloc& a);

}  // namespace ldf
