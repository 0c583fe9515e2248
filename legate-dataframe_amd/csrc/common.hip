// Shared kernels and host utilities: device-wide scans, fills, column gather (the
// cudf::gather step of cpp/src/join.cu:129-146 with NULLIFY for negative ids), Legate-layout
// adapters (cpp/src/core/null_mask.cu:38-114, cpp/src/core/ranges.cu:33-138), and the
// error / version entry points of the C ABI.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "ldf_common.hpp"
#include "ldf_device.hpp"

namespace ldf {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// ------------------------------------------------------------------------------------------
// default pool + kernel timing
// ------------------------------------------------------------------------------------------
void configure_default_pool()
{
  static std::mutex mu;
  static bool done[64] = {};
  int dev             = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
  std::lock_guard<std::mutex> lk(mu);
  if (done[dev]) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t thr = ~0ull;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  (void)hipGetLastError();
  done[dev] = true;
}

namespace {
struct ProfState {
  std::mutex mu;
  bool on = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<int64_t, double>> acc;
  void drain()
  {
    for (auto& p : pending) {
      float ms = 0.f;
      if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        auto& e = acc[p.name];
        e.first += 1;
        e.second += ms;
      }
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending.clear();
  }
};
ProfState& prof()
{
  static ProfState s;
  return s;
}
}  // namespace

ProfScope::ProfScope(const char* name, hipStream_t s) : name_(name), s_(s)
{
  if (!prof().on) return;
  if (hipEventCreate(&a_) != hipSuccess || hipEventRecord(a_, s_) != hipSuccess) a_ = nullptr;
}

ProfScope::~ProfScope()
{
  if (a_ == nullptr) return;
  hipEvent_t b;
  if (hipEventCreate(&b) != hipSuccess || hipEventRecord(b, s_) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(prof().mu);
  prof().pending.push_back({name_, a_, b});
}

void validate_view(const ldf_column_view& c, int64_t nrows, const char* what)
{
  int w = type_width(c.type_id);
  if (w < 0) throw Error(LDF_E_UNSUPPORTED, std::string(what) + ": unsupported type id " + std::to_string(c.type_id));
  if (c.length != nrows) throw Error(LDF_E_INVALID, std::string(what) + ": column lengths differ");
  if (c.length > 0 && c.data == nullptr && !(w == 0 && c.offsets != nullptr))
    throw Error(LDF_E_INVALID, std::string(what) + ": NULL data pointer");
  if (w == 0 && c.offsets == nullptr) throw Error(LDF_E_INVALID, std::string(what) + ": string column without offsets");
}

// ------------------------------------------------------------------------------------------
// fills
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
void fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, s, p, v, n);
  LDF_HIP_TRY(hipGetLastError());
}
void fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, s, p, v, n);
  LDF_HIP_TRY(hipGetLastError());
}
void fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<uint8_t>, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, s, p, v, n);
  LDF_HIP_TRY(hipGetLastError());
}

uint64_t read_u64(const uint64_t* dev, hipStream_t s)
{
  uint64_t h = 0;
  LDF_HIP_TRY(hipMemcpyAsync(&h, dev, sizeof(h), hipMemcpyDeviceToHost, s));
  LDF_HIP_TRY(hipStreamSynchronize(s));
  return h;
}

// ------------------------------------------------------------------------------------------
// device-wide exclusive scan (out has n+1 entries; out[n] = total)
// ------------------------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = (int64_t)kScanBlock * kScanItems;

template <typename In>
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_sums(const In* __restrict__ in, int64_t n,
                                                               uint64_t* __restrict__ sums)
{
  __shared__ uint64_t smem[17];
  int64_t base = blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint64_t s   = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + j;
    if (i < n) s += (uint64_t)in[i];
  }
  uint64_t t = block_reduce_sum(s, smem);
  if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

template <typename In>
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_apply(const In* __restrict__ in, int64_t n,
                                                                const uint64_t* __restrict__ tile_off,
                                                                uint64_t* __restrict__ out)
{
  __shared__ uint64_t smem[17];
  int64_t base = blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + j;
    v[j]      = (i < n) ? (uint64_t)in[i] : 0;
    s += v[j];
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan(s, smem, &total) + (tile_off ? tile_off[blockIdx.x] : 0);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + j;
    if (i <= n) out[i] = run;  // i == n writes the total
    run += v[j];
  }
}

template <typename In>
static void scan_impl(const In* in, uint64_t* out, int64_t n, Alloc& a)
{
  hipStream_t s  = a.stream();
  int64_t m      = n + 1;  // out[n] = total
  int64_t ntiles = (m + kScanTile - 1) / kScanTile;
  if (ntiles == 1) {
    hipLaunchKernelGGL(k_scan_tile_apply<In>, dim3(1), dim3(kScanBlock), 0, s, in, n, (const uint64_t*)nullptr, out);
    LDF_HIP_TRY(hipGetLastError());
    return;
  }
  Buf<uint64_t> sums(a, ntiles);
  Buf<uint64_t> offs(a, ntiles + 1);
  hipLaunchKernelGGL(k_scan_tile_sums<In>, dim3(ntiles), dim3(kScanBlock), 0, s, in, n, sums.get());
  LDF_HIP_TRY(hipGetLastError());
  scan_impl<uint64_t>(sums.get(), offs.get(), ntiles, a);
  hipLaunchKernelGGL(k_scan_tile_apply<In>, dim3(ntiles), dim3(kScanBlock), 0, s, in, n,
                     (const uint64_t*)offs.get(), out);
  LDF_HIP_TRY(hipGetLastError());
}

void exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, Alloc& a) { scan_impl<uint64_t>(in, out, n, a); }
void exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, int64_t n, Alloc& a)
{
  scan_impl<uint32_t>(in, out, n, a);
}

// ------------------------------------------------------------------------------------------
// column allocation / gather
// ------------------------------------------------------------------------------------------
void alloc_column(int32_t type, int64_t n, bool nullable, int64_t string_bytes, Alloc& a, ldf_column* out)
{
  std::memset(out, 0, sizeof(*out));
  out->type_id    = type;
  out->length     = n;
  out->null_count = 0;
  int w           = type_width(type);
  if (w == 0) {
    out->offsets    = static_cast<int64_t*>(a.allocate((size_t)(n + 1) * sizeof(int64_t)));
    out->data_bytes = string_bytes;
    out->data       = a.allocate((size_t)(string_bytes > 0 ? string_bytes : 1));
  } else {
    out->data_bytes = n * w;
    out->data       = a.allocate((size_t)(n > 0 ? n * w : 1));
  }
  if (nullable) {
    // whole 64-bit words so kernels can store one ballot per wave
    out->validity_bytes = ((n + 63) / 64) * 8;
    out->validity       = static_cast<uint8_t*>(a.allocate((size_t)(out->validity_bytes > 0 ? out->validity_bytes : 8)));
  }
}

void free_column(ldf_column* c, Alloc& a)
{
  int w = type_width(c->type_id);
  if (c->data) a.deallocate(c->data, (size_t)(c->data_bytes > 0 ? c->data_bytes : 1));
  if (c->validity) a.deallocate(c->validity, (size_t)(c->validity_bytes > 0 ? c->validity_bytes : 8));
  if (w == 0 && c->offsets) a.deallocate(c->offsets, (size_t)(c->length + 1) * sizeof(int64_t));
  c->data = nullptr;
  c->validity = nullptr;
  c->offsets = nullptr;
}

// One wave covers 64 consecutive output rows so the validity of those rows is one ballot.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_fixed(const T* __restrict__ src, const uint8_t* __restrict__ src_valid,
                                                      const int32_t* __restrict__ map, int64_t n, T* __restrict__ dst,
                                                      uint64_t* __restrict__ dst_valid, unsigned long long* null_count)
{
  const int lane = threadIdx.x & 63;
  uint64_t nulls = 0;
  for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + lane;
    bool ok   = false;
    T v       = T(0);
    if (i < n) {
      int32_t m = map[i];
      if (m >= 0) {
        ok = valid_at(src_valid, m);
        v  = src[m];
      }
      dst[i] = v;
    }
    if (dst_valid != nullptr) {
      uint64_t b = __ballot(ok);
      if (lane == 0) {
        dst_valid[base >> 6] = b;
        int64_t in_range = n - base < 64 ? n - base : 64;
        nulls += (uint64_t)in_range - (uint64_t)__popcll(b);
      }
    }
  }
  if (dst_valid != nullptr && lane == 0 && nulls) atomicAdd(null_count, (unsigned long long)nulls);
}

__global__ __launch_bounds__(256) void k_gather_str_len(const int64_t* __restrict__ src_off, const uint8_t* __restrict__ src_valid,
                                                        const int32_t* __restrict__ map, int64_t n, uint64_t* __restrict__ len,
                                                        uint64_t* __restrict__ dst_valid, unsigned long long* null_count)
{
  const int lane = threadIdx.x & 63;
  uint64_t nulls = 0;
  for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = base + lane;
    bool ok   = false;
    if (i < n) {
      int32_t m  = map[i];
      uint64_t l = 0;
      if (m >= 0) {
        ok = valid_at(src_valid, m);
        if (ok) l = (uint64_t)(src_off[m + 1] - src_off[m]);
      }
      len[i] = l;
    }
    if (dst_valid != nullptr) {
      uint64_t b = __ballot(ok);
      if (lane == 0) {
        dst_valid[base >> 6] = b;
        int64_t in_range = n - base < 64 ? n - base : 64;
        nulls += (uint64_t)in_range - (uint64_t)__popcll(b);
      }
    }
  }
  if (dst_valid != nullptr && lane == 0 && nulls) atomicAdd(null_count, (unsigned long long)nulls);
}

__global__ __launch_bounds__(256) void k_gather_str_chars(const int64_t* __restrict__ src_off, const uint8_t* __restrict__ src_chars,
                                                          const int32_t* __restrict__ map, int64_t n,
                                                          const int64_t* __restrict__ dst_off, uint8_t* __restrict__ dst_chars)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t d = dst_off[i], e = dst_off[i + 1];
    if (e == d) continue;
    int64_t s = src_off[map[i]];
    for (int64_t k = 0; k < e - d; ++k) dst_chars[d + k] = src_chars[s + k];
  }
}

void gather_column(const ldf_column_view& src, const int32_t* map, int64_t n_out, bool force_nullable, Alloc& a,
                   ldf_column* out)
{
  hipStream_t s  = a.stream();
  bool nullable  = force_nullable || src.validity != nullptr;
  int w          = type_width(src.type_id);
  Buf<unsigned long long> nc(a, 1);
  fill_u64(reinterpret_cast<uint64_t*>(nc.get()), 0, 1, s);
  unsigned grid = grid_for(n_out, 256, 256 * 16);
  if (w == 0) {
    alloc_column(src.type_id, n_out, nullable, 0, a, out);
    a.deallocate(out->data, 1);
    out->data = nullptr;
    Buf<uint64_t> len(a, n_out > 0 ? n_out : 1);
    hipLaunchKernelGGL(k_gather_str_len, dim3(grid), dim3(256), 0, s, src.offsets, src.validity, map, n_out, len.get(),
                       (uint64_t*)out->validity, nc.get());
    LDF_HIP_TRY(hipGetLastError());
    exclusive_scan_u64(len.get(), reinterpret_cast<uint64_t*>(out->offsets), n_out, a);
    int64_t total = (int64_t)read_u64(reinterpret_cast<uint64_t*>(out->offsets) + n_out, s);
    out->data_bytes = total;
    out->data       = a.allocate((size_t)(total > 0 ? total : 1));
    if (total > 0) {
      hipLaunchKernelGGL(k_gather_str_chars, dim3(grid), dim3(256), 0, s, src.offsets,
                         static_cast<const uint8_t*>(src.data), map, n_out, out->offsets, static_cast<uint8_t*>(out->data));
      LDF_HIP_TRY(hipGetLastError());
    }
  } else {
    alloc_column(src.type_id, n_out, nullable, 0, a, out);
    if (n_out > 0) {
      ProfScope prof("gather", s);
      switch (w) {
        case 1:
          hipLaunchKernelGGL(k_gather_fixed<uint8_t>, dim3(grid), dim3(256), 0, s, (const uint8_t*)src.data, src.validity, map,
                             n_out, (uint8_t*)out->data, (uint64_t*)out->validity, nc.get());
          break;
        case 2:
          hipLaunchKernelGGL(k_gather_fixed<uint16_t>, dim3(grid), dim3(256), 0, s, (const uint16_t*)src.data, src.validity,
                             map, n_out, (uint16_t*)out->data, (uint64_t*)out->validity, nc.get());
          break;
        case 4:
          hipLaunchKernelGGL(k_gather_fixed<uint32_t>, dim3(grid), dim3(256), 0, s, (const uint32_t*)src.data, src.validity,
                             map, n_out, (uint32_t*)out->data, (uint64_t*)out->validity, nc.get());
          break;
        default:
          hipLaunchKernelGGL(k_gather_fixed<uint64_t>, dim3(grid), dim3(256), 0, s, (const uint64_t*)src.data, src.validity,
                             map, n_out, (uint64_t*)out->data, (uint64_t*)out->validity, nc.get());
      }
      LDF_HIP_TRY(hipGetLastError());
    }
  }
  out->null_count = nullable ? (int64_t)read_u64(reinterpret_cast<const uint64_t*>(nc.get()), s) : 0;
}

// ------------------------------------------------------------------------------------------
// Legate-layout adapters
// ------------------------------------------------------------------------------------------
__global__ void k_bools_to_bitmask(const uint8_t* __restrict__ bools, int64_t n, uint8_t* __restrict__ mask,
                                   unsigned long long* null_count)
{
  uint64_t nulls = 0;
  int64_t nbytes = (n + 7) / 8;
  for (int64_t byte = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; byte < nbytes; byte += (int64_t)gridDim.x * blockDim.x) {
    uint8_t m = 0;
    for (int b = 0; b < 8; ++b) {
      int64_t i = byte * 8 + b;
      if (i < n) {
        if (bools[i]) m |= (uint8_t)(1u << b);
        else ++nulls;
      }
    }
    mask[byte] = m;
  }
  if (null_count && nulls) atomicAdd(null_count, (unsigned long long)nulls);
}

__global__ void k_bitmask_to_bools(const uint8_t* __restrict__ mask, int64_t n, uint8_t* __restrict__ bools)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bools[i] = (mask[i >> 3] >> (i & 7)) & 1;
}

// Rect<1>{lo, hi} with inclusive hi and a global origin -> local offsets (offsets[0] = 0).
__global__ void k_ranges_to_offsets(const int64_t* __restrict__ ranges, int64_t n, int64_t* __restrict__ offsets)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x) {
    if (n == 0) {
      offsets[i] = 0;
      continue;
    }
    int64_t base = ranges[0];
    offsets[i]   = (i < n) ? ranges[2 * i] - base : ranges[2 * (n - 1) + 1] + 1 - base;
  }
}

__global__ void k_offsets_to_ranges(const int64_t* __restrict__ offsets, int64_t n, int64_t base, int64_t* __restrict__ ranges)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    ranges[2 * i]     = base + offsets[i];
    ranges[2 * i + 1] = base + offsets[i + 1] - 1;
  }
}

}  // namespace ldf

using namespace ldf;

extern "C" {

const char* ldf_last_error(void) { return g_last_error.c_str(); }

void ldf_profile_enable(int on)
{
  std::lock_guard<std::mutex> lk(prof().mu);
  prof().on = on != 0;
}

void ldf_profile_reset(void)
{
  std::lock_guard<std::mutex> lk(prof().mu);
  prof().drain();
  prof().acc.clear();
}

int ldf_profile_get(const char* kernel, int64_t* launches, double* total_ms)
{
  std::lock_guard<std::mutex> lk(prof().mu);
  prof().drain();
  auto it = prof().acc.find(kernel ? kernel : "");
  if (launches) *launches = it == prof().acc.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == prof().acc.end() ? 0.0 : it->second.second;
  return it == prof().acc.end() ? LDF_E_INVALID : LDF_OK;
}
const char* ldf_version(void) { return "ldf-hip 0.1.0 (gfx950)"; }

void ldf_free_table(ldf_table_out* out, const ldf_allocator* alloc, void* stream)
{
  if (out == nullptr || out->cols == nullptr) return;
  Alloc a(alloc, (hipStream_t)stream);
  for (int i = 0; i < out->ncols; ++i) free_column(&out->cols[i], a);
}

int ldf_bools_to_bitmask(const uint8_t* bools, int64_t n, uint8_t* bitmask, int64_t* null_count_dev, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(n >= 0, "negative length");
    if (n == 0) return;
    hipLaunchKernelGGL(k_bools_to_bitmask, dim3(grid_for((n + 7) / 8, 256)), dim3(256), 0, (hipStream_t)stream, bools, n,
                       bitmask, (unsigned long long*)null_count_dev);
    LDF_HIP_TRY(hipGetLastError());
  });
}

int ldf_bitmask_to_bools(const uint8_t* bitmask, int64_t n, uint8_t* bools, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(n >= 0, "negative length");
    if (n == 0) return;
    hipLaunchKernelGGL(k_bitmask_to_bools, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, bitmask, n, bools);
    LDF_HIP_TRY(hipGetLastError());
  });
}

int ldf_ranges_to_offsets(const int64_t* ranges, int64_t n, int64_t* offsets, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(n >= 0, "negative length");
    hipLaunchKernelGGL(k_ranges_to_offsets, dim3(grid_for(n + 1, 256)), dim3(256), 0, (hipStream_t)stream, ranges, n, offsets);
    LDF_HIP_TRY(hipGetLastError());
  });
}

int ldf_offsets_to_ranges(const int64_t* offsets, int64_t n, int64_t global_base, int64_t* ranges, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(n >= 0, "negative length");
    if (n == 0) return;
    hipLaunchKernelGGL(k_offsets_to_ranges, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, offsets, n,
                       global_base, ranges);
    LDF_HIP_TRY(hipGetLastError());
  });
}

}  // extern "C"
