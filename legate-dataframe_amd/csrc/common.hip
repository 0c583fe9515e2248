// Shared kernels and host utilities: device-wide scans, fills, column gather (the
// cudf::gather step of cpp/src/join.cu:129-146 with NULLIFY for negative ids), Legate-layout
// adapters (cpp/src/core/null_mask.cu:38-114, cpp/src/core/ranges.cu:33-138), and the
// error / version entry points of the C ABI.
#include <hip/hip_runtime.h>

#include <cstring>

#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "ldf_common.hpp"
#include "ldf_device.hpp"
#include "ldf_rec.hpp"

namespace ldf {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// ------------------------------------------------------------------------------------------
// default pool + kernel timing
// ------------------------------------------------------------------------------------------
namespace {
// size classes: multiples of 4 KiB up to 1 MiB, then 8 classes per power of two
size_t pool_class(size_t b)
{
  if (b <= (size_t(1) << 20)) return (b + 4095) & ~size_t(4095);
  const int msb     = 63 - __builtin_clzll((unsigned long long)b);
  const size_t step = size_t(1) << (msb - 3);
  return (b + step - 1) & ~(step - 1);
}

struct DevicePool {
  std::mutex mu;
  struct Key {
    int dev;
    hipStream_t s;
    bool operator<(const Key& o) const { return dev != o.dev ? dev < o.dev : s < o.s; }
  };
  std::map<Key, std::map<size_t, std::vector<void*>>> free_blocks;
  std::map<void*, std::pair<int, size_t>> live;  // block -> (device, class)

  void release_device(int dev)  // caller holds mu; `dev` is the current device
  {
    (void)hipDeviceSynchronize();  // queued work may still use a cached block
    for (auto it = free_blocks.begin(); it != free_blocks.end();) {
      if (it->first.dev == dev) {
        for (auto& kv : it->second)
          for (void* p : kv.second) (void)hipFree(p);
        it = free_blocks.erase(it);
      } else {
        ++it;
      }
    }
    (void)hipGetLastError();
  }
};
DevicePool& device_pool()
{
  static DevicePool* p = new DevicePool();  // never destroyed: blocks may outlive static teardown
  return *p;
}
}  // namespace

void* pool_allocate(size_t bytes, hipStream_t s)
{
  int dev = 0;
  LDF_HIP_TRY(hipGetDevice(&dev));
  const size_t cls = pool_class(bytes);
  DevicePool& P    = device_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto fit = P.free_blocks.find({dev, s});
  if (fit != P.free_blocks.end()) {
    // the smallest cached class that holds the request, up to twice its size (sizes of
    // consecutive calls vary, and exact classes alone would keep growing the cache)
    for (auto cit = fit->second.lower_bound(cls); cit != fit->second.end() && cit->first <= 2 * cls; ++cit) {
      if (cit->second.empty()) continue;
      void* p = cit->second.back();
      cit->second.pop_back();
      P.live[p] = {dev, cit->first};
      return p;
    }
  }
  void* p      = nullptr;
  hipError_t e = hipMalloc(&p, cls);
  if (e != hipSuccess || p == nullptr) {
    (void)hipGetLastError();
    P.release_device(dev);  // cached blocks of every stream on this device, then once more
    p = nullptr;
    e = hipMalloc(&p, cls);
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      throw Error(LDF_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
  }
  P.live[p] = {dev, cls};
  return p;
}

void pool_deallocate(void* p, hipStream_t s)
{
  DevicePool& P = device_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.live.find(p);
  if (it == P.live.end()) return;  // not a pool block
  P.free_blocks[{it->second.first, s}][it->second.second].push_back(p);
  P.live.erase(it);
}

void pool_stats(int64_t* live_blocks, int64_t* live_bytes, int64_t* cached_blocks, int64_t* cached_bytes)
{
  DevicePool& P = device_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  int64_t lb = 0, lB = 0, cb = 0, cB = 0;
  for (auto& kv : P.live) {
    ++lb;
    lB += (int64_t)kv.second.second;
  }
  for (auto& per_stream : P.free_blocks)
    for (auto& cls : per_stream.second) {
      cb += (int64_t)cls.second.size();
      cB += (int64_t)(cls.first * cls.second.size());
    }
  if (live_blocks) *live_blocks = lb;
  if (live_bytes) *live_bytes = lB;
  if (cached_blocks) *cached_blocks = cb;
  if (cached_bytes) *cached_bytes = cB;
}

void pool_release_all()
{
  DevicePool& P = device_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  std::vector<int> devs;
  for (auto& kv : P.free_blocks) devs.push_back(kv.first.dev);
  for (int d : devs) {
    (void)hipSetDevice(d);
    P.release_device(d);
  }
  (void)hipSetDevice(cur);
  (void)hipGetLastError();
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
  if (name == nullptr || !prof().on) return;
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
// several fills in one launch: blockIdx.y picks the job
__global__ void k_fill_jobs(FillJobs f)
{
  const FillJob jb = f.j[blockIdx.y];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < jb.n; i += (int64_t)gridDim.x * blockDim.x) {
    if (jb.w == 8) static_cast<uint64_t*>(jb.p)[i] = jb.v;
    else static_cast<uint32_t*>(jb.p)[i] = (uint32_t)jb.v;
  }
}
void fill_many(const FillJob* jobs, int n, hipStream_t s)
{
  LDF_EXPECTS(n >= 0 && n <= kMaxFillJobs, "fill_many: too many jobs");
  FillJobs f{};
  int64_t most = 0;
  for (int i = 0; i < n; ++i) {
    LDF_EXPECTS(jobs[i].w == 4 || jobs[i].w == 8, "fill_many: 4- or 8-byte words");
    f.j[f.n++] = jobs[i];
    most       = std::max<int64_t>(most, jobs[i].n);
  }
  if (f.n == 0 || most <= 0) return;
  hipLaunchKernelGGL(k_fill_jobs, dim3(grid_for(most, 256, 256 * 32), f.n), dim3(256), 0, s, f);
  LDF_HIP_TRY(hipGetLastError());
}
void fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<uint8_t>, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, s, p, v, n);
  LDF_HIP_TRY(hipGetLastError());
}

void* pinned_stage(size_t bytes)
{
  struct Stage {
    void* p    = nullptr;
    size_t cap = 0;
  };
  thread_local Stage st;  // kept for the thread's lifetime
  if (bytes > st.cap) {
    size_t cap = st.cap > 0 ? st.cap : 4096;
    while (cap < bytes) cap *= 2;
    void* p = nullptr;
    LDF_HIP_TRY(hipHostMalloc(&p, cap, hipHostMallocDefault));
    if (st.p != nullptr) (void)hipHostFree(st.p);
    st.p   = p;
    st.cap = cap;
  }
  return st.p;
}

void copy_to_host(void* dst, const void* dev, size_t bytes, hipStream_t s)
{
  void* stage = pinned_stage(bytes);
  LDF_HIP_TRY(hipMemcpyAsync(stage, dev, bytes, hipMemcpyDeviceToHost, s));
  LDF_HIP_TRY(hipStreamSynchronize(s));
  std::memcpy(dst, stage, bytes);
}

uint64_t read_u64(const uint64_t* dev, hipStream_t s)
{
  uint64_t h = 0;
  copy_to_host(&h, dev, sizeof(h), s);
  return h;
}

// ------------------------------------------------------------------------------------------
// device-wide exclusive scan (out has n+1 entries; out[n] = total)
// ------------------------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = (int64_t)kScanBlock * kScanItems;

// Tile sums: the order of the adds is irrelevant, so lanes read consecutive elements (coalesced).
template <typename In>
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_sums(const In* __restrict__ in, int64_t n,
                                                               uint64_t* __restrict__ sums)
{
  __shared__ uint64_t smem[17];
  const int64_t base = blockIdx.x * kScanTile + threadIdx.x;
  uint64_t s         = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int64_t i = base + (int64_t)j * kScanBlock;
    if (i < n) s += (uint64_t)in[i];
  }
  uint64_t t = block_reduce_sum(s, smem);
  if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

// Tile scan: coalesced loads into LDS, each thread scans kScanItems consecutive elements there
// (one pad word per kScanItems keeps those reads off a single bank), coalesced stores back.
__device__ __forceinline__ int scan_pad(int idx) { return idx + (idx >> 4); }

template <typename In>
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_apply(const In* __restrict__ in, int64_t n,
                                                                const uint64_t* __restrict__ tile_off,
                                                                uint64_t* __restrict__ out)
{
  static_assert(kScanItems == 16, "scan_pad assumes 16 items per thread");
  __shared__ uint64_t smem[17];
  __shared__ uint64_t tile[kScanTile + kScanTile / kScanItems];
  const int64_t base = blockIdx.x * kScanTile;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int idx   = j * kScanBlock + threadIdx.x;
    const int64_t i = base + idx;
    tile[scan_pad(idx)] = (i < n) ? (uint64_t)in[i] : 0;
  }
  __syncthreads();
  uint64_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = tile[scan_pad(threadIdx.x * kScanItems + j)];
    s += v[j];
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan(s, smem, &total) + (tile_off ? tile_off[blockIdx.x] : 0);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    tile[scan_pad(threadIdx.x * kScanItems + j)] = run;
    run += v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int idx   = j * kScanBlock + threadIdx.x;
    const int64_t i = base + idx;
    if (i <= n) out[i] = tile[scan_pad(idx)];  // i == n writes the total
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

// Same recursion with scratch taken from a caller-provided workspace (no allocation, so the
// launch sequence can be captured in a graph and replayed).
template <typename In>
static void scan_ws_impl(const In* in, uint64_t* out, int64_t n, uint64_t* ws, hipStream_t s)
{
  int64_t m      = n + 1;
  int64_t ntiles = (m + kScanTile - 1) / kScanTile;
  if (ntiles == 1) {
    hipLaunchKernelGGL(k_scan_tile_apply<In>, dim3(1), dim3(kScanBlock), 0, s, in, n, (const uint64_t*)nullptr, out);
    LDF_HIP_TRY(hipGetLastError());
    return;
  }
  uint64_t* sums = ws;
  uint64_t* offs = ws + ntiles;
  hipLaunchKernelGGL(k_scan_tile_sums<In>, dim3(ntiles), dim3(kScanBlock), 0, s, in, n, sums);
  LDF_HIP_TRY(hipGetLastError());
  scan_ws_impl<uint64_t>(sums, offs, ntiles, ws + 2 * ntiles + 1, s);
  hipLaunchKernelGGL(k_scan_tile_apply<In>, dim3(ntiles), dim3(kScanBlock), 0, s, in, n, (const uint64_t*)offs, out);
  LDF_HIP_TRY(hipGetLastError());
}

int64_t scan_workspace_elems(int64_t n)
{
  int64_t total = 0;
  for (int64_t m = n + 1, nt = (m + kScanTile - 1) / kScanTile; nt > 1; m = nt + 1, nt = (m + kScanTile - 1) / kScanTile)
    total += 2 * nt + 1;
  return total + 1;
}
void exclusive_scan_u64_ws(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* ws, hipStream_t s)
{
  scan_ws_impl<uint64_t>(in, out, n, ws, s);
}
void exclusive_scan_u32_to_u64_ws(const uint32_t* in, uint64_t* out, int64_t n, uint64_t* ws, hipStream_t s)
{
  scan_ws_impl<uint32_t>(in, out, n, ws, s);
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

// Gather of several columns by one map, so the map (4 B/row) is read once per output row for all
// of them.  One wave covers 64 consecutive output rows so each column's validity for those rows
// is one ballot.  String columns get the sum of each 64-row group's lengths here (gsum[c][group])
// and, through packed records, each row's length (uint32, into dst[c]) and source start; once the
// group sums are scanned, k_gather_str_chars writes each row's offset and characters.
struct GatherSet {
  int ncols;
  int width[kMaxGather];  // 1, 2, 4, 8, or 0 for large_utf8
  const void* src[kMaxGather];
  const uint8_t* sval[kMaxGather];
  const int64_t* soff[kMaxGather];
  void* dst[kMaxGather];
  uint64_t* dval[kMaxGather];
  int64_t* start[kMaxGather];  // packed path: each output row's source start of a string column
  uint64_t* gsum[kMaxGather];  // string column: the character count of each 64-row output group
};

// Sum over the wave's 64 lanes, in every lane.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Packed source rows for a gather through a random map: each source row's fields -- fixed-width
// values, a string's start (8 bytes) and length (4 bytes), and a byte of validity bits when any
// column has validity -- in one record, so an output row costs one random record read instead of
// one per column (and per string offset).
struct PackSet {
  int ncols;
  int32_t off[kMaxGather];   // value, or string start (low 56 bits of an 8-byte field)
  int32_t loff[kMaxGather];  // string length
  int32_t voff;              // validity bits (bit c = column c), -1 when no column has validity;
                             // with a string column, the top byte of the first one's start field
  int32_t stride;
};
constexpr uint64_t kStartMask = (1ull << 56) - 1;

// Columns are visited by a loop that is not unrolled (each descriptor read with scalar loads
// when its turn comes), which keeps the scalar registers from spilling.
// Each thread packs kPackRows rows per pass: a column's loads for all of them are issued (branch
// free, rows past the end read row n - 1) before any is put into its record.
constexpr int kPackRows = 4;

__global__ __launch_bounds__(256) void k_pack_rows(GatherSet g, PackSet ps, int64_t n, unsigned char* __restrict__ rec)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r0 < n; r0 += kPackRows * stride) {
    RecWords w[kPackRows];
    uint32_t vb[kPackRows];
    int64_t rr[kPackRows];
#pragma unroll
    for (int k = 0; k < kPackRows; ++k) {
      const int64_t r = r0 + k * stride;
      rr[k]           = r < n ? r : n - 1;
      vb[k]           = 0;
    }
#pragma unroll 1
    for (int c = 0; c < g.ncols; ++c) {
      const int width   = g.width[c];
      const void* any   = width > 0 ? g.src[c] : static_cast<const void*>(g.soff[c]);
      const uint8_t* sv = g.sval[c];
      uint32_t v[kPackRows];
#pragma unroll
      for (int k = 0; k < kPackRows; ++k) v[k] = valid_bit_nb(sv, any, rr[k]);
      if (width > 0) {
        uint64_t x[kPackRows];
#pragma unroll
        for (int k = 0; k < kPackRows; ++k) x[k] = load_bits_nb(g.src[c], width, rr[k]);
#pragma unroll
        for (int k = 0; k < kPackRows; ++k) w[k].put(ps.off[c], x[k]);
      } else {
        const int64_t* soff = g.soff[c];
        int64_t a0[kPackRows], a1[kPackRows];
#pragma unroll
        for (int k = 0; k < kPackRows; ++k) {
          a0[k] = soff[rr[k]];
          a1[k] = soff[rr[k] + 1];
        }
#pragma unroll
        for (int k = 0; k < kPackRows; ++k) {
          w[k].put(ps.off[c], (uint64_t)a0[k]);
          w[k].put(ps.loff[c], (uint64_t)(uint32_t)(a1[k] - a0[k]));
        }
      }
#pragma unroll
      for (int k = 0; k < kPackRows; ++k) vb[k] |= v[k] << c;
    }
#pragma unroll
    for (int k = 0; k < kPackRows; ++k) {
      if (r0 + k * stride >= n) break;
      if (ps.voff >= 0) w[k].put(ps.voff, vb[k]);
      w[k].store(rec, ps.stride, r0 + k * stride);
    }
  }
}

__device__ __forceinline__ void store_width(void* dst, int width, int64_t i, uint64_t v)
{
  switch (width) {
    case 1: static_cast<uint8_t*>(dst)[i] = (uint8_t)v; break;
    case 2: static_cast<uint16_t*>(dst)[i] = (uint16_t)v; break;
    case 4: static_cast<uint32_t*>(dst)[i] = (uint32_t)v; break;
    default: static_cast<uint64_t*>(dst)[i] = v; break;
  }
}

// k_gather_multi through packed records (W = record stride / 8).  The column loop is not
// unrolled: each column's descriptor is read with scalar loads when its turn comes (unrolled,
// the descriptors of every column slot stayed live at once and spilled scalar registers), and
// null counts gather in LDS, one device atomic per column and block at the end.
template <int W>
__global__ __launch_bounds__(256) void k_gather_packed(GatherSet g, PackSet ps, const unsigned char* __restrict__ rec,
                                                       const int32_t* __restrict__ map, int64_t n,
                                                       unsigned long long* __restrict__ null_count)
{
  __shared__ unsigned long long s_nulls[kMaxGather];
  if (threadIdx.x < kMaxGather) s_nulls[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + lane;
    const int32_t m = i < n ? map[i] : -1;
    uint64_t w[W];
#pragma unroll
    for (int k = 0; k < W; ++k) w[k] = 0;
    if (m >= 0) rec_load<W>(rec, m, w);
    const uint32_t vb = ps.voff >= 0 ? (uint32_t)rec_get<W>(w, ps.voff, 1) : 0xFFu;
#pragma unroll 1
    for (int c = 0; c < g.ncols; ++c) {
      const bool ok    = m >= 0 && ((vb >> c) & 1u);
      const int width  = g.width[c];
      if (width == 0) {
        const uint64_t len = ok ? rec_get<W>(w, ps.loff[c], 4) : 0;  // lanes past n hold no row
        if (i < n) {
          static_cast<uint32_t*>(g.dst[c])[i] = (uint32_t)len;
          g.start[c][i]                       = (int64_t)(rec_get<W>(w, ps.off[c], 8) & kStartMask);
        }
        const uint64_t t = wave_sum_u64(len);
        if (lane == 0) g.gsum[c][base >> 6] = t;
      } else if (i < n) {
        store_width(g.dst[c], width, i, m >= 0 ? rec_get<W>(w, ps.off[c], width) : 0);
      }
      uint64_t* dval = g.dval[c];
      if (dval != nullptr) {
        const uint64_t b = __ballot(ok);
        if (lane == 0) {
          dval[base >> 6]        = b;
          const int64_t in_range = n - base < 64 ? n - base : 64;
          const uint64_t nn      = (uint64_t)in_range - (uint64_t)__popcll(b);
          if (nn != 0) atomicAdd(&s_nulls[c], (unsigned long long)nn);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < g.ncols && s_nulls[threadIdx.x] != 0) atomicAdd(&null_count[threadIdx.x], s_nulls[threadIdx.x]);
}

// kGatherGroups groups of 64 output rows per wave and iteration: the map loads of all groups are
// issued, then each column's value loads for all groups, before the stores that use them, so a
// wave keeps several random reads in flight.  Columns are visited by a loop that is not unrolled,
// as in k_gather_packed.
constexpr int kGatherGroups = 4;

__global__ __launch_bounds__(256) void k_gather_multi(GatherSet g, const int32_t* __restrict__ map, int64_t n,
                                                      int64_t n_src, unsigned long long* __restrict__ null_count)
{
  __shared__ unsigned long long s_nulls[kMaxGather];
  if (threadIdx.x < kMaxGather) s_nulls[threadIdx.x] = 0;
  __syncthreads();
  const int lane       = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n;
       base += kGatherGroups * stride) {
    int32_t m[kGatherGroups];
#pragma unroll
    for (int j = 0; j < kGatherGroups; ++j) {  // unconditional loads (rows past n read row n - 1)
      const int64_t i  = base + j * stride + lane;
      const int32_t mv = map[i < n ? i : n - 1];
      m[j]             = i < n ? mv : -1;
    }
#pragma unroll 1
    for (int c = 0; c < g.ncols; ++c) {
      const int width       = g.width[c];
      const uint8_t* sval   = g.sval[c];
      const void* src       = g.src[c];
      const int64_t* soff   = g.soff[c];
      void* dst             = g.dst[c];
      uint64_t* dval        = g.dval[c];
      uint64_t val[kGatherGroups];
      bool ok[kGatherGroups];
      // every group's loads before any use: branch-free (a row without a source row loads source
      // row 0 and is masked); no loads at all from an empty source (every map entry is then -1)
#pragma unroll
      for (int j = 0; j < kGatherGroups; ++j) {
        ok[j]  = false;
        val[j] = 0;
      }
      if (n_src > 0) {
        const void* any = width == 0 ? static_cast<const void*>(soff) : src;
        if (width == 0) {
#pragma unroll
          for (int j = 0; j < kGatherGroups; ++j) {
            const int32_t mm  = m[j] >= 0 ? m[j] : 0;
            const uint32_t vb = valid_bit_nb(sval, any, mm);
            const int64_t a0 = soff[mm], a1 = soff[mm + 1];
            ok[j]             = m[j] >= 0 && vb != 0;
            val[j]            = ok[j] ? (uint64_t)(a1 - a0) : 0;
          }
        } else {
#pragma unroll
          for (int j = 0; j < kGatherGroups; ++j) {
            const int32_t mm  = m[j] >= 0 ? m[j] : 0;
            const uint32_t vb = valid_bit_nb(sval, any, mm);
            const uint64_t x  = load_bits_nb(src, width, mm);
            ok[j]             = m[j] >= 0 && vb != 0;
            val[j]            = m[j] >= 0 ? x : 0;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kGatherGroups; ++j) {
        const int64_t gb = base + j * stride;
        if (gb >= n) break;  // the same for the whole wave
        const int64_t i = gb + lane;
        if (i < n && width != 0) store_width(dst, width, i, val[j]);
        if (width == 0) {  // lanes past n hold 0; the characters kernel reads the length again
          const uint64_t t = wave_sum_u64(val[j]);
          if (lane == 0) g.gsum[c][gb >> 6] = t;
        }
        if (dval != nullptr) {
          const uint64_t bal = __ballot(ok[j]);
          if (lane == 0) {
            dval[gb >> 6]          = bal;
            const int64_t in_range = n - gb < 64 ? n - gb : 64;
            const uint64_t nn      = (uint64_t)in_range - (uint64_t)__popcll(bal);
            if (nn != 0) atomicAdd(&s_nulls[c], (unsigned long long)nn);
          }
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < g.ncols && s_nulls[threadIdx.x] != 0) atomicAdd(&null_count[threadIdx.x], s_nulls[threadIdx.x]);
}

// Characters of gathered strings.  Each wave owns 64 consecutive output rows, whose characters
// form one contiguous destination range [D0, D1): D0 is the scanned sum of the lengths of the
// groups before (gbase[group]), and a wave scan of the rows' lengths gives each row's offset,
// which the wave also stores (dst_off; the last row's wave stores dst_off[n] too).
//  * Rows of at most kShortStr bytes (the usual case): each lane copies its own row into the
//    wave's LDS window -- aligned 16-byte source chunks, up to four loads in flight (an aligned
//    chunk holding a byte of the string never leaves that byte's page), one LDS byte store per
//    character -- and the wave then writes the window as aligned 16-byte stores.
//  * A wave with a longer row: lanes fill aligned 4-byte destination words, finding each byte's
//    row by binary search over the 65 row offsets staged in LDS.
// Destination pieces shared with a neighbouring wave's range are written with byte stores, so no
// two waves write the same byte.  `row_start`, when given, holds each output row's source start
// (read in order instead of the map and the source offsets).
constexpr int kStrBlock = 256;
constexpr int kStrWin   = 2048;  // LDS bytes per wave
constexpr int kShortStr = 256;

__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kStrBlock) void k_gather_str_chars(const int64_t* __restrict__ src_off,
                                                                const uint8_t* __restrict__ src_chars,
                                                                const int32_t* __restrict__ map, int64_t n,
                                                                const uint32_t* __restrict__ lens,
                                                                const uint64_t* __restrict__ dvalid,
                                                                const uint64_t* __restrict__ gbase,
                                                                int64_t* __restrict__ dst_off,
                                                                uint8_t* __restrict__ dst_chars,
                                                                const int64_t* __restrict__ row_start)
{
  __shared__ int64_t s_doff[kStrBlock / 64][65];
  __shared__ int64_t s_soff[kStrBlock / 64][64];
  __shared__ __align__(16) uint8_t s_win[kStrBlock / 64][kStrWin];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t align4  = (int64_t)((uintptr_t)dst_chars & 3);
  const int64_t align16 = (int64_t)((uintptr_t)dst_chars & 15);
  // Waves work independently (all LDS state is per wave): wave w of the grid takes the 64-row
  // groups w, w + waves, ...  The group headers are software-pipelined: while a group's
  // characters are copied, the next group's lengths, base and source starts and the following
  // group's map entries are already in flight.
  const int64_t step = (int64_t)gridDim.x * kStrBlock;
  int64_t base       = ((int64_t)blockIdx.x * (kStrBlock / 64) + wv) * 64;
  auto map_at = [&](int64_t b) -> int32_t { return b + lane < n ? map[b + lane] : -1; };
  // A group's base `d` and this lane's length `l` and source start `st`: with `row_start`, the
  // packed gather's start and uint32 length of each output row; otherwise from the map entry `m`
  // and the source offsets, 0 for a null row (the output validity word the gather wrote, one
  // per group, or no row at all: m < 0).
  auto head = [&](int64_t b, int32_t m, uint64_t& l, int64_t& d, int64_t& st) {
    l  = 0;
    d  = 0;
    st = 0;
    if (b < n) {
      d = (int64_t)gbase[b >> 6];
      if (b + lane < n) {
        if (row_start != nullptr) {
          l  = lens[b + lane];
          st = row_start[b + lane];
        } else {
          const bool ok = m >= 0 && (dvalid == nullptr || ((dvalid[b >> 6] >> lane) & 1));
          if (ok) {
            st = src_off[m];
            l  = (uint64_t)(src_off[m + 1] - st);
          }
        }
      }
    }
  };
  uint64_t hl;             // this lane's length
  int64_t hd, hs;          // the group's base, this lane's source start
  int32_t m_next = -1;     // the next group's map entry of this lane
  head(base, row_start == nullptr ? map_at(base) : -1, hl, hd, hs);
  if (row_start == nullptr) m_next = map_at(base + step);
  for (; base < n; base += step) {
    const int nrows = (int)(n - base < 64 ? n - base : 64);
    const int64_t nb = base + step;
    uint64_t nhl;
    int64_t nhd, nhs;
    int32_t m_after = -1;
    head(nb, m_next, nhl, nhd, nhs);
    if (row_start == nullptr) m_after = map_at(nb + step);
    // this lane's destination range: the group's base plus the lengths of the lanes below
    uint64_t incl = hl;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    const int64_t hb = hd + (int64_t)(incl - hl), he = hd + (int64_t)incl;
    if (lane < nrows) dst_off[base + lane] = hb;
    if (base + nrows == n && lane == nrows - 1) dst_off[n] = he;
    {
      const int64_t D0 = __shfl(hb, 0), D1 = __shfl(he, nrows - 1);
      const int64_t myb = lane < nrows ? hb : D1;
      const int64_t mye = lane < nrows ? he : D1;
      const int64_t mys = lane < nrows ? hs : 0;
      int64_t longest   = mye - myb;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const int64_t t = __shfl_xor(longest, o);
        longest         = t > longest ? t : longest;
      }
      if (longest <= kShortStr) {
        for (int64_t W = ((D0 + align16) & ~(int64_t)15) - align16; W < D1; W += kStrWin) {
          const int64_t We = W + kStrWin;
          const int64_t lo = myb > W ? myb : W, hi = mye < We ? mye : We;
          // the window starts zeroed: rows OR their bytes in, whole dwords at a time, so a dword
          // shared by the end of one row and the start of the next needs no byte stores
          reinterpret_cast<uint4*>(&s_win[wv][0])[lane]      = make_uint4(0, 0, 0, 0);
          reinterpret_cast<uint4*>(&s_win[wv][0])[lane + 64] = make_uint4(0, 0, 0, 0);
          static_assert(kStrWin == 64 * 32, "two 16-byte stores per lane clear the window");
          wave_lds_sync();
          const uintptr_t u0 = (uintptr_t)(src_chars + mys + (lo - myb));
          const uintptr_t u1 = lo < hi ? u0 + (uintptr_t)(hi - lo) : u0;  // this lane's source bytes [u0, u1)
          const bool has     = lo < hi;
          const int64_t off  = (lo - W) - (int64_t)u0;                     // window index of source byte u: u + off
          uint32_t* wdw      = reinterpret_cast<uint32_t*>(&s_win[wv][0]);
          uint32_t carry     = 0;  // the source dword before the batch
          // 64-byte batches of aligned source: four 16-byte loads in flight, then window dwords
          // V_k (k = 0..16) = source bytes [a - 4 + 4k + sig, +4), sig chosen so that V_k lands on
          // a whole window dword, each one byte-aligned out of two consecutive source dwords
          for (uintptr_t a = u0 & ~(uintptr_t)15; __any(has && a < u1); a += 64) {
            uint4 q[4] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0),
                          make_uint4(0, 0, 0, 0)};
#pragma unroll
            for (int k = 0; k < 4; ++k)  // only chunks holding a byte of the row (never leaves its page)
              if (has && a + 16 * k < u1) q[k] = *reinterpret_cast<const uint4*>(a + 16 * k);
            const uint32_t d[18] = {carry,  q[0].x, q[0].y, q[0].z, q[0].w, q[1].x, q[1].y, q[1].z, q[1].w,
                                    q[2].x, q[2].y, q[2].z, q[2].w, q[3].x, q[3].y, q[3].z, q[3].w, 0u};
            // 32-bit offsets from a: the row's bytes [r0, r1), window index of byte a is wa
            const int r0       = (int)((int64_t)u0 - (int64_t)a);
            const int r1       = (int)((int64_t)u1 - (int64_t)a);
            const int wa       = (int)((int64_t)a + off);
            const uint32_t sig = (uint32_t)(-wa) & 3u;
#pragma unroll
            for (int k = 0; k <= 16; ++k) {
              const int t = 4 * k - 4 + (int)sig;  // V_k's first byte, from a
              if (!__any(t < r1)) break;         // every lane is past its row's end
              const int b0 = min(max(r0 - t, 0), 4), b1 = min(max(r1 - t, 0), 4);
              if (b1 > b0) {
                const uint32_t m = (b1 == 4 ? 0xFFFFFFFFu : (1u << (8 * b1)) - 1u) & ~((1u << (8 * b0)) - 1u);
                const uint32_t v = __builtin_amdgcn_alignbyte(d[k + 1], d[k], sig);
                atomicOr(&wdw[(wa + t) >> 2], v & m);
              }
            }
            carry = q[3].w;
          }
          wave_lds_sync();
          const int64_t lim = We < D1 ? We : D1;
          for (int64_t c = W + 16 * lane; c < lim; c += 16 * 64) {
            const int64_t c0 = c > D0 ? c : D0, c1 = c + 16 < lim ? c + 16 : lim;
            if (c0 == c && c1 == c + 16) {
              *reinterpret_cast<uint4*>(dst_chars + c) = *reinterpret_cast<const uint4*>(&s_win[wv][c - W]);
            } else {
              for (int64_t d = c0; d < c1; ++d) dst_chars[d] = s_win[wv][d - W];
            }
          }
          wave_lds_sync();
        }
      } else {
        if (lane < nrows) {
          s_doff[wv][lane] = myb;
          s_soff[wv][lane] = mys;
        }
        if (lane == 0) s_doff[wv][nrows] = D1;
        wave_lds_sync();
        const int64_t W0 = ((D0 + align4) & ~(int64_t)3) - align4;
        for (int64_t wd = W0 + 4 * lane; wd < D1; wd += 256) {
          const int64_t d0 = wd > D0 ? wd : D0;
          int lo = 0, hi = nrows - 1;  // last row whose first byte is <= d0
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_doff[wv][mid] <= d0) lo = mid;
            else hi = mid - 1;
          }
          uint32_t word = 0, bytes = 0;
          int r         = lo;
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const int64_t d = wd + b;
            if (d >= D0 && d < D1) {
              while (d >= s_doff[wv][r + 1]) ++r;
              word |= (uint32_t)src_chars[s_soff[wv][r] + (d - s_doff[wv][r])] << (8 * b);
              bytes |= 1u << b;
            }
          }
          if (bytes == 0xFu) {
            *reinterpret_cast<uint32_t*>(dst_chars + wd) = word;
          } else {
#pragma unroll
            for (int b = 0; b < 4; ++b)
              if ((bytes >> b) & 1u) dst_chars[wd + b] = (uint8_t)(word >> (8 * b));
          }
        }
      }
    }
    wave_lds_sync();  // the wave's LDS state is rewritten by its next group
    hl     = nhl;
    hd     = nhd;
    hs     = nhs;
    m_next = m_after;
  }
}

void gather_columns(const ldf_column_view* const* src, int ncols, const int32_t* map, int64_t n_out,
                    bool force_nullable, Alloc& a, ldf_column* out, bool random_map,
                    unsigned long long* null_counts)
{
  hipStream_t s = a.stream();
  int done      = 0;
  std::vector<Buf<uint32_t>> lens;
  std::vector<Buf<uint64_t>> gsums;
  const int64_t ngroups = (n_out + 63) / 64;
  try {
    for (int c0 = 0; c0 < ncols; c0 += kMaxGather) {
      const int nc = ncols - c0 < kMaxGather ? ncols - c0 : kMaxGather;
      GatherSet g{};
      g.ncols = nc;
      Buf<unsigned long long> own;  // this batch's null counts, unless the caller collects them
      unsigned long long* cnt_p = null_counts != nullptr ? null_counts + c0 : nullptr;
      if (cnt_p == nullptr) {
        own   = Buf<unsigned long long>(a, kMaxGather);
        cnt_p = own.get();
        fill_u64(reinterpret_cast<uint64_t*>(cnt_p), 0, kMaxGather, s);
      }
      for (int j = 0; j < nc; ++j) {
        const ldf_column_view& v = *src[c0 + j];
        ldf_column* o            = &out[c0 + j];
        const int w              = type_width(v.type_id);
        alloc_column(v.type_id, n_out, force_nullable || v.validity != nullptr, 0, a, o);
        ++done;
        g.width[j] = w;
        g.src[j]   = v.data;
        g.sval[j]  = v.validity;
        g.soff[j]  = v.offsets;
        g.dval[j]  = reinterpret_cast<uint64_t*>(o->validity);
        if (w == 0) {
          g.dst[j] = nullptr;  // lengths only on the packed path (below)
          gsums.emplace_back(a, ngroups + 1);
          g.gsum[j] = gsums.back().get();
        } else {
          g.dst[j] = o->data;
        }
      }
      // A random map reads each source row through one packed record when that saves random reads
      // (two or more per output row otherwise) and the output reads most source rows.
      const int64_t n_src = src[c0]->length;
      PackSet ps{};
      ps.voff = -1;
      int reads = 0, nf = 0, widths[2 * kMaxGather + 1], offs[2 * kMaxGather + 1];
      for (int j = 0; j < nc; ++j) {
        if (g.width[j] > 0) {
          widths[nf++] = g.width[j];
          reads += 1;
        } else {
          widths[nf++] = 8;
          widths[nf++] = 4;
          reads += 2;  // the offsets here, the start again for the characters
        }
        if (g.sval[j] != nullptr) reads += 1;
      }
      bool anyval = false, anystr = false;
      for (int j = 0; j < nc; ++j) {
        anyval |= g.sval[j] != nullptr;
        anystr |= g.width[j] == 0;
      }
      // the validity bits ride in the top byte of a string start (starts stay below 2^56)
      if (anyval && !anystr) widths[nf++] = 1;
      ps.stride         = rec_layout(widths, nf, offs);
      const bool packed = random_map && n_out > 0 && n_src > 0 && n_out >= n_src / 4 && reads >= 2 && ps.stride > 0;
      Buf<unsigned char> recs;
      std::vector<Buf<int64_t>> starts;
      if (packed) {
        ps.ncols = nc;
        int f    = 0;
        for (int j = 0; j < nc; ++j) {
          ps.off[j] = offs[f++];
          if (g.width[j] == 0) {
            ps.loff[j] = offs[f++];
            starts.emplace_back(a, n_out);
            g.start[j] = starts.back().get();
            lens.emplace_back(a, n_out);
            g.dst[j] = lens.back().get();
            if (anyval && ps.voff < 0) ps.voff = ps.off[j] + 7;
          }
        }
        if (anyval && !anystr) ps.voff = offs[f++];
        recs = Buf<unsigned char>(a, (size_t)n_src * (size_t)ps.stride);
        hipLaunchKernelGGL(k_pack_rows, dim3(grid_for(n_src, 256, 256 * 16)), dim3(256), 0, s, g, ps, n_src,
                           recs.get());
        LDF_HIP_TRY(hipGetLastError());
      }
      if (n_out > 0) {
        ProfScope prof("gather", s);
        const dim3 grid(grid_for(n_out, 256, 256 * 16));
        if (!packed)
          hipLaunchKernelGGL(k_gather_multi, grid, dim3(256), 0, s, g, map, n_out, n_src, cnt_p);
        else if (ps.stride == 8)
          hipLaunchKernelGGL(k_gather_packed<1>, grid, dim3(256), 0, s, g, ps, (const unsigned char*)recs.get(), map,
                             n_out, cnt_p);
        else if (ps.stride == 16)
          hipLaunchKernelGGL(k_gather_packed<2>, grid, dim3(256), 0, s, g, ps, (const unsigned char*)recs.get(), map,
                             n_out, cnt_p);
        else if (ps.stride == 32)
          hipLaunchKernelGGL(k_gather_packed<4>, grid, dim3(256), 0, s, g, ps, (const unsigned char*)recs.get(), map,
                             n_out, cnt_p);
        else
          hipLaunchKernelGGL(k_gather_packed<8>, grid, dim3(256), 0, s, g, ps, (const unsigned char*)recs.get(), map,
                             n_out, cnt_p);
        LDF_HIP_TRY(hipGetLastError());
      }
      recs.reset();
      // strings: the group sums scanned into group bases, then each row's offset and characters
      size_t gi = gsums.size(), li = lens.size();
      for (int j = nc - 1; j >= 0; --j) {
        if (g.width[j] != 0) continue;
        ldf_column* o            = &out[c0 + j];
        const ldf_column_view& v = *src[c0 + j];
        const uint32_t* len      = packed ? lens[--li].get() : nullptr;
        Buf<uint64_t> gbase(a, ngroups + 1);
        exclusive_scan_u64(gsums[--gi].get(), gbase.get(), ngroups, a);
        gsums[gi].reset();
        const int64_t total = (int64_t)read_u64(gbase.get() + ngroups, s);
        a.deallocate(o->data, 1);
        o->data       = nullptr;
        o->data_bytes = total;
        o->data       = a.allocate((size_t)(total > 0 ? total : 1));
        if (total > 0) {
          ProfScope prof("gather_str", s);
          hipLaunchKernelGGL(k_gather_str_chars, dim3(grid_for(n_out, kStrBlock, 256 * 16)), dim3(kStrBlock), 0, s,
                             v.offsets, static_cast<const uint8_t*>(v.data), map, n_out, len,
                             (const uint64_t*)o->validity, (const uint64_t*)gbase.get(), o->offsets,
                             static_cast<uint8_t*>(o->data), (const int64_t*)g.start[j]);
          LDF_HIP_TRY(hipGetLastError());
        } else {
          fill_u64(reinterpret_cast<uint64_t*>(o->offsets), 0, n_out + 1, s);  // every row empty
        }
      }
      lens.clear();
      gsums.clear();
      starts.clear();
      if (null_counts != nullptr) continue;  // the caller reads them back
      uint64_t hc[kMaxGather];
      copy_to_host(hc, cnt_p, sizeof(hc), s);
      for (int j = 0; j < nc; ++j) out[c0 + j].null_count = out[c0 + j].validity ? (int64_t)hc[j] : 0;
    }
  } catch (...) {
    for (int i = 0; i < done; ++i) free_column(&out[i], a);
    throw;
  }
}

void gather_column(const ldf_column_view& src, const int32_t* map, int64_t n_out, bool force_nullable, Alloc& a,
                   ldf_column* out)
{
  const ldf_column_view* p = &src;
  gather_columns(&p, 1, map, n_out, force_nullable, a, out);
}

// ------------------------------------------------------------------------------------------
// Legate-layout adapters
// ------------------------------------------------------------------------------------------
// One wave per 64 rows: a ballot of the row bytes is the bitmap word; lanes 0..7 store its bytes
// (only the (n+7)/8 bytes of the bitmap are written) and lane 0 counts the word's nulls.
__global__ void k_bools_to_bitmask(const uint8_t* __restrict__ bools, int64_t n, uint8_t* __restrict__ mask,
                                   unsigned long long* null_count)
{
  const int lane         = threadIdx.x & (kWave - 1);
  const int64_t nwords   = (n + 63) / 64;
  const int64_t nwaves   = (int64_t)gridDim.x * (blockDim.x / kWave);
  unsigned long long nul = 0;
  for (int64_t w = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kWave; w < nwords; w += nwaves) {
    const int64_t i     = w * 64 + lane;
    const uint64_t bits = __ballot(i < n && bools[i] != 0);
    const int64_t rem   = n - w * 64;  // rows in this word (1..64)
    const int nbytes    = rem >= 64 ? 8 : (int)((rem + 7) / 8);
    if (lane < nbytes) mask[w * 8 + lane] = (uint8_t)(bits >> (8 * lane));
    if (lane == 0) nul += (unsigned long long)((rem >= 64 ? 64 : rem) - __popcll(bits));
  }
  if (null_count && lane == 0 && nul) atomicAdd(null_count, nul);
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

// HBM ceilings of the running GPU: 16-byte vector loads, four in flight per thread, grid-stride.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_hbm_read(const u32x4* __restrict__ a, int64_t n16, uint32_t* __restrict__ sink)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t acc         = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += 4 * stride) {
    u32x4 v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = i + j * stride < n16 ? a[i + j * stride] : u32x4{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j) acc ^= v[j].x ^ v[j].y ^ v[j].z ^ v[j].w;
  }
  if (acc == 0x9e3779b9u) sink[0] = acc;  // keeps the loads; almost never taken
}

__global__ __launch_bounds__(256) void k_hbm_copy(const u32x4* __restrict__ a, u32x4* __restrict__ b, int64_t n16)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += 4 * stride) {
    u32x4 v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i + j * stride < n16) v[j] = a[i + j * stride];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i + j * stride < n16) b[i + j * stride] = v[j];
  }
}

}  // namespace ldf

using namespace ldf;

extern "C" {

const char* ldf_last_error(void) { return g_last_error.c_str(); }

int ldf_hbm_probe(const void* src, void* dst, int64_t nbytes, int32_t copy, int32_t nblocks, double* ms, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(src != nullptr && dst != nullptr && ms != nullptr, "NULL argument");
    LDF_EXPECTS(nbytes >= 16 && nbytes % 16 == 0, "hbm probe: nbytes must be a positive multiple of 16");
    LDF_EXPECTS(nblocks >= 1 && nblocks <= (1 << 20), "hbm probe: nblocks out of range");
    LDF_EXPECTS(((uintptr_t)src | (uintptr_t)dst) % 16 == 0, "hbm probe: buffers must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    LDF_HIP_TRY(hipEventCreate(&e0));
    LDF_HIP_TRY(hipEventCreate(&e1));
    LDF_HIP_TRY(hipEventRecord(e0, s));
    if (copy)
      hipLaunchKernelGGL(k_hbm_copy, dim3(nblocks), dim3(256), 0, s, (const u32x4*)src, (u32x4*)dst, nbytes / 16);
    else
      hipLaunchKernelGGL(k_hbm_read, dim3(nblocks), dim3(256), 0, s, (const u32x4*)src, nbytes / 16, (uint32_t*)dst);
    hipError_t le = hipGetLastError();
    LDF_HIP_TRY(hipEventRecord(e1, s));
    LDF_HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    LDF_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    LDF_HIP_TRY(le);
    *ms = t;
  });
}

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

void ldf_pool_release(void) { pool_release_all(); }

void* ldf_pool_alloc(size_t bytes, void* stream)
{
  try {
    return pool_allocate(bytes ? bytes : 1, (hipStream_t)stream);
  } catch (...) {
    return nullptr;
  }
}

void ldf_pool_free(void* ptr, void* stream)
{
  if (ptr) pool_deallocate(ptr, (hipStream_t)stream);
}

void ldf_pool_stats(int64_t* live_blocks, int64_t* live_bytes, int64_t* cached_blocks, int64_t* cached_bytes)
{
  pool_stats(live_blocks, live_bytes, cached_blocks, cached_bytes);
}

int ldf_bools_to_bitmask(const uint8_t* bools, int64_t n, uint8_t* bitmask, int64_t* null_count_dev, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(n >= 0, "negative length");
    if (null_count_dev) fill_u64(reinterpret_cast<uint64_t*>(null_count_dev), 0, 1, (hipStream_t)stream);
    if (n == 0) return;
    hipLaunchKernelGGL(k_bools_to_bitmask, dim3(grid_for((n + 63) / 64 * 64, 256)), dim3(256), 0, (hipStream_t)stream,
                       bools, n, bitmask, (unsigned long long*)null_count_dev);
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
