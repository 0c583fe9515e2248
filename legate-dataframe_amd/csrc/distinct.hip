// Distinct rows of a table by a set of key columns.
//
// Reference: DistinctTask::gpu_variant (cpp/src/stream_compaction.cu:38-72) runs
// cudf::distinct(table, keys, KEEP_ANY, null_equality::EQUAL, nan_equality::ALL_EQUAL) on the
// rows a hash repartition brought to this rank; the client (cpp/src/stream_compaction.cpp:120-146)
// returns every column of the table, in the table's order.  On one GPU the repartition keeps every
// row, so `high_cardinality` (which only decides whether a local distinct runs before the
// shuffle) changes nothing.
//
// One row per distinct key tuple is kept: the first row of each key (KEEP_ANY allows any; the
// first makes the output deterministic), and the output rows keep their input order.  Keys compare
// as cudf::distinct compares them there: a null equals a null, every NaN equals every NaN
// (nan_equality::ALL_EQUAL) and -0.0 equals 0.0; other values by their bits.
//
// Layout in HBM: an open-addressing set of 64-bit slots (2^k >= n/4 first, for inputs of 2^20
// rows or more whose sampled key count is at most about n/5; 2^k >= 2n when that set overflows,
// for more keys or for smaller inputs), each (high 32 bits of the
// row's 64-bit key hash << 32 | row index), installed with one atomicCAS; a row whose key is
// already present lowers the slot's row with atomicMin (same hash bits, so the minimum is the
// smaller row), so once every row is in, each key's slot holds its first row.  Slot values only
// decrease, so a row that reads its key's slot with a smaller row already there skips the atomic.
// The kept rows are then flagged from a sequential scan of the slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ldf_agg.hpp"
#include "ldf_common.hpp"
#include "ldf_device.hpp"

namespace ldf {

namespace {

constexpr unsigned long long kDEmpty = ~0ull;  // row index 0xFFFFFFFF never occurs (< 2^31 rows)
constexpr int kDThreads              = 256;
constexpr int kDRowsPerThread        = 16;
constexpr int kDBlockRows            = kDThreads * kDRowsPerThread;
constexpr int64_t kDSmallFrom        = int64_t(1) << 20;  // inputs from this size try the n/4-slot set first
constexpr int kDProbeLimit           = 128;               // probes per row in that set before it gives up

// Key bits with every NaN mapped to one NaN and -0.0 to 0.0 (other values unchanged).
__device__ __forceinline__ uint64_t key_bits(const DCol& c, int64_t row)
{
  const uint64_t b = load_bits(c.data, c.width, row);
  return is_float_type(c.type) ? murmur_input_bits(b, c.width) : b;
}

// The key columns' descriptors, in device memory (any number of key columns).
struct DKeysP {
  const DCol* c;
  int32_t n;
};

// Up to 32 descriptors per launch, passed by value and stored into the device array.
constexpr int kDColChunk = 32;
struct DColChunk {
  DCol c[kDColChunk];
  int32_t n;
  int32_t offset;
};
__global__ void k_put_dcols(DColChunk ch, DCol* __restrict__ dst)
{
  if ((int)threadIdx.x < ch.n) dst[ch.offset + threadIdx.x] = ch.c[threadIdx.x];
}

__device__ inline uint64_t key_hash(const DKeysP& k, int64_t row)
{
  uint64_t h = 0;
  for (int i = 0; i < k.n; ++i) {
    const DCol& c = k.c[i];
    uint64_t e;
    if (!valid_at(c.valid, row)) e = 0x2545f4914f6cdd1dull;
    else if (c.width == 0) e = string_hash64(static_cast<const uint8_t*>(c.data) + c.offsets[row],
                                             c.offsets[row + 1] - c.offsets[row]);
    else e = mix64(key_bits(c, row) + 0x9e3779b97f4a7c15ull);
    h = mix64(h ^ (e + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)));
  }
  return h;
}

__device__ inline bool keys_equal(const DKeysP& k, int64_t ra, int64_t rb)
{
  for (int i = 0; i < k.n; ++i) {
    const DCol& c = k.c[i];
    const bool va = valid_at(c.valid, ra), vb = valid_at(c.valid, rb);
    if (va != vb) return false;
    if (!va) continue;
    if (c.width == 0) {
      if (!strings_equal(c, ra, c, rb)) return false;
    } else if (key_bits(c, ra) != key_bits(c, rb)) {
      return false;
    }
  }
  return true;
}

// Inserts every row into the set; err = 1 when a row found no room within `plimit` probes (with
// the full-size set, plimit = its capacity and load <= 1/2: cannot happen, kept as a guard).
// Once err is set the pass is void (it reruns with the full-size set), so rows stop probing.
__global__ __launch_bounds__(kDThreads) void k_distinct_insert(DKeysP k, int64_t n, unsigned long long* __restrict__ set,
                                                               uint64_t mask, uint64_t plimit, int32_t* __restrict__ err)
{
  const volatile int32_t* failed = err;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    if (*failed) return;
    const uint64_t h              = key_hash(k, r);
    const unsigned long long mine = ((h >> 32) << 32) | (uint64_t)(uint32_t)r;
    uint64_t i                    = h & mask;
    bool placed                   = false;
    for (uint64_t probes = 0; probes < plimit; ++probes) {
      unsigned long long cur = set[i];
      if (cur == kDEmpty) {
        cur = atomicCAS(&set[i], kDEmpty, mine);
        if (cur == kDEmpty) {
          placed = true;
          break;
        }
      }
      // cur holds a row of the slot's key (its row may still be lowered, never its key); a stale
      // read only shows a larger row than the slot holds, so cur's row < r means nothing to lower
      if ((cur >> 32) == (h >> 32) && keys_equal(k, (int64_t)(uint32_t)cur, r)) {
        if ((uint32_t)cur > (uint32_t)r) atomicMin(&set[i], mine);
        placed = true;
        break;
      }
      i = (i + 1) & mask;
    }
    if (!placed) *err = 1;
  }
}

// Cardinality sample: S rows at a stride, their key hashes into a set of 2^20 slots; dups
// counts the sampled rows whose hash was already there.  Sampling a fraction f of rows whose D keys
// repeat evenly gives about S^2 / (2 D) repeats (the birthday count), so few repeats mean many
// keys -- more than the n/4-slot set holds.
constexpr int kDSample     = 1 << 18;
constexpr int kDSampleBits = 20;
__global__ __launch_bounds__(kDThreads) void k_distinct_sample(DKeysP k, int64_t S, int64_t stride,
                                                               unsigned long long* __restrict__ set,
                                                               unsigned long long* __restrict__ dups)
{
  unsigned long long mine = 0;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < S; j += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h = key_hash(k, j * stride) | 1ull;  // never the empty marker 0
    uint64_t i       = h & ((1ull << kDSampleBits) - 1);
    while (true) {
      const unsigned long long old = atomicCAS(&set[i], 0ull, (unsigned long long)h);
      if (old == 0ull) break;
      if (old == h) {
        ++mine;
        break;
      }
      i = (i + 1) & ((1ull << kDSampleBits) - 1);
    }
  }
  if (mine) atomicAdd(dups, mine);
}

// keep[row] = 1 for the row each occupied slot holds (keep cleared before).
__global__ __launch_bounds__(kDThreads) void k_distinct_mark(const unsigned long long* __restrict__ set, uint64_t cap,
                                                             uint8_t* __restrict__ keep)
{
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long v = set[i];
    if (v != kDEmpty) keep[(uint32_t)v] = 1;
  }
}

// Per-block counts of kept rows.  Block b covers rows [b * kDBlockRows, (b + 1) * kDBlockRows),
// thread t the kDRowsPerThread rows from b * kDBlockRows + t * kDRowsPerThread, so a block scan
// of the per-thread counts orders the kept rows by row index (as k_distinct_emit does).
__global__ __launch_bounds__(kDThreads) void k_distinct_count(int64_t n, const uint8_t* __restrict__ keep,
                                                              uint32_t* __restrict__ blk_cnt)
{
  __shared__ uint64_t smem[17];
  const int64_t r0 = (int64_t)blockIdx.x * kDBlockRows + (int64_t)threadIdx.x * kDRowsPerThread;
  uint32_t cnt     = 0;
#pragma unroll
  for (int j = 0; j < kDRowsPerThread; ++j) cnt += r0 + j < n ? keep[r0 + j] : 0u;
  const uint64_t tot = block_reduce_sum(cnt, smem);
  if (threadIdx.x == 0) blk_cnt[blockIdx.x] = (uint32_t)tot;
}

// map[blk_off[b] + rank] = r for every kept row r, in row order.
__global__ __launch_bounds__(kDThreads) void k_distinct_emit(int64_t n, const uint8_t* __restrict__ keep,
                                                             const uint64_t* __restrict__ blk_off,
                                                             int32_t* __restrict__ map)
{
  __shared__ uint64_t smem[17];
  const int64_t r0 = (int64_t)blockIdx.x * kDBlockRows + (int64_t)threadIdx.x * kDRowsPerThread;
  uint8_t kp[kDRowsPerThread];
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < kDRowsPerThread; ++j) {
    kp[j] = r0 + j < n ? keep[r0 + j] : 0;
    cnt += kp[j];
  }
  uint64_t total;
  uint64_t pos = blk_off[blockIdx.x] + block_exclusive_scan(cnt, smem, &total);
#pragma unroll
  for (int j = 0; j < kDRowsPerThread; ++j)
    if (kp[j]) map[pos++] = (int32_t)(r0 + j);
}

}  // namespace

static void run_distinct(const ldf_table_view* t, const int32_t* keys, int32_t nkeys, Alloc& a, ldf_table_out* out)
{
  hipStream_t s = a.stream();
  LDF_EXPECTS(t != nullptr && out != nullptr && (out->cols != nullptr || t->ncols == 0), "distinct: NULL argument");
  LDF_EXPECTS(nkeys >= 1, "distinct: at least one key column");
  LDF_EXPECTS(keys != nullptr, "distinct: NULL key list");
  LDF_EXPECTS(out->ncols >= t->ncols, "distinct: output table has too few column slots");
  const int64_t n = t->ncols > 0 ? t->cols[0].length : 0;
  for (int i = 0; i < t->ncols; ++i) validate_view(t->cols[i], n, "distinct");
  LDF_EXPECTS(n < INT32_MAX, "distinct: limited to 2^31-1 rows per call");
  // the key columns, each once (a repeated key column changes nothing)
  std::vector<DCol> kcols;
  std::vector<int32_t> seen;
  for (int i = 0; i < nkeys; ++i) {
    if (keys[i] < 0 || keys[i] >= t->ncols) throw Error(LDF_E_INVALID, "distinct: key index out of range");
    if (std::find(seen.begin(), seen.end(), keys[i]) != seen.end()) continue;
    seen.push_back(keys[i]);
    kcols.push_back(to_dcol(t->cols[keys[i]]));
  }
  Buf<DCol> kdev(a, kcols.size());
  for (size_t off = 0; off < kcols.size(); off += kDColChunk) {
    DColChunk ch;
    ch.n      = (int32_t)std::min<size_t>(kDColChunk, kcols.size() - off);
    ch.offset = (int32_t)off;
    for (int i = 0; i < ch.n; ++i) ch.c[i] = kcols[off + i];
    hipLaunchKernelGGL(k_put_dcols, dim3(1), dim3(kDColChunk), 0, s, ch, kdev.get());
    LDF_HIP_TRY(hipGetLastError());
  }
  const DKeysP k{kdev.get(), (int32_t)kcols.size()};

  int64_t n_out = 0;
  Buf<int32_t> map;
  if (n > 0) {
    // The set first gets n/4 slots (for up to ~n/8 keys at load <= 1/2; a quarter of the memory
    // traffic of the full size) with a probe limit per row; if any row hits the limit (more keys),
    // the whole pass runs again with 2n slots, which always hold every key.
    const uint64_t full_cap = next_pow2((uint64_t)(n > 32 ? 2 * n : 64));
    uint64_t cap            = n >= kDSmallFrom ? next_pow2((uint64_t)(n / 4)) : full_cap;
    if (cap < full_cap) {
      // estimated keys above n/5 (fewer than 2.5 S^2 / n repeats in the sample): the n/4-slot set
      // would overflow, so go to the full-size set at once
      const int64_t S = std::min<int64_t>(n, kDSample);
      Buf<unsigned long long> sset(a, (size_t)1 << kDSampleBits);
      Buf<unsigned long long> dups(a, 1);
      fill_u64(reinterpret_cast<uint64_t*>(sset.get()), 0, (int64_t)1 << kDSampleBits, s);
      fill_u64(reinterpret_cast<uint64_t*>(dups.get()), 0, 1, s);
      hipLaunchKernelGGL(k_distinct_sample, dim3(grid_for(S, kDThreads)), dim3(kDThreads), 0, s, k, S, n / S, sset.get(),
                         dups.get());
      LDF_HIP_TRY(hipGetLastError());
      unsigned long long d = 0;
      copy_to_host(&d, dups.get(), sizeof(d), s);
      if ((double)d < 2.5 * (double)S * (double)S / (double)n) cap = full_cap;
    }
    const int64_t nblk      = (n + kDBlockRows - 1) / kDBlockRows;
    Buf<uint8_t> keep(a, n);
    Buf<uint32_t> blk_cnt(a, nblk);
    Buf<uint64_t> blk_off(a, nblk + 1);
    Buf<int32_t> err(a, 1);
    while (true) {
      const bool last = cap >= full_cap;
      Buf<unsigned long long> set(a, cap);
      fill_u64(reinterpret_cast<uint64_t*>(set.get()), kDEmpty, (int64_t)cap, s);
      LDF_HIP_TRY(hipMemsetAsync(err.get(), 0, sizeof(int32_t), s));
      {
        ProfScope prof("distinct_insert", s);
        hipLaunchKernelGGL(k_distinct_insert, dim3(grid_for(n, kDThreads)), dim3(kDThreads), 0, s, k, n, set.get(),
                           cap - 1, last ? cap : (uint64_t)kDProbeLimit, err.get());
        LDF_HIP_TRY(hipGetLastError());
      }
      LDF_HIP_TRY(hipMemsetAsync(keep.get(), 0, (size_t)n, s));
      hipLaunchKernelGGL(k_distinct_mark, dim3(grid_for((int64_t)cap, kDThreads)), dim3(kDThreads), 0, s,
                         (const unsigned long long*)set.get(), cap, keep.get());
      LDF_HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(k_distinct_count, dim3((unsigned)nblk), dim3(kDThreads), 0, s, n, (const uint8_t*)keep.get(),
                         blk_cnt.get());
      LDF_HIP_TRY(hipGetLastError());
      exclusive_scan_u32_to_u64(blk_cnt.get(), blk_off.get(), nblk, a);
      uint64_t* st = static_cast<uint64_t*>(pinned_stage(16));
      st[0] = st[1] = 0;
      LDF_HIP_TRY(hipMemcpyAsync(st, blk_off.get() + nblk, 8, hipMemcpyDeviceToHost, s));
      LDF_HIP_TRY(hipMemcpyAsync(st + 1, err.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
      LDF_HIP_TRY(hipStreamSynchronize(s));
      n_out             = (int64_t)st[0];
      const bool failed = st[1] != 0;
      if (!failed) break;
      if (last) throw Error(LDF_E_INTERNAL, "distinct: hash set overflow");
      cap = full_cap;
    }
    map = Buf<int32_t>(a, n_out > 0 ? n_out : 1);
    if (n_out > 0) {
      hipLaunchKernelGGL(k_distinct_emit, dim3((unsigned)nblk), dim3(kDThreads), 0, s, n, (const uint8_t*)keep.get(),
                         (const uint64_t*)blk_off.get(), map.get());
      LDF_HIP_TRY(hipGetLastError());
    }
  } else {
    map = Buf<int32_t>(a, 1);
  }
  // every column, gathered at the kept rows (in row order, so the reads are nearly sequential)
  out->ncols = t->ncols;
  std::vector<const ldf_column_view*> src(t->ncols);
  for (int i = 0; i < t->ncols; ++i) src[i] = &t->cols[i];
  gather_columns(src.data(), t->ncols, map.get(), n_out, false, a, out->cols);
}

}  // namespace ldf

using namespace ldf;

extern "C" int ldf_distinct(const ldf_table_view* table, const int32_t* keys, int32_t nkeys,
                            const ldf_allocator* alloc, void* stream, ldf_table_out* out)
{
  return guarded([&] {
    Alloc a(alloc, (hipStream_t)stream);
    run_distinct(table, keys, nkeys, a, out);
  });
}
