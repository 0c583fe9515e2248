// Groupby-aggregate: SUM, PRODUCT, MIN, MAX, COUNT, MEAN, VARIANCE, STDDEV, COUNT_DISTINCT,
// COUNT_ALL over one or more key columns.
//
// Reference: GroupByAggregationTask::gpu_variant (cpp/src/groupby_aggregation.cu:90-167) runs
// cudf::groupby(keys, null_policy::INCLUDE).aggregate and casts to the Arrow result types that
// cpp/src/groupby_aggregation.cpp:100-132 derives from Arrow's hash_<kind> kernels; the CPU
// variant (cpp/src/groupby_aggregation.cpp:41-94) is Arrow Acero's hash aggregate.  Semantics
// kept here are Arrow's: a null key is a group of its own; SUM/MEAN/MIN/MAX/PRODUCT/VARIANCE
// of a group without valid values is null; MIN/MAX ignore NaN unless every valid value is
// NaN; COUNT / COUNT_DISTINCT honour CountOptions modes; integer SUM wraps (order-free).
//
// Structure:
//   accumulators  the aggregations are lowered to deduplicated accumulators per input column
//                 (row count, valid count, wrapping integer sum, exact FP sum, min, max,
//                 product, distinct count): SUM and MEAN of a column share one exact sum, and
//                 every aggregation of a column shares its valid count.
//   key table     open addressing, 8-byte slots installed by 64-bit atomicCAS.  A single
//                 fixed-width key stores its bits (two extra slots: the key equal to the empty
//                 marker, and the null key); composite / string keys store
//                 (fingerprint32 << 32 | representative row).  Global accumulator arrays are
//                 indexed by slot.
//   hot keys      a 65536-row sample gives both the table size (GEE distinct estimate) and the
//                 most frequent keys.  Every block of k_gb_main preloads those keys into an LDS
//                 hash table with LDS accumulators, aggregates its rows of hot keys with LDS
//                 atomics, and flushes once at the end; rows of other keys update the global
//                 arrays directly.  Skewed keys (Zipf) therefore never serialise on global
//                 atomics.
//   exact FP sum  x * 2^96 accumulated exactly in a 192-bit two's-complement integer over three
//                 64-bit words; the carry out of each word is read from the old value the
//                 atomic add returns, so the sum is exact and independent of the order of the
//                 atomics, and the final double is the correctly rounded exact sum.  Values with
//                 |x| >= 2^62 or non-finite values divert that aggregation to an FP64 atomic
//                 fallback pass.  Bits below 2^-96 are truncated (|error| <= n * 2^-96).
//   distinct      COUNT_DISTINCT inserts (slot, value) into an open-addressing set; a
//                 successful CAS counts a new distinct value for the group.
// If the key table still fills beyond 3/4, the call is re-run with four times the capacity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>

#include "ldf_agg.hpp"
#include "ldf_common.hpp"
#include "ldf_device.hpp"
#include "ldf_rec.hpp"

namespace ldf {

constexpr unsigned long long kGEmpty = ~0ull;
constexpr int kGbThreads             = 1024;
constexpr int kGbRowsPerThread       = 4;
constexpr int kMaxAcc                = 16;
constexpr int kMaxCols               = 4;  // distinct value columns read per launch
constexpr int kMaxOut                = 16;
constexpr int kSampleRows            = 65536;

enum AccKind : int32_t {
  ACC_ROWS     = 0,
  ACC_VCOUNT   = 1,
  ACC_ISUM     = 2,
  ACC_FSUM     = 3,
  ACC_MIN      = 4,
  ACC_MAX      = 5,
  ACC_PROD     = 6,
  ACC_DISTINCT = 7
};

struct AccDev {
  int32_t kind;
  int32_t col;       // index into GbPlan::col (-1 for ACC_ROWS)
  int32_t lds_off;   // byte offset of this accumulator's LDS array, -1 when global only
  int32_t flag_bit;  // MIN/MAX: bit in the NaN-flag word
  unsigned long long* g0;  // count / sum / min / max / product / limb 0
  unsigned long long* g1;  // limb 1
  unsigned long long* g2;  // limb 2
  uint32_t* gflag;         // MIN/MAX: NaN seen
  unsigned long long* pset;
  uint64_t pmask;
  int32_t* err;  // FSUM: a value left the fixed-point range
  unsigned long long* pset2;  // DISTINCT pairs of keys outside the hot set (partitioned path)
  uint64_t pmask2;
};

struct ColDev {
  const void* data;
  const uint8_t* valid;
  const int64_t* offsets;
  int32_t type;
  int32_t width;
};

// Where the per-slot LDS accumulators of a table of `cap` slots live (byte offsets).
struct LdsLayout {
  int32_t rows_off;           // u32 rows [cap]
  int32_t flags_off;          // u32 NaN flags [cap]
  int32_t acc_off[kMaxAcc];   // per accumulator, -1 = none
};

struct GbPlan {
  AccDev a[kMaxAcc];
  ColDev col[kMaxCols];
  int32_t nacc;
  int32_t ncol;
  int32_t hot_cap;     // LDS hot slots (power of two); 0 = no LDS table
  int32_t tab_off;     // LDS: u64 keys [2*hot_cap]
  int32_t tidx_off;    // LDS: u16 hot index [2*hot_cap]
  int32_t flags_off;   // LDS: u32 NaN flags [hot_cap]
  int32_t rows_off;    // LDS: u32 rows [hot_cap]
  int32_t lds_bytes;
  int64_t rows_per_block;
  int32_t dpart;  // COUNT_DISTINCT through pair partitions (partitioned path, one such accumulator)
  int32_t dcol;   // its value column (index into col)
  int32_t part_bits;  // partitioned path: 2^part_bits partitions
  int32_t acc_end;    // LDS bytes of the hot accumulators (the hot key table follows them)
  int32_t stage;      // k_gb_main3 packs partition records through LDS (0: in registers)
  // per value column, the index in `a` of its VCOUNT, ISUM, FSUM, MIN and MAX accumulator (-1:
  // none; the lowering keeps at most one accumulator per kind and column)
  int8_t cacc[kMaxCols][5];
};

struct HotList {
  unsigned long long* keys;  // key bits
  int32_t* special;          // 1 = key equal to the empty marker, 2 = null key
  int32_t* gslot;            // global slot
  unsigned int* count;
};

struct KeyTab {
  unsigned long long* slots;
  uint64_t mask;
  uint64_t cap;
  unsigned long long* nkeys;  // distinct keys inserted
  int32_t* overflow;
};

// ---- key table --------------------------------------------------------------------------
// Probes an insertion may take before it gives up and flags an overflow.  A table filled to its
// 3/4 limit with well-mixed hashes essentially never has a run this long, while a table that
// has filled up (the distinct-count estimate was low; the call is then run again with four times
// the capacity) would otherwise cost every further new key a scan of the whole table.
__device__ __forceinline__ uint64_t probe_limit(const KeyTab& t) { return t.mask < 4095 ? t.mask + 1 : 4096; }

// slot of a single fixed-width key given by its bits (kv = the key is valid)
// COUNT = false: the table's key count is not kept (a table that holds at most half its slots,
// such as the sample's, cannot pass the 3/4 load check; a single counter that every new key
// increments serialises the inserts)
template <bool COUNT = true>
__device__ __forceinline__ int64_t gb_insert_bits(unsigned long long key, bool kv, const KeyTab& t)
{
  if (!kv) {
    if (t.slots[t.cap + 1] == kGEmpty) t.slots[t.cap + 1] = 0;  // mark used (idempotent)
    return (int64_t)t.cap + 1;
  }
  if (key == kGEmpty) {
    if (t.slots[t.cap] == kGEmpty) t.slots[t.cap] = 0;
    return (int64_t)t.cap;
  }
  uint64_t i = mix64(key) & t.mask;
  for (uint64_t probes = 0, lim = probe_limit(t); probes < lim; ++probes) {
    unsigned long long cur = atomic_load_relaxed(&t.slots[i]);
    if (cur == key) return (int64_t)i;
    if (cur == kGEmpty) {
      unsigned long long old = atomicCAS(&t.slots[i], kGEmpty, key);
      if (old == kGEmpty) {
        if (COUNT) {
          unsigned long long nk = atomicAdd(t.nkeys, 1ull);
          if (nk + 1 > (t.cap / 4) * 3) *t.overflow = 1;
        }
        return (int64_t)i;
      }
      if (old == key) return (int64_t)i;
    }
    i = (i + 1) & t.mask;
  }
  *t.overflow = 1;
  return -1;
}

template <bool GENERIC>
__device__ __forceinline__ int64_t gb_find_or_insert(const DKeys& k, int64_t r, const KeyTab& t)
{
  if (!GENERIC) {
    const DCol& c = k.c[0];
    if (!valid_at(c.valid, r)) {
      if (t.slots[t.cap + 1] == kGEmpty) t.slots[t.cap + 1] = 0;  // mark used (idempotent)
      return (int64_t)t.cap + 1;
    }
    unsigned long long key = load_bits(c.data, c.width, r);
    if (key == kGEmpty) {
      if (t.slots[t.cap] == kGEmpty) t.slots[t.cap] = 0;
      return (int64_t)t.cap;
    }
    uint64_t i = mix64(key) & t.mask;
    for (uint64_t probes = 0, lim = probe_limit(t); probes < lim; ++probes) {
      unsigned long long cur = atomic_load_relaxed(&t.slots[i]);
      if (cur == key) return (int64_t)i;
      if (cur == kGEmpty) {
        unsigned long long old = atomicCAS(&t.slots[i], kGEmpty, key);
        if (old == kGEmpty) {
          unsigned long long nk = atomicAdd(t.nkeys, 1ull);
          if (nk + 1 > (t.cap / 4) * 3) *t.overflow = 1;
          return (int64_t)i;
        }
        if (old == key) return (int64_t)i;
      }
      i = (i + 1) & t.mask;
    }
    *t.overflow = 1;
    return -1;
  } else {
    uint64_t h                = row_hash64(k, r);
    uint32_t fp               = (uint32_t)(h >> 32);
    unsigned long long packed = ((unsigned long long)fp << 32) | (uint32_t)r;
    uint64_t i                = h & t.mask;
    for (uint64_t probes = 0, lim = probe_limit(t); probes < lim; ++probes) {
      unsigned long long cur = atomic_load_relaxed(&t.slots[i]);
      if (cur == kGEmpty) {
        unsigned long long old = atomicCAS(&t.slots[i], kGEmpty, packed);
        if (old == kGEmpty) {
          unsigned long long nk = atomicAdd(t.nkeys, 1ull);
          if (nk + 1 > (t.cap / 4) * 3) *t.overflow = 1;
          return (int64_t)i;
        }
        cur = old;
      }
      if ((uint32_t)(cur >> 32) == fp && rows_equal(k, r, k, (int64_t)(uint32_t)cur)) return (int64_t)i;
      i = (i + 1) & t.mask;
    }
    *t.overflow = 1;
    return -1;
  }
}

template <bool GENERIC>
__device__ __forceinline__ int64_t gb_find(const DKeys& k, int64_t r, const KeyTab& t)
{
  if (!GENERIC) {
    const DCol& c = k.c[0];
    if (!valid_at(c.valid, r)) return (int64_t)t.cap + 1;
    unsigned long long key = load_bits(c.data, c.width, r);
    if (key == kGEmpty) return (int64_t)t.cap;
    uint64_t i = mix64(key) & t.mask;
    while (true) {
      unsigned long long cur = t.slots[i];
      if (cur == key) return (int64_t)i;
      if (cur == kGEmpty) return -1;
      i = (i + 1) & t.mask;
    }
  } else {
    uint64_t h  = row_hash64(k, r);
    uint32_t fp = (uint32_t)(h >> 32);
    uint64_t i  = h & t.mask;
    while (true) {
      unsigned long long cur = t.slots[i];
      if (cur == kGEmpty) return -1;
      if ((uint32_t)(cur >> 32) == fp && rows_equal(k, r, k, (int64_t)(uint32_t)cur)) return (int64_t)i;
      i = (i + 1) & t.mask;
    }
  }
}

__device__ __forceinline__ void atomic_mul_u64(unsigned long long* p, uint64_t v)
{
  unsigned long long old = *p, assumed;
  do {
    assumed = old;
    old     = atomicCAS(p, assumed, (unsigned long long)(assumed * v));
  } while (old != assumed);
}

__device__ __forceinline__ void atomic_mul_f64(unsigned long long* p, double v)
{
  unsigned long long old = *p, assumed;
  do {
    assumed  = old;
    double r = __longlong_as_double((long long)assumed) * v;
    old      = atomicCAS(p, assumed, (unsigned long long)__double_as_longlong(r));
  } while (old != assumed);
}

// Inserts (slot, value) into the distinct set; true when the pair is new.
__device__ inline bool distinct_insert(const AccDev& a, const ColDev& c, const DKeys& keys, int64_t r, int64_t slot,
                                       uint64_t vbits)
{
  if (c.width > 0 && c.width <= 4) {
    unsigned long long pk = ((unsigned long long)(uint32_t)slot << 32) | (uint32_t)vbits;
    uint64_t i            = mix64(pk) & a.pmask;
    while (true) {  // most pairs are new: try the CAS first, one round trip
      unsigned long long old = atomicCAS(&a.pset[i], kGEmpty, pk);
      if (old == kGEmpty) return true;
      if (old == pk) return false;
      i = (i + 1) & a.pmask;
    }
  }
  // (fingerprint, representative row): equality re-checks the key tuple and the value
  uint64_t vh;
  if (c.width == 0) {
    int64_t b = c.offsets[r], e = c.offsets[r + 1];
    vh        = string_hash64(static_cast<const uint8_t*>(c.data) + b, e - b);
  } else {
    vh = mix64(vbits ^ 0x51afd7ed558ccdffull);
  }
  uint64_t h                = mix64(vh ^ mix64((uint64_t)slot + 0x9e3779b97f4a7c15ull));
  uint32_t fp               = (uint32_t)(h >> 32);
  unsigned long long packed = ((unsigned long long)fp << 32) | (uint32_t)r;
  uint64_t i                = h & a.pmask;
  while (true) {
    unsigned long long cur = atomic_load_relaxed(&a.pset[i]);
    if (cur == kGEmpty) {
      unsigned long long old = atomicCAS(&a.pset[i], kGEmpty, packed);
      if (old == kGEmpty) return true;
      cur = old;
    }
    if ((uint32_t)(cur >> 32) == fp) {
      int64_t o = (int64_t)(uint32_t)cur;
      bool same = rows_equal(keys, r, keys, o);
      if (same) {
        if (c.width == 0) {
          DCol d;
          d.data    = c.data;
          d.offsets = c.offsets;
          d.valid   = nullptr;
          d.width   = 0;
          same      = strings_equal(d, r, d, o);
        } else {
          same = load_bits(c.data, c.width, o) == vbits;
        }
      }
      if (same) return false;
    }
    i = (i + 1) & a.pmask;
  }
}

// ---- per-row accumulation ---------------------------------------------------------------
// Values of one row for the plan's (up to 4) input columns, passed by value so they stay in
// registers.
struct RowVals {
  uint64_t b0, b1, b2, b3;
  uint32_t valid;  // bit c = column c valid
};

__device__ __forceinline__ uint64_t pick(const RowVals v, int c)
{
  uint64_t x = v.b0;
  x          = c == 1 ? v.b1 : x;
  x          = c == 2 ? v.b2 : x;
  x          = c == 3 ? v.b3 : x;
  return x;
}

// global accumulators at slot s
__device__ __forceinline__ void acc_global(const GbPlan& p, const DKeys& keys, int64_t r, int64_t s, const RowVals v)
{
  for (int i = 0; i < p.nacc; ++i) {
    const AccDev& a = p.a[i];
    const bool ok   = a.col < 0 || ((v.valid >> a.col) & 1);
    if (a.kind == ACC_ROWS) {
      if (a.g0) atomicAdd(&a.g0[s], 1ull);
      continue;
    }
    if (!ok) continue;
    const ColDev& c   = p.col[a.col];
    const uint64_t vb = pick(v, a.col);
    switch (a.kind) {
      case ACC_VCOUNT: atomicAdd(&a.g0[s], 1ull); break;
      case ACC_ISUM: atomicAdd(&a.g0[s], (unsigned long long)bits_as_int(vb, c.type, c.width)); break;
      case ACC_FSUM: {
        uint64_t l0, l1, l2;
        if (to_fixed192(bits_as_double(vb, c.type, c.width), l0, l1, l2)) add_fixed192(&a.g0[s], &a.g1[s], &a.g2[s], l0, l1, l2);
        else *a.err = 1;
        break;
      }
      case ACC_MIN:
      case ACC_MAX: {
        if (is_float_type(c.type) && isnan(bits_as_double(vb, c.type, c.width))) {
          atomicOr(&a.gflag[s], 1u);
          break;
        }
        uint64_t o = ordered_of(vb, c.type, c.width);
        if (a.kind == ACC_MIN) atomicMin(&a.g0[s], (unsigned long long)o);
        else atomicMax(&a.g0[s], (unsigned long long)o);
        break;
      }
      case ACC_PROD:
        if (is_float_type(c.type)) atomic_mul_f64(&a.g0[s], bits_as_double(vb, c.type, c.width));
        else atomic_mul_u64(&a.g0[s], bits_as_int(vb, c.type, c.width));
        break;
      case ACC_DISTINCT:
        if (distinct_insert(a, c, keys, r, s, vb)) atomicAdd(&a.g0[s], 1ull);
        break;
      default: break;
    }
  }
}

// LDS accumulators of hot slot h (global slot s for the global-only accumulators)
__device__ __forceinline__ void acc_lds(const GbPlan& p, unsigned char* lds, const DKeys& keys, int64_t r, int h,
                                        int64_t s, const RowVals v)
{
  uint32_t* rows = reinterpret_cast<uint32_t*>(lds + p.rows_off);
  // DISTINCT pairs are keyed by the global slot here
  atomicAdd(&rows[h], 1u);
  for (int i = 0; i < p.nacc; ++i) {
    const AccDev& a = p.a[i];
    if (a.kind == ACC_ROWS) continue;
    const bool ok = (v.valid >> a.col) & 1;
    if (!ok) continue;
    const ColDev& c   = p.col[a.col];
    const uint64_t vb = pick(v, a.col);
    switch (a.kind) {
      case ACC_VCOUNT: atomicAdd(&reinterpret_cast<uint32_t*>(lds + a.lds_off)[h], 1u); break;
      case ACC_ISUM:
        atomicAdd(&reinterpret_cast<unsigned long long*>(lds + a.lds_off)[h],
                  (unsigned long long)bits_as_int(vb, c.type, c.width));
        break;
      case ACC_FSUM: {
        unsigned long long* base = reinterpret_cast<unsigned long long*>(lds + a.lds_off);
        if (!lds_fsum_add(&base[6 * h], bits_as_double(vb, c.type, c.width))) *a.err = 1;
        break;
      }
      case ACC_MIN:
      case ACC_MAX: {
        if (is_float_type(c.type) && isnan(bits_as_double(vb, c.type, c.width))) {
          atomicOr(&reinterpret_cast<uint32_t*>(lds + p.flags_off)[h], 1u << a.flag_bit);
          break;
        }
        uint64_t o               = ordered_of(vb, c.type, c.width);
        unsigned long long* slot = &reinterpret_cast<unsigned long long*>(lds + a.lds_off)[h];
        if (a.kind == ACC_MIN) atomicMin(slot, (unsigned long long)o);
        else atomicMax(slot, (unsigned long long)o);
        break;
      }
      case ACC_PROD:
        if (is_float_type(c.type)) atomic_mul_f64(&a.g0[s], bits_as_double(vb, c.type, c.width));
        else atomic_mul_u64(&a.g0[s], bits_as_int(vb, c.type, c.width));
        break;
      case ACC_DISTINCT:
        if (distinct_insert(a, c, keys, r, s, vb)) atomicAdd(&reinterpret_cast<uint32_t*>(lds + a.lds_off)[h], 1u);
        break;
      default: break;
    }
  }
}

template <bool GENERIC>
__global__ __launch_bounds__(kGbThreads) void k_gb_main(DKeys keys, int64_t n, KeyTab t, GbPlan p, HotList hot)
{
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int s_null_hot, s_sent_hot;
  const int H = p.hot_cap;
  int nhot    = 0;
  if (!GENERIC && H > 0) {
    nhot = (int)min(*hot.count, (unsigned int)H);
    const int T                 = 2 * H;
    unsigned long long* tkeys   = reinterpret_cast<unsigned long long*>(lds + p.tab_off);
    uint16_t* tidx              = reinterpret_cast<uint16_t*>(lds + p.tidx_off);
    for (int i = threadIdx.x; i < T; i += blockDim.x) tkeys[i] = kGEmpty;
    // accumulator arrays: zero, MIN = all ones
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
      reinterpret_cast<uint32_t*>(lds + p.rows_off)[i]  = 0;
      reinterpret_cast<uint32_t*>(lds + p.flags_off)[i] = 0;
    }
    for (int j = 0; j < p.nacc; ++j) {
      const AccDev& a = p.a[j];
      if (a.lds_off < 0) continue;
      if (a.kind == ACC_VCOUNT || a.kind == ACC_DISTINCT) {
        for (int i = threadIdx.x; i < H; i += blockDim.x) reinterpret_cast<uint32_t*>(lds + a.lds_off)[i] = 0;
      } else {
        const int words          = a.kind == ACC_FSUM ? 6 * H : H;
        const unsigned long long v = a.kind == ACC_MIN ? ~0ull : 0ull;
        for (int i = threadIdx.x; i < words; i += blockDim.x) reinterpret_cast<unsigned long long*>(lds + a.lds_off)[i] = v;
      }
    }
    if (threadIdx.x == 0) {
      s_null_hot = -1;
      s_sent_hot = -1;
    }
    __syncthreads();
    for (int h = threadIdx.x; h < nhot; h += blockDim.x) {
      int sp = hot.special[h];
      if (sp == 1) s_sent_hot = h;
      else if (sp == 2) s_null_hot = h;
      else {
        unsigned long long k = hot.keys[h];
        uint32_t i           = (uint32_t)(mix64(k) & (uint64_t)(T - 1));
        while (true) {
          unsigned long long old = atomicCAS(&tkeys[i], kGEmpty, k);
          if (old == kGEmpty) {
            tidx[i] = (uint16_t)h;
            break;
          }
          i = (i + 1) & (T - 1);
        }
      }
    }
    __syncthreads();
  }

  const int64_t begin = (int64_t)blockIdx.x * p.rows_per_block;
  const int64_t end   = begin + p.rows_per_block < n ? begin + p.rows_per_block : n;
  const DCol& kc      = keys.c[0];
  for (int64_t tile = begin; tile < end; tile += (int64_t)kGbRowsPerThread * blockDim.x) {
    RowVals v[kGbRowsPerThread];
    uint64_t kb[kGbRowsPerThread];
    bool kv[kGbRowsPerThread];
#pragma unroll
    for (int j = 0; j < kGbRowsPerThread; ++j) {
      const int64_t r = tile + (int64_t)j * blockDim.x + threadIdx.x;
      v[j]  = RowVals{0, 0, 0, 0, 0};
      kb[j] = 0;
      kv[j] = false;
      if (r < end) {
        if (!GENERIC) {
          kv[j] = valid_at(kc.valid, r);
          kb[j] = load_bits(kc.data, kc.width, r);
        }
        uint32_t vm = 0;
        uint64_t x[kMaxCols] = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < kMaxCols; ++c) {
          if (c < p.ncol) {
            const ColDev& cd = p.col[c];
            if (valid_at(cd.valid, r)) vm |= 1u << c;
            x[c] = cd.width > 0 ? load_bits(cd.data, cd.width, r) : 0;
          }
        }
        v[j] = RowVals{x[0], x[1], x[2], x[3], vm};
      }
    }
#pragma unroll
    for (int j = 0; j < kGbRowsPerThread; ++j) {
      const int64_t r = tile + (int64_t)j * blockDim.x + threadIdx.x;
      if (r >= end) continue;
      int h = -1;
      if (!GENERIC && nhot > 0) {
        if (!kv[j]) h = s_null_hot;
        else if (kb[j] == kGEmpty) h = s_sent_hot;
        else {
          const int T                     = 2 * H;
          const unsigned long long* tkeys = reinterpret_cast<const unsigned long long*>(lds + p.tab_off);
          uint32_t i                      = (uint32_t)(mix64(kb[j]) & (uint64_t)(T - 1));
          while (true) {
            unsigned long long cur = tkeys[i];
            if (cur == kb[j]) {
              h = reinterpret_cast<const uint16_t*>(lds + p.tidx_off)[i];
              break;
            }
            if (cur == kGEmpty) break;
            i = (i + 1) & (T - 1);
          }
        }
      }
      if (h >= 0) {
        acc_lds(p, lds, keys, r, h, hot.gslot[h], v[j]);
      } else {
        int64_t s = gb_find_or_insert<GENERIC>(keys, r, t);
        if (s >= 0) acc_global(p, keys, r, s, v[j]);
      }
    }
  }

  if (!GENERIC && H > 0) {
    __syncthreads();
    const uint32_t* rows = reinterpret_cast<const uint32_t*>(lds + p.rows_off);
    const uint32_t* flg  = reinterpret_cast<const uint32_t*>(lds + p.flags_off);
    for (int h = threadIdx.x; h < nhot; h += blockDim.x) {
      const uint32_t nr = rows[h];
      if (nr == 0) continue;
      const int64_t s = hot.gslot[h];
      for (int j = 0; j < p.nacc; ++j) {
        const AccDev& a = p.a[j];
        switch (a.kind) {
          case ACC_ROWS:
            if (a.g0) atomicAdd(&a.g0[s], (unsigned long long)nr);
            break;
          case ACC_VCOUNT:
          case ACC_DISTINCT: {
            uint32_t c = reinterpret_cast<const uint32_t*>(lds + a.lds_off)[h];
            if (c) atomicAdd(&a.g0[s], (unsigned long long)c);
            break;
          }
          case ACC_ISUM: {
            unsigned long long x = reinterpret_cast<const unsigned long long*>(lds + a.lds_off)[h];
            if (x) atomicAdd(&a.g0[s], x);
            break;
          }
          case ACC_FSUM: {
            uint64_t b0, b1, b2;
            lds_fsum_value(reinterpret_cast<const unsigned long long*>(lds + a.lds_off) + 6 * h, b0, b1, b2);
            if (b0 | b1 | b2) add_fixed192(&a.g0[s], &a.g1[s], &a.g2[s], b0, b1, b2);
            break;
          }
          case ACC_MIN:
          case ACC_MAX: {
            unsigned long long x = reinterpret_cast<const unsigned long long*>(lds + a.lds_off)[h];
            if (a.kind == ACC_MIN && x != ~0ull) atomicMin(&a.g0[s], x);
            if (a.kind == ACC_MAX && x != 0ull) atomicMax(&a.g0[s], x);
            if ((flg[h] >> a.flag_bit) & 1) atomicOr(&a.gflag[s], 1u);
            break;
          }
          default: break;
        }
      }
    }
  }
}

// ---- partitioned path for keys outside the hot set ---------------------------------------
// Rows whose key is not hot are not aggregated with global atomics: the main pass scatters them
// (key, value bits, a flags byte) into 2^part_bits hash partitions, LDS-staged so each partition
// receives contiguous runs, and k_gb_part3 aggregates every (partition, sub-pass) in one block's
// LDS table.  Each key therefore lives in exactly one place -- a hot LDS slot, a partition
// block's LDS slot, or (only when a partition block's table is full) the global table -- and
// the final values are stored, not added, into the global slot arrays.
constexpr int kMain3Lds    = 160 * 1024 - 256;  // k_gb_main3 dynamic LDS limit (a CU holds 160 KB; a little for its static LDS)
constexpr int kMinPartBits = 10;
constexpr int kMaxPartBits = 12;
constexpr int kSub3     = kGbThreads * kGbRowsPerThread;  // rows staged per step
constexpr int kFlagKeyValid = 1 << 4;
constexpr int kFlagDistinct = 5;  // bits 5..7: the row's (key, value) pair was new, per DISTINCT
constexpr int kPartProbes   = 64;

// A partition row is one record of `stride` bytes (8, 16, 32 or 64): the key and value bits,
// each field at a multiple of its width (widest first), then the flags byte (bits 0-3 value
// valid, bit 4 key valid, bits 5-7 new distinct pair).  One aligned record per row keeps the
// scattered stores of the main pass to whole 16-byte (or 8-byte) pieces of one 32-byte segment.
struct PartBufs {
  unsigned char* rec;
  const uint64_t* offsets;  // exclusive scan of counts[part][blk]
  int32_t stride;
  int32_t key_off;
  int32_t val_off[kMaxCols];
  int32_t flags_off;
  int32_t key_width;
  int32_t nblk;  // histogram blocks: the stride of offsets / hoffsets
  // dpart: the hot rows' pairs (hot index << 32 | value) in row order, as the image under
  // pair_mix: its low 32 bits (hwords) and its pair partition, bits 32..41 (hparts)
  uint32_t* hwords;
  uint16_t* hparts;
  const uint64_t* hoffsets;    // hot pairs before each histogram block (exclusive scan)
  uint32_t* pcounts;           // k_gb_main3's hot pairs per (pair partition, block)
  // dpart with one sub-pass: k_gb_part3 appends each partition's (key << 32 | value) pairs of rows
  // with a valid value at the start of the partition's range, and their number per partition
  unsigned long long* cpairs;
  uint32_t* cpair_count;
};


struct HotParts {  // per-block partials of the hot slots, index blk * H + h (FSUM: 3 words)
  uint32_t* rows;
  uint32_t* flags;
  unsigned long long* acc[kMaxAcc];
};

struct PartPlan {
  int32_t tcap;       // LDS table slots (power of two)
  int32_t S;          // sub-passes per partition
  int32_t key_off;    // u64 keys [tcap]
  int32_t rows_off;   // u32 [tcap]
  int32_t flags_off;  // u32 [tcap]
  int32_t acc_off[kMaxAcc];
  int32_t lds_bytes;
};

// partition of a key from its hash hk = mix64(key): the top `bits` bits (the LDS tables index by
// low bits of the same hash)
__device__ __forceinline__ uint32_t cold_part(uint64_t hk, int bits) { return (uint32_t)(hk >> (64 - bits)); }
// sub-pass of a key in [0, S) from bits 22..53 of the same hash (multiply-high range reduction)
__device__ __forceinline__ uint32_t cold_sub(uint64_t hk, int S)
{
  return (uint32_t)((((hk >> 22) & 0xffffffffull) * (uint64_t)S) >> 32);
}

// COUNT_DISTINCT through pair partitions: the (hot index, value) pairs of hot rows are spread by
// a hash of the pair over kHPairParts coarse partitions (then fine ones), so each fine
// partition's pairs fit an LDS hash set.
constexpr int kHPairBits = 10;
constexpr int kHPairParts = 1 << kHPairBits;
constexpr int kPairSet    = 8192;  // LDS pair-set slots of the dedupe kernels (64 KB)
__device__ __forceinline__ uint64_t pair_hash(unsigned long long pr) { return mix64(pr ^ 0x94d049bb133111ebull); }

// A hot row's pair p = (hot index << 32 | value) has at most 42 bits (hot index < 2^10).  The
// hot pairs go to their coarse partition by the top 10 bits of a bijection of the 42-bit pair
// space (xor-shifts by half the width and multiplications by odd constants, each invertible), and
// a partition stores only the low 32 bits of the image: the pair is recovered from the partition
// index and those bits.  Inside a coarse partition the top kHFineBits of the stored word pick the
// fine partition and its low bits the LDS set slot.
constexpr int kPairHotMax = 1 << (42 - 32);  // hot keys of a groupby with a COUNT_DISTINCT
static_assert(kHPairBits == 42 - 32, "coarse partition bits + stored bits = pair bits");
constexpr uint64_t kP42 = (1ull << 42) - 1;
constexpr uint64_t kPM1 = 0x9E3779B97F4A7C15ull & kP42;  // odd
constexpr uint64_t kPM2 = 0xBF58476D1CE4E5B9ull & kP42;  // odd
constexpr uint64_t inv_odd(uint64_t c)  // c^-1 mod 2^64 by Newton's iteration (c odd)
{
  uint64_t x = c;
  for (int i = 0; i < 6; ++i) x *= 2 - c * x;
  return x;
}
constexpr uint64_t kPM1i = inv_odd(kPM1) & kP42;
constexpr uint64_t kPM2i = inv_odd(kPM2) & kP42;
static_assert(((kPM1 * kPM1i) & kP42) == 1 && ((kPM2 * kPM2i) & kP42) == 1, "inverse multipliers");
__host__ __device__ __forceinline__ uint64_t pair_mix(uint64_t p)
{
  p ^= p >> 21;
  p = (p * kPM1) & kP42;
  p ^= p >> 21;
  p = (p * kPM2) & kP42;
  return p ^ (p >> 21);
}
__host__ __device__ __forceinline__ uint64_t pair_unmix(uint64_t x)
{
  x ^= x >> 21;
  x = (x * kPM2i) & kP42;
  x ^= x >> 21;
  x = (x * kPM1i) & kP42;
  return x ^ (x >> 21);
}
constexpr int kHFineBits   = 5;
constexpr int kHFine       = 1 << kHFineBits;  // fine partitions of a coarse hot-pair partition
constexpr int kPairSet32   = 8192;             // LDS set slots of the hot dedupe (32 KB of words)
constexpr uint32_t kWEmpty = 0xffffffffu;      // a set word holds the stored word without its fine bits
constexpr uint32_t kWMask  = (1u << (32 - kHFineBits)) - 1;

// Builds the LDS table of hot keys when `table` and clears the hot accumulators when `accs`.
__device__ inline void hot_setup(const GbPlan& p, unsigned char* lds, const HotList& hot, int nhot, bool accs,
                                 bool table, int* s_null_hot, int* s_sent_hot)
{
  const int H               = p.hot_cap;
  const int T               = 2 * H;
  unsigned long long* tkeys = reinterpret_cast<unsigned long long*>(lds + p.tab_off);
  uint16_t* tidx            = reinterpret_cast<uint16_t*>(lds + p.tidx_off);
  if (table)
    for (int i = threadIdx.x; i < T; i += blockDim.x) tkeys[i] = kGEmpty;
  if (accs) {
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
      reinterpret_cast<uint32_t*>(lds + p.rows_off)[i]  = 0;
      reinterpret_cast<uint32_t*>(lds + p.flags_off)[i] = 0;
    }
    for (int j = 0; j < p.nacc; ++j) {
      const AccDev& a = p.a[j];
      if (a.lds_off < 0) continue;
      if (a.kind == ACC_VCOUNT || a.kind == ACC_DISTINCT) {
        for (int i = threadIdx.x; i < H; i += blockDim.x) reinterpret_cast<uint32_t*>(lds + a.lds_off)[i] = 0;
      } else {
        const int words            = a.kind == ACC_FSUM ? 6 * H : H;
        const unsigned long long v = a.kind == ACC_MIN ? ~0ull : 0ull;
        for (int i = threadIdx.x; i < words; i += blockDim.x) reinterpret_cast<unsigned long long*>(lds + a.lds_off)[i] = v;
      }
    }
  }
  if (threadIdx.x == 0) {
    *s_null_hot = -1;
    *s_sent_hot = -1;
  }
  __syncthreads();
  for (int h = threadIdx.x; h < nhot && table; h += blockDim.x) {
    int sp = hot.special[h];
    if (sp == 1) *s_sent_hot = h;
    else if (sp == 2) *s_null_hot = h;
    else {
      unsigned long long k = hot.keys[h];
      uint32_t i           = (uint32_t)(mix64(k) & (uint64_t)(T - 1));
      while (true) {
        unsigned long long old = atomicCAS(&tkeys[i], kGEmpty, k);
        if (old == kGEmpty) {
          tidx[i] = (uint16_t)h;
          break;
        }
        i = (i + 1) & (T - 1);
      }
    }
  }
  __syncthreads();
}

// hot index of key k (hash hk = mix64(k)), -1 when k is not hot
__device__ __forceinline__ int hot_lookup(const GbPlan& p, const unsigned char* lds, uint64_t k, uint64_t hk, bool kv,
                                          int null_hot, int sent_hot)
{
  if (!kv) return null_hot;
  if (k == kGEmpty) return sent_hot;
  const int T                     = 2 * p.hot_cap;
  const unsigned long long* tkeys = reinterpret_cast<const unsigned long long*>(lds + p.tab_off);
  uint32_t i                      = (uint32_t)(hk & (uint64_t)(T - 1));
  while (true) {
    unsigned long long cur = tkeys[i];
    if (cur == k) return reinterpret_cast<const uint16_t*>(lds + p.tidx_off)[i];
    if (cur == kGEmpty) return -1;
    i = (i + 1) & (T - 1);
  }
}

// Classifies every row: codes[r] = its hot index, or kCodeCold | its partition; counts
// counts[part * nh + hb] of the rows whose key is not hot and, with dpart, hcounts[hpart * nh + hb]
// of the hot rows' distinct pairs, for each of the nh = gridDim.x histogram blocks (plain
// stores).  kHistSplit histogram blocks of kHistThreads cover one main-pass block's rows, in
// order, so the main-pass block b's rows of a partition are the scanned range that starts at
// histogram block b * kHistSplit.  The kernel's LDS holds only the hot key table and the
// histograms, so several blocks share a CU.
constexpr int kHistSplit   = 3;  // 3 blocks of 512 per CU (76 VGPRs): the grid is one round
constexpr int kHistThreads = 512;
constexpr uint32_t kCodeCold = 0x8000u;

// Class table of a single key of at most 4 bytes whose sampled values lie in [kmin, kmin + range):
// code[key - kmin] is the key's row class, a hot index or kCodeCold | its partition (range 0: no
// table; code then still points at one readable entry).  Keys outside the range are classified by
// hashing and probing the hot key table.
struct DenseCodes {
  const uint16_t* code;
  uint32_t kmin;
  uint32_t range;
};
constexpr uint32_t kDenseMaxRange = 1u << 22;  // 8 MB of codes at most

__global__ void k_gb_dense_fill(uint16_t* code, uint32_t kmin, uint32_t range, int part_bits)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < range; i += gridDim.x * blockDim.x)
    code[i] = (uint16_t)(kCodeCold | cold_part(mix64((uint64_t)(uint32_t)(kmin + i)), part_bits));
}

// after k_gb_dense_fill: each hot key (other than the null key) inside the range takes its index
__global__ void k_gb_dense_hot(uint16_t* code, uint32_t kmin, uint32_t range, HotList hot, int cap)
{
  const int nh = (int)min(*hot.count, (unsigned int)cap);
  for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += gridDim.x * blockDim.x) {
    if (hot.special[h] != 0) continue;
    const uint32_t off = (uint32_t)hot.keys[h] - kmin;
    if (off < range) code[off] = (uint16_t)h;
  }
}

// Hot key table of k_gb_hist3 for keys of at most 4 bytes: one 8-byte LDS word per slot holding
// (key << 32 | hot index), so a probe is one LDS read; 4 * hot_cap slots, at most 4096.
constexpr int kHot32MaxSlots = 4096;
__host__ __device__ inline int hot32_slots(int hot_cap) { return 4 * hot_cap < kHot32MaxSlots ? 4 * hot_cap : kHot32MaxSlots; }

template <bool PACK32>
__global__ __launch_bounds__(kHistThreads) void k_gb_hist3(DKeys keys, int64_t n, GbPlan p, HotList hot, uint32_t* counts,
                                                         uint32_t* hcounts, uint16_t* __restrict__ codes,
                                                         DenseCodes dense)
{
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int s_null_hot, s_sent_hot;
  __shared__ uint64_t s_red[17];
  const int nparts = 1 << p.part_bits;
  uint32_t* hist   = reinterpret_cast<uint32_t*>(lds + p.lds_bytes);  // after the hot table
  const int nhot   = (int)min(*hot.count, (unsigned int)p.hot_cap);
  uint64_t npairs  = 0;  // this thread's hot rows with a COUNT_DISTINCT value
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) hist[i] = 0;
  const int t32             = hot32_slots(p.hot_cap);
  unsigned long long* tab32 = reinterpret_cast<unsigned long long*>(lds);
  if (PACK32) {
    for (int i = threadIdx.x; i < t32; i += blockDim.x) tab32[i] = kGEmpty;
    if (threadIdx.x == 0) {
      s_null_hot = -1;
      s_sent_hot = -1;
    }
    __syncthreads();
    for (int h = threadIdx.x; h < nhot; h += blockDim.x) {
      const int sp = hot.special[h];
      if (sp == 1) s_sent_hot = h;
      else if (sp == 2) s_null_hot = h;
      else {
        const unsigned long long k = hot.keys[h];
        uint32_t i                 = (uint32_t)(mix64(k) & (uint64_t)(t32 - 1));
        while (atomicCAS(&tab32[i], kGEmpty, (k << 32) | (uint32_t)h) != kGEmpty) i = (i + 1) & (uint32_t)(t32 - 1);
      }
    }
    __syncthreads();
  } else {
    hot_setup(p, lds, hot, nhot, false, true, &s_null_hot, &s_sent_hot);
  }
  // the class of a null key: its hot index, or the partition of key bits 0 (where the partition
  // pass keeps null keys apart by their flag)
  const uint32_t null_code = s_null_hot >= 0 ? (uint32_t)s_null_hot : (kCodeCold | cold_part(mix64(0ull), p.part_bits));
  const int64_t mb0   = (int64_t)(blockIdx.x / kHistSplit) * p.rows_per_block;  // main-pass block's rows
  const int64_t share = (p.rows_per_block + kHistSplit - 1) / kHistSplit;
  const int sub       = (int)(blockIdx.x % kHistSplit);
  const int64_t begin = mb0 + (int64_t)sub * share;
  int64_t end         = mb0 + ((int64_t)(sub + 1) * share < p.rows_per_block ? (int64_t)(sub + 1) * share : p.rows_per_block);
  end                 = end < n ? end : n;
  const int nh        = (int)gridDim.x;
  const DCol& kc      = keys.c[0];
  const ColDev& dc    = p.col[p.dpart ? p.dcol : 0];
  constexpr int R = PACK32 ? 4 : 8;  // rows per thread whose loads are in flight together
  const BlockCol bk = block_col(kc.data, kc.valid, kc.width, begin < end ? begin : 0);
  const BlockCol bd = block_col(kc.data, p.dpart ? dc.valid : nullptr, kc.width, begin < end ? begin : 0);
  using KT = typename std::conditional<PACK32, uint32_t, uint64_t>::type;
  for (int64_t r0 = begin + threadIdx.x; r0 < end; r0 += (int64_t)R * blockDim.x) {
    KT k[R];
    bool kv[R], dv[R];  // dv: the row has a COUNT_DISTINCT value (only its presence is counted here)
    // branch-free loads, all in flight before any is used (rows past the end load row `begin`)
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int64_t r   = r0 + (int64_t)j * blockDim.x;
      const uint32_t in = r < end ? 1u : 0u;
      const uint32_t o  = in ? (uint32_t)(r - begin) : 0u;
      const uint32_t k1 = bc_valid(bk, o) & in;
      kv[j]             = k1 != 0;
      k[j]              = (KT)(bc_bits(bk, o) & (0ull - (uint64_t)k1));
      dv[j]             = (bc_valid(bd, o) & in & (uint32_t)p.dpart) != 0;
    }
    int hh[R];
    uint32_t cpj[R];
    bool tabd[R];  // the row's code came from the dense table (or it is a null key)
    if (PACK32) {
      // Keys inside the dense range take their code from the table (loads all in flight); a null
      // key takes the block's null code; only other keys (outside the sampled range) are hashed
      // and probed below, by the whole wave when any of its rows needs it.
      uint32_t tc[R];
      bool need = false;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const uint32_t off = (uint32_t)k[j] - (uint32_t)dense.kmin;
        tabd[j]            = !kv[j] || off < dense.range;
        tc[j]              = dense.code[off < dense.range ? off : 0u];
        need |= !tabd[j];
      }
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const uint32_t c = kv[j] ? tc[j] : null_code;
        hh[j]            = (c & kCodeCold) ? -1 : (int)c;
        cpj[j]           = c & (kCodeCold - 1);
      }
      if (__any(need)) {  // every row's first probe is issued before any is resolved
      unsigned long long e[R];
      uint32_t slot[R];
      uint32_t cph[R];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const uint64_t hk = mix64((uint64_t)k[j]);
        slot[j]           = (uint32_t)(hk & (uint64_t)(t32 - 1));
        cph[j]            = cold_part(hk, p.part_bits);
        e[j]              = tab32[slot[j]];
      }
      // every row's probe sequence advanced together: each round reads the next slot of all rows
      // still looking (their LDS reads in flight at once) until a row finds its key or an empty slot
      bool look[R];
      int hp[R];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        hp[j]   = -1;
        look[j] = !tabd[j];
      }
      for (int round = 0; round < t32; ++round) {
        bool any = false;
#pragma unroll
        for (int j = 0; j < R; ++j) {
          if (!look[j]) continue;
          if (e[j] == kGEmpty) look[j] = false;
          else if ((e[j] >> 32) == k[j]) {
            hp[j]   = (int)(uint32_t)e[j];
            look[j] = false;
          } else {
            slot[j] = (slot[j] + 1) & (uint32_t)(t32 - 1);
          }
          any |= look[j];
        }
        if (!any) break;
#pragma unroll
        for (int j = 0; j < R; ++j) e[j] = tab32[look[j] ? slot[j] : 0u];
      }
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (!tabd[j]) {
          hh[j]  = hp[j];
          cpj[j] = cph[j];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int64_t r = r0 + (int64_t)j * blockDim.x;
      if (r >= end) break;
      int h;
      uint32_t cp;
      if (PACK32) {
        h  = hh[j];
        cp = cpj[j];
      } else {
        const uint64_t hk = mix64((uint64_t)k[j]);
        h                 = hot_lookup(p, lds, (uint64_t)k[j], hk, kv[j], s_null_hot, s_sent_hot);
        cp                = cold_part(hk, p.part_bits);
      }
      codes[r]          = (uint16_t)(h >= 0 ? (uint32_t)h : (kCodeCold | cp));
      if (h < 0) atomicAdd(&hist[cp], 1u);
      else if (dv[j]) ++npairs;
    }
  }
  const uint64_t block_pairs = block_reduce_sum(npairs, s_red);  // synchronises the block
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) counts[(int64_t)i * nh + blockIdx.x] = hist[i];
  if (p.dpart && threadIdx.x == 0) hcounts[blockIdx.x] = (uint32_t)block_pairs;
}

template <int C>
__device__ __forceinline__ uint64_t colv(const RowVals& v)
{
  if constexpr (C == 0) return v.b0;
  else if constexpr (C == 1) return v.b1;
  else if constexpr (C == 2) return v.b2;
  else return v.b3;
}

// The VCOUNT / ISUM / FSUM / MIN / MAX accumulators of value column C (a compile-time index, so a
// row's value is a fixed register) over R rows of one thread; s[j] is the row's LDS slot (< 0:
// not aggregated here).  `off` gives each accumulator's LDS byte offset (the hot table's or the
// partition table's layout).
template <int C, int R, typename Off>
__device__ __forceinline__ void acc_col_rows(const GbPlan& p, const Off& off, unsigned char* lds, uint32_t* flags,
                                             const int (&s)[R], const RowVals (&v)[R])
{
  const int ct = p.col[C].type;
  const int cw = p.col[C].width;
  bool ok[R];
#pragma unroll
  for (int j = 0; j < R; ++j) ok[j] = s[j] >= 0 && ((v[j].valid >> C) & 1);
  int ai = p.cacc[C][0];
  if (ai >= 0) {  // VCOUNT
    uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + off[ai]);
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (ok[j]) atomicAdd(&cnt[s[j]], 1u);
  }
  ai = p.cacc[C][1];
  if (ai >= 0) {  // ISUM
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds + off[ai]);
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (ok[j]) atomicAdd(&acc[s[j]], (unsigned long long)bits_as_int(colv<C>(v[j]), ct, cw));
  }
  ai = p.cacc[C][2];
  if (ai >= 0) {  // FSUM
    unsigned long long* base = reinterpret_cast<unsigned long long*>(lds + off[ai]);
    bool bad                 = false;
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (ok[j] && !lds_fsum_add(&base[6 * s[j]], bits_as_double(colv<C>(v[j]), ct, cw))) bad = true;
    if (bad) *p.a[ai].err = 1;
  }
#pragma unroll
  for (int mm = 0; mm < 2; ++mm) {  // MIN, MAX
    ai = p.cacc[C][3 + mm];
    if (ai < 0) continue;
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(lds + off[ai]);
    const uint32_t fbit     = 1u << p.a[ai].flag_bit;
    const bool fl           = is_float_type(ct);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (!ok[j]) continue;
      const uint64_t vb = colv<C>(v[j]);
      if (fl && isnan(bits_as_double(vb, ct, cw))) {
        atomicOr(&flags[s[j]], fbit);
        continue;
      }
      const uint64_t o = ordered_of(vb, ct, cw);
      if (mm == 0) atomicMin(&acc[s[j]], (unsigned long long)o);
      else atomicMax(&acc[s[j]], (unsigned long long)o);
    }
  }
}

template <int R, typename Off>
__device__ __forceinline__ void acc_cols_rows(const GbPlan& p, const Off& off, unsigned char* lds, uint32_t* flags,
                                              const int (&s)[R], const RowVals (&v)[R])
{
  if (p.ncol > 0) acc_col_rows<0, R>(p, off, lds, flags, s, v);
  if (p.ncol > 1) acc_col_rows<1, R>(p, off, lds, flags, s, v);
  if (p.ncol > 2) acc_col_rows<2, R>(p, off, lds, flags, s, v);
  if (p.ncol > 3) acc_col_rows<3, R>(p, off, lds, flags, s, v);
}

// LDS byte offset of accumulator i: in the hot-key tables (the plan's own offsets) or in a
// partition table (k_gb_part3's layout)
struct HotOff {
  const GbPlan& p;
  __device__ int32_t operator[](int i) const { return p.a[i].lds_off; }
};
struct PartOff {
  const int32_t* o;
  __device__ int32_t operator[](int i) const { return o[i]; }
};

// LDS accumulators of R rows of one thread at once (h[j] = hot index, < 0 for rows that are not
// hot): per value column (acc_col_rows), then the PRODUCT and COUNT_DISTINCT accumulators.
template <bool DPART, int R>
__device__ __forceinline__ void acc_hot_rows(const GbPlan& p, unsigned char* lds, const DKeys& keys, int64_t r0,
                                             const int (&h)[R], const HotList& hot, const RowVals (&v)[R])
{
  const HotOff off{p};
  uint32_t* rows = reinterpret_cast<uint32_t*>(lds + p.rows_off);
#pragma unroll
  for (int j = 0; j < R; ++j)
    if (h[j] >= 0) atomicAdd(&rows[h[j]], 1u);
  acc_cols_rows<R>(p, off, lds, reinterpret_cast<uint32_t*>(lds + p.flags_off), h, v);
  for (int i = 0; i < p.nacc; ++i) {
    const AccDev& a = p.a[i];
    if (a.kind != ACC_PROD && a.kind != ACC_DISTINCT) continue;
    const int col = a.col;
    const int ct  = p.col[col].type;
    const int cw  = p.col[col].width;
    bool ok[R];
#pragma unroll
    for (int j = 0; j < R; ++j) ok[j] = h[j] >= 0 && ((v[j].valid >> col) & 1);
    switch (a.kind) {
      case ACC_PROD: {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          if (!ok[j]) continue;
          const uint64_t vb = pick(v[j], col);
          const int64_t s   = hot.gslot[h[j]];
          if (is_float_type(ct)) atomic_mul_f64(&a.g0[s], bits_as_double(vb, ct, cw));
          else atomic_mul_u64(&a.g0[s], bits_as_int(vb, ct, cw));
        }
        break;
      }
      case ACC_DISTINCT: {
        if (DPART) break;  // the pair goes to a pair partition instead
        uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + a.lds_off);
#pragma unroll
        for (int j = 0; j < R; ++j) {
          if (!ok[j]) continue;
          const int64_t r = r0 + (int64_t)j * blockDim.x + threadIdx.x;
          if (distinct_insert(a, p.col[col], keys, r, hot.gslot[h[j]], pick(v[j], col))) atomicAdd(&cnt[h[j]], 1u);
        }
        break;
      }
      default: break;
    }
  }
}

// a field of `width` bytes into LDS (its offset is a multiple of its width)
__device__ __forceinline__ void lds_put(unsigned char* at, int width, uint64_t bits)
{
  switch (width) {
    case 1: *at = (uint8_t)bits; break;
    case 2: *reinterpret_cast<uint16_t*>(at) = (uint16_t)bits; break;
    case 4: *reinterpret_cast<uint32_t*>(at) = (uint32_t)bits; break;
    default: *reinterpret_cast<uint64_t*>(at) = bits; break;
  }
}

// Main pass: hot rows into LDS accumulators, other rows written straight into their partition:
// each takes the next position of its partition's region with an LDS atomic on the region
// cursor (no barrier per tile; order inside a partition is irrelevant because every
// accumulator is exact and order-independent).  LDS: [0, p.acc_end) hot accumulators, then
// cursor u32[2^part_bits] (record indices: fewer than 2^31 rows), and with p.stage one record-sized slot per thread where a record's
// fields are written before it is stored as 16-byte pieces.  With DPART the hot rows'
// (hot index, value) pairs are appended in row order to the block's range of pb.hwords /
// pb.hparts as their image under pair_mix (one LDS counter, coalesced stores); k_gb_pair_part then
// splits them by pair partition.
template <bool DPART, int NC>
__global__ __launch_bounds__(kGbThreads) void k_gb_main3(DKeys keys, int64_t n, GbPlan p, HotList hot, PartBufs pb,
                                                         HotParts hp, const uint16_t* __restrict__ codes)
{
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int s_null_hot, s_sent_hot;
  uint32_t* cursor           = reinterpret_cast<uint32_t*>(lds + p.acc_end);
  const int nparts           = 1 << p.part_bits;
  uint32_t* phist            = reinterpret_cast<uint32_t*>(cursor + nparts);  // DPART: pairs per pair partition
  unsigned char* stage = reinterpret_cast<unsigned char*>(phist + (DPART ? kHPairParts : 0)) +
                         (size_t)threadIdx.x * (size_t)pb.stride;

  const int nhot = (int)min(*hot.count, (unsigned int)p.hot_cap);
  const int64_t hb0 = (int64_t)blockIdx.x * kHistSplit;  // first histogram block of this block's rows
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) cursor[i] = (uint32_t)pb.offsets[(int64_t)i * pb.nblk + hb0];
  __shared__ uint32_t s_pcur;  // hot pairs this block has appended
  if (threadIdx.x == 0) s_pcur = 0;
  if (DPART)
    for (int i = threadIdx.x; i < kHPairParts; i += blockDim.x) phist[i] = 0;
  const int64_t pair0 = DPART ? (int64_t)pb.hoffsets[hb0] : 0;  // the block's first hot pair
  hot_setup(p, lds, hot, nhot, true, false, &s_null_hot, &s_sent_hot);  // rows come classified

  const int64_t begin = (int64_t)blockIdx.x * p.rows_per_block;
  const int64_t end   = begin + p.rows_per_block < n ? begin + p.rows_per_block : n;
  const DCol& kc      = keys.c[0];
  // the block's rows through 32-bit offsets from `begin` (rows_per_block < 2^31)
  const BlockCol bk         = block_col(kc.data, kc.valid, kc.width, begin);
  const uint16_t* cbase     = codes + begin;
  BlockCol bcol[NC > 0 ? NC : 1];
#pragma unroll
  for (int c = 0; c < NC; ++c) bcol[c] = block_col(p.col[c].data, p.col[c].valid, p.col[c].width, begin);
  for (int64_t sub = begin; sub < end; sub += kSub3) {
    RowVals v[kGbRowsPerThread];
    uint64_t kb[kGbRowsPerThread];
    bool kv[kGbRowsPerThread];
    uint32_t code[kGbRowsPerThread];
    // every load of the thread's rows is issued before any is used (one branch-free load per
    // value, combined by masks; rows past the end load the block's first row and are masked)
#pragma unroll
    for (int j = 0; j < kGbRowsPerThread; ++j) {
      const int64_t r   = sub + (int64_t)j * blockDim.x + threadIdx.x;
      const uint32_t in = r < end ? 1u : 0u;
      const uint32_t o  = in ? (uint32_t)(r - begin) : 0u;
      code[j]           = (uint32_t)cbase[o] & (0u - in);
      const uint32_t k1 = bc_valid(bk, o) & in;
      kv[j]             = k1 != 0;
      kb[j]             = bc_bits(bk, o) & (0ull - (uint64_t)k1);
      uint32_t vm       = 0;
      uint64_t x[kMaxCols] = {0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        x[c] = bc_bits(bcol[c], o) & (0ull - (uint64_t)in);
        vm |= (bc_valid(bcol[c], o) & in) << c;
      }
      v[j] = RowVals{x[0], x[1], x[2], x[3], vm};
    }
    int hx[kGbRowsPerThread];  // hot index; -1 not hot; -2 past the end
    uint32_t part[kGbRowsPerThread];
#pragma unroll
    for (int j = 0; j < kGbRowsPerThread; ++j) {
      const int64_t r = sub + (int64_t)j * blockDim.x + threadIdx.x;
      part[j]         = code[j] & (kCodeCold - 1);
      hx[j]           = r >= end ? -2 : ((code[j] & kCodeCold) ? -1 : (int)code[j]);
    }
    acc_hot_rows<DPART, kGbRowsPerThread>(p, lds, keys, sub, hx, hot, v);
    if (DPART) {  // hot pairs appended in row order, wave by wave (coalesced stores)
      const int lane       = threadIdx.x & (kWave - 1);
      const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
      for (int j = 0; j < kGbRowsPerThread; ++j) {
        const bool has   = hx[j] >= 0 && ((v[j].valid >> p.dcol) & 1);
        const uint64_t m = __ballot(has);
        if (m == 0) continue;
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t at      = 0;
        if (lane == leader) at = atomicAdd(&s_pcur, (uint32_t)__popcll(m));
        at = __shfl(at, leader, kWave);
        if (has) {
          const uint64_t x = pair_mix(((uint64_t)(uint32_t)hx[j] << 32) | (uint32_t)pick(v[j], p.dcol));
          const int64_t o  = pair0 + at + (uint32_t)__popcll(m & below);
          pb.hwords[o]     = (uint32_t)x;
          pb.hparts[o]     = (uint16_t)(x >> 32);
          atomicAdd(&phist[(uint32_t)(x >> 32)], 1u);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kGbRowsPerThread; ++j) {
      if (hx[j] != -1) continue;
      uint32_t fl = (v[j].valid & 0xF) | (kv[j] ? kFlagKeyValid : 0);
      int di      = 0;
      for (int i = 0; i < p.nacc && !DPART; ++i) {
        const AccDev& a = p.a[i];
        if (a.kind != ACC_DISTINCT) continue;
        if ((v[j].valid >> a.col) & 1) {
          // (key, value) pair of a key outside the hot set: both fit 32 bits on this path
          unsigned long long pk = ((unsigned long long)(uint32_t)kb[j] << 32) | (uint32_t)pick(v[j], a.col);
          uint64_t q            = mix64(pk) & a.pmask2;
          while (true) {  // try the CAS first: one round trip for a new pair
            unsigned long long old = atomicCAS(&a.pset2[q], kGEmpty, pk);
            if (old == kGEmpty) {
              fl |= 1u << (kFlagDistinct + di);
              break;
            }
            if (old == pk) break;
            q = (q + 1) & a.pmask2;
          }
        }
        ++di;
      }
      const int64_t dest = (int64_t)atomicAdd(&cursor[part[j]], 1u);
      if (p.stage) {  // fields into this thread's LDS slot, then the record as 16-byte pieces
        lds_put(stage + pb.key_off, pb.key_width, kb[j]);
        stage[pb.flags_off] = (uint8_t)fl;
        if (p.ncol > 0) lds_put(stage + pb.val_off[0], p.col[0].width, v[j].b0);
        if (p.ncol > 1) lds_put(stage + pb.val_off[1], p.col[1].width, v[j].b1);
        if (p.ncol > 2) lds_put(stage + pb.val_off[2], p.col[2].width, v[j].b2);
        if (p.ncol > 3) lds_put(stage + pb.val_off[3], p.col[3].width, v[j].b3);
        unsigned char* d = pb.rec + dest * (int64_t)pb.stride;
        if (pb.stride == 8) {
          *reinterpret_cast<uint64_t*>(d) = *reinterpret_cast<const uint64_t*>(stage);
        } else {
          for (int k = 0; k < pb.stride; k += 16)
            *reinterpret_cast<uint4*>(d + k) = *reinterpret_cast<const uint4*>(stage + k);
        }
      } else {
        RecWords w;
        w.put(pb.key_off, kb[j]);
        w.put(pb.flags_off, fl);
        if (p.ncol > 0) w.put(pb.val_off[0], v[j].b0);
        if (p.ncol > 1) w.put(pb.val_off[1], v[j].b1);
        if (p.ncol > 2) w.put(pb.val_off[2], v[j].b2);
        if (p.ncol > 3) w.put(pb.val_off[3], v[j].b3);
        w.store(pb.rec, pb.stride, dest);
      }
    }
  }
  __syncthreads();
  if (DPART)  // this block's pairs per pair partition (partition-major, for the split's scan)
    for (int i = threadIdx.x; i < kHPairParts; i += blockDim.x) pb.pcounts[(int64_t)i * gridDim.x + blockIdx.x] = phist[i];

  // hot partials of this block
  const int H = p.hot_cap;
  for (int h = threadIdx.x; h < nhot; h += blockDim.x) {
    const int64_t o = (int64_t)blockIdx.x * H + h;
    hp.rows[o]      = reinterpret_cast<const uint32_t*>(lds + p.rows_off)[h];
    hp.flags[o]     = reinterpret_cast<const uint32_t*>(lds + p.flags_off)[h];
    for (int j = 0; j < p.nacc; ++j) {
      const AccDev& a = p.a[j];
      if (a.lds_off < 0) continue;
      if (a.kind == ACC_VCOUNT || a.kind == ACC_DISTINCT) {
        hp.acc[j][o] = reinterpret_cast<const uint32_t*>(lds + a.lds_off)[h];
      } else if (a.kind == ACC_FSUM) {
        uint64_t b0, b1, b2;
        lds_fsum_value(reinterpret_cast<const unsigned long long*>(lds + a.lds_off) + 6 * h, b0, b1, b2);
        hp.acc[j][3 * o]     = b0;
        hp.acc[j][3 * o + 1] = b1;
        hp.acc[j][3 * o + 2] = b2;
      } else {
        hp.acc[j][o] = reinterpret_cast<const unsigned long long*>(lds + a.lds_off)[h];
      }
    }
  }
}

// Combines the per-block partials of each hot key (one block per key) and stores the totals in
// its global slot.
__global__ __launch_bounds__(256) void k_gb_hot_reduce(GbPlan p, HotList hot, HotParts hp, int nblk)
{
  __shared__ unsigned long long red[3][256];
  const int H    = p.hot_cap;
  const int nhot = (int)min(*hot.count, (unsigned int)H);
  for (int h = blockIdx.x; h < nhot; h += gridDim.x) {
    const int64_t s = hot.gslot[h];
    for (int j = -1; j < p.nacc; ++j) {
      const int kind = j < 0 ? -1 : p.a[j].kind;
      if (j >= 0 && (kind == ACC_PROD || p.a[j].lds_off < 0) && kind != ACC_ROWS) continue;
      uint64_t x0 = 0, x1 = 0, x2 = 0;
      if (kind == ACC_MIN) x0 = ~0ull;
      for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const int64_t o = (int64_t)b * H + h;
        switch (kind) {
          case -1: x0 |= hp.flags[o]; break;
          case ACC_ROWS: x0 += hp.rows[o]; break;
          case ACC_FSUM: {
            uint64_t a0 = hp.acc[j][3 * o], a1 = hp.acc[j][3 * o + 1], a2 = hp.acc[j][3 * o + 2];
            uint64_t t0 = x0 + a0, c0 = t0 < x0;
            uint64_t t1 = x1 + a1, c1 = t1 < x1;
            uint64_t t1c = t1 + c0;
            c1 |= t1c < t1;
            x0 = t0;
            x1 = t1c;
            x2 = x2 + a2 + c1;
            break;
          }
          case ACC_MIN: x0 = min(x0, (uint64_t)hp.acc[j][o]); break;
          case ACC_MAX: x0 = max(x0, (uint64_t)hp.acc[j][o]); break;
          default: x0 += hp.acc[j][o]; break;
        }
      }
      red[0][threadIdx.x] = x0;
      red[1][threadIdx.x] = x1;
      red[2][threadIdx.x] = x2;
      __syncthreads();
      for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
          uint64_t a0 = red[0][threadIdx.x], b0 = red[0][threadIdx.x + w];
          switch (kind) {
            case -1: red[0][threadIdx.x] = a0 | b0; break;
            case ACC_MIN: red[0][threadIdx.x] = min(a0, b0); break;
            case ACC_MAX: red[0][threadIdx.x] = max(a0, b0); break;
            case ACC_FSUM: {
              uint64_t a1 = red[1][threadIdx.x], b1 = red[1][threadIdx.x + w];
              uint64_t t0 = a0 + b0, c0 = t0 < a0;
              uint64_t t1 = a1 + b1, c1 = t1 < a1;
              uint64_t t1c = t1 + c0;
              c1 |= t1c < t1;
              red[0][threadIdx.x] = t0;
              red[1][threadIdx.x] = t1c;
              red[2][threadIdx.x] = red[2][threadIdx.x] + red[2][threadIdx.x + w] + c1;
              break;
            }
            default: red[0][threadIdx.x] = a0 + b0; break;
          }
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        if (kind == -1) {
          for (int q = 0; q < p.nacc; ++q)
            if ((p.a[q].kind == ACC_MIN || p.a[q].kind == ACC_MAX) && ((red[0][0] >> p.a[q].flag_bit) & 1))
              p.a[q].gflag[s] = 1u;
        } else {
          p.a[j].g0[s] = red[0][0];
          if (kind == ACC_FSUM) {
            p.a[j].g1[s] = red[1][0];
            p.a[j].g2[s] = red[2][0];
          }
        }
      }
      __syncthreads();
    }
  }
}

// LDS accumulators of R partition rows of one thread (slot[j] < 0: the row is not accumulated
// here); DISTINCT counts come from the rows' flag bits.  Accumulator loop outside, row loop inside.
template <int R>
__device__ __forceinline__ void acc_part_rows(const GbPlan& p, const PartPlan& q, unsigned char* lds,
                                              const int (&slot)[R], const RowVals (&v)[R], const uint32_t (&fl)[R])
{
  uint32_t* rows = reinterpret_cast<uint32_t*>(lds + q.rows_off);
#pragma unroll
  for (int j = 0; j < R; ++j)
    if (slot[j] >= 0) atomicAdd(&rows[slot[j]], 1u);
  int di = 0;
  for (int i = 0; i < p.nacc; ++i) {
    if (p.a[i].kind != ACC_DISTINCT) continue;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(lds + q.acc_off[i]);
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (slot[j] >= 0 && ((fl[j] >> (kFlagDistinct + di)) & 1)) atomicAdd(&cnt[slot[j]], 1u);
    ++di;
  }
  acc_cols_rows<R>(p, PartOff{q.acc_off}, lds, reinterpret_cast<uint32_t*>(lds + q.flags_off), slot, v);
}

// the same into the global arrays at slot s (rows a full partition table could not take)
__device__ __forceinline__ void acc_spill(const GbPlan& p, int64_t s, const RowVals v, uint32_t fl)
{
  int di = 0;
  for (int i = 0; i < p.nacc; ++i) {
    const AccDev& a = p.a[i];
    if (a.kind == ACC_ROWS) {
      atomicAdd(&a.g0[s], 1ull);
      continue;
    }
    if (a.kind == ACC_DISTINCT) {
      if ((fl >> (kFlagDistinct + di)) & 1) atomicAdd(&a.g0[s], 1ull);
      ++di;
      continue;
    }
    if (!((v.valid >> a.col) & 1)) continue;
    const ColDev& c   = p.col[a.col];
    const uint64_t vb = pick(v, a.col);
    switch (a.kind) {
      case ACC_VCOUNT: atomicAdd(&a.g0[s], 1ull); break;
      case ACC_ISUM: atomicAdd(&a.g0[s], (unsigned long long)bits_as_int(vb, c.type, c.width)); break;
      case ACC_FSUM: {
        uint64_t l0, l1, l2;
        if (to_fixed192(bits_as_double(vb, c.type, c.width), l0, l1, l2)) add_fixed192(&a.g0[s], &a.g1[s], &a.g2[s], l0, l1, l2);
        else *a.err = 1;
        break;
      }
      case ACC_MIN:
      case ACC_MAX: {
        if (is_float_type(c.type) && isnan(bits_as_double(vb, c.type, c.width))) {
          atomicOr(&a.gflag[s], 1u);
          break;
        }
        uint64_t o = ordered_of(vb, c.type, c.width);
        if (a.kind == ACC_MIN) atomicMin(&a.g0[s], (unsigned long long)o);
        else atomicMax(&a.g0[s], (unsigned long long)o);
        break;
      }
      default: break;
    }
  }
}

// One block per (partition, sub-pass); the S blocks of a partition share an XCD (block % 8) so
// the partition's rows are re-read from that XCD's L2.
template <int W>
__global__ __launch_bounds__(kGbThreads) void k_gb_part3(GbPlan p, PartPlan q, PartBufs pb, KeyTab t)
{
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ uint32_t s_pcur;
  const int xcd   = blockIdx.x & 7;
  const int rest  = blockIdx.x >> 3;
  const int sub   = rest % q.S;
  const int part  = (rest / q.S) * 8 + xcd;
  if (part >= (1 << p.part_bits)) return;
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(lds + q.key_off);
  for (int i = threadIdx.x; i < q.tcap; i += blockDim.x) {
    keys[i]                                           = kGEmpty;
    reinterpret_cast<uint32_t*>(lds + q.rows_off)[i]  = 0;
    reinterpret_cast<uint32_t*>(lds + q.flags_off)[i] = 0;
  }
  if (threadIdx.x == 0) s_pcur = 0;
  for (int j = 0; j < p.nacc; ++j) {
    const AccDev& a = p.a[j];
    if (q.acc_off[j] < 0) continue;
    if (a.kind == ACC_VCOUNT || a.kind == ACC_DISTINCT) {
      for (int i = threadIdx.x; i < q.tcap; i += blockDim.x) reinterpret_cast<uint32_t*>(lds + q.acc_off[j])[i] = 0;
    } else {
      const int words            = a.kind == ACC_FSUM ? 6 * q.tcap : q.tcap;
      const unsigned long long v = a.kind == ACC_MIN ? ~0ull : 0ull;
      for (int i = threadIdx.x; i < words; i += blockDim.x) reinterpret_cast<unsigned long long*>(lds + q.acc_off[j])[i] = v;
    }
  }
  __syncthreads();

  const int64_t begin = (int64_t)pb.offsets[(int64_t)part * pb.nblk];
  const int64_t end   = (int64_t)pb.offsets[(int64_t)(part + 1) * pb.nblk];
  const uint32_t mask = (uint32_t)q.tcap - 1;
  // the LDS exact-sum words take fewer than 2^31 addends: a larger partition uses the global slots
  const bool big  = end - begin >= (int64_t(1) << 31);
  // this block's dense slots (after the hash table and its two special slots): every key outside
  // the hot set lives in exactly one (partition, sub-pass) block, which stores its group there
  const int64_t dbase = (int64_t)t.cap + 2 + ((int64_t)part * q.S + sub) * q.tcap;
  constexpr int R = W >= 8 ? 2 : 4;  // rows per thread in flight (R * W record words in registers)
  for (int64_t i0 = begin + threadIdx.x; i0 < end; i0 += (int64_t)R * blockDim.x) {
    uint64_t w[R][W];
#pragma unroll
    for (int j = 0; j < R; ++j) {  // issue the loads of R rows before using any (unconditional:
      const int64_t i = i0 + (int64_t)j * blockDim.x;  // a row past the end loads row `begin`)
      rec_load<W>(pb.rec, i < end ? i : begin, w[j]);
    }
    int slot[R];  // LDS slot; -1 spilled to the global table; -2 not a row of this block
    uint32_t fl[R];
    uint64_t kr[R];
    RowVals vr[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      slot[j] = -2;
      fl[j]   = 0;
      kr[j]   = 0;
      vr[j]   = RowVals{0, 0, 0, 0, 0};
      if (i0 + (int64_t)j * blockDim.x >= end) continue;
      const uint64_t k  = rec_get<W>(w[j], pb.key_off, pb.key_width);
      const uint64_t hk = mix64(k);
      if (q.S > 1 && (int)cold_sub(hk, q.S) != sub) continue;
      kr[j]                = k;
      fl[j]                = (uint32_t)rec_get<W>(w[j], pb.flags_off, 1);
      uint64_t x[kMaxCols] = {0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < kMaxCols; ++c)
        if (c < p.ncol) x[c] = rec_get<W>(w[j], pb.val_off[c], p.col[c].width);
      vr[j]         = RowVals{x[0], x[1], x[2], x[3], fl[j] & 0xF};
      const bool kv = fl[j] & kFlagKeyValid;
      // A key spills when the first kPartProbes slots of its probe sequence hold other keys.
      // Slots never become empty again, so every thread reaches the same verdict for a key and a
      // key is either in this table or spilled, never both.
      slot[j] = -1;
      if (kv && k != kGEmpty && !big) {
        uint32_t h = (uint32_t)(hk & mask);
        for (int probe = 0; probe < kPartProbes; ++probe) {
          unsigned long long cur = keys[h];
          if (cur == k) {
            slot[j] = (int)h;
            break;
          }
          if (cur == kGEmpty) {
            unsigned long long old = atomicCAS(&keys[h], kGEmpty, (unsigned long long)k);
            if (old == kGEmpty || old == k) {
              slot[j] = (int)h;
              break;
            }
          }
          h = (h + 1) & mask;
        }
      }
    }
    acc_part_rows<R>(p, q, lds, slot, vr, fl);
    int64_t gs[R];  // the row's accumulator slot: its key's dense slot, or its global slot if spilled
#pragma unroll
    for (int j = 0; j < R; ++j) {
      gs[j] = slot[j] >= 0 ? dbase + slot[j] : -1;
      if (slot[j] != -1) continue;
      const int64_t s = gb_insert_bits(kr[j], (fl[j] & kFlagKeyValid) != 0, t);
      if (s >= 0) acc_spill(p, s, vr[j], fl[j]);
      gs[j] = s;
    }
    if (pb.cpairs != nullptr) {  // the rows' COUNT_DISTINCT pairs (slot, value), appended wave by wave
      const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const bool has = gs[j] >= 0 && ((fl[j] >> p.dcol) & 1);
        const uint64_t m = __ballot(has);
        if (m == 0) continue;
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base    = 0;
        if (lane == leader) base = atomicAdd(&s_pcur, (uint32_t)__popcll(m));
        base = __shfl(base, leader, kWave);
        if (has) {
          const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
          pb.cpairs[begin + base + (uint32_t)__popcll(m & below)] =
              ((unsigned long long)(uint32_t)gs[j] << 32) | (uint32_t)pick(vr[j], p.dcol);
        }
      }
    }
  }
  __syncthreads();
  if (pb.cpairs != nullptr && threadIdx.x == 0) pb.cpair_count[part] = s_pcur;
  // Store this block's groups in its own dense slots, after the hash table and its two special
  // slots: every key outside the hot set lives in exactly one (partition, sub-pass) block, so the
  // slot needs no hashing and no atomics.
  for (int h = threadIdx.x; h < q.tcap; h += blockDim.x) {
    const unsigned long long k = keys[h];
    if (k == kGEmpty) continue;
    const int64_t s = dbase + h;
    t.slots[s]      = k;
    const uint32_t nr = reinterpret_cast<const uint32_t*>(lds + q.rows_off)[h];
    const uint32_t fg = reinterpret_cast<const uint32_t*>(lds + q.flags_off)[h];
    for (int j = 0; j < p.nacc; ++j) {
      const AccDev& a = p.a[j];
      switch (a.kind) {
        case ACC_ROWS: a.g0[s] = nr; break;
        case ACC_VCOUNT:
        case ACC_DISTINCT: a.g0[s] = reinterpret_cast<const uint32_t*>(lds + q.acc_off[j])[h]; break;
        case ACC_ISUM: a.g0[s] = reinterpret_cast<const unsigned long long*>(lds + q.acc_off[j])[h]; break;
        case ACC_FSUM: {
          uint64_t b0, b1, b2;
          lds_fsum_value(reinterpret_cast<const unsigned long long*>(lds + q.acc_off[j]) + 6 * h, b0, b1, b2);
          a.g0[s] = b0;
          a.g1[s] = b1;
          a.g2[s] = b2;
          break;
        }
        case ACC_MIN:
        case ACC_MAX:
          a.g0[s] = reinterpret_cast<const unsigned long long*>(lds + q.acc_off[j])[h];
          if ((fg >> a.flag_bit) & 1) a.gflag[s] = 1u;
          break;
        default: break;
      }
    }
  }
}

// FP64 fallback (sums that left the fixed-point range) and the VARIANCE second pass.
// ---- COUNT_DISTINCT through pair partitions -----------------------------------------------
// slot of a single fixed-width key given by its bits, -1 when absent
__device__ __forceinline__ int64_t gb_find_bits(unsigned long long key, const KeyTab& t)
{
  if (key == kGEmpty) return (int64_t)t.cap;
  uint64_t i = mix64(key) & t.mask;
  for (uint64_t probes = 0, lim = probe_limit(t); probes < lim; ++probes) {
    const unsigned long long cur = t.slots[i];
    if (cur == key) return (int64_t)i;
    if (cur == kGEmpty) return -1;
    i = (i + 1) & t.mask;
  }
  return -1;
}

// Inserts pr into an LDS set of kPairSet slots; 1 = new, 0 = present, -1 = no room.
__device__ __forceinline__ int pair_set_insert(unsigned long long* set, unsigned long long pr, uint64_t hsh)
{
  uint32_t i = (uint32_t)(hsh & (kPairSet - 1));
  for (int t = 0; t < kPairSet; ++t) {
    const unsigned long long old = atomicCAS(&set[i], kGEmpty, pr);
    if (old == kGEmpty) return 1;
    if (old == pr) return 0;
    i = (i + 1) & (kPairSet - 1);
  }
  return -1;
}

// The hot pairs, appended in row order by k_gb_main3 as their stored words and pair partitions
// (block b's at [hoffs[b * split], hoffs[(b + 1) * split])), split into the kHPairParts pair
// partitions, one block per main-pass block (1024 threads = kHPairParts): k_gb_main3 counted its
// pairs per partition, and `off` is the scan of those counts (partition-major).  Each tile of
// kPairTile pairs is sorted by partition in LDS and every partition's words of the tile are
// written as one contiguous run (order inside a partition does not matter to the deduplication).
constexpr int kPairBatch = 8;   // pairs a thread loads before using any (deduplication loops)
constexpr int kPairTile = 20480;  // 20 pairs per partition and tile on average (LDS: 6 bytes a pair)

__global__ __launch_bounds__(1024) void k_gb_pair_part(const uint32_t* __restrict__ words,
                                                       const uint16_t* __restrict__ parts,
                                                       const uint64_t* __restrict__ hoffs, int split,
                                                       const uint64_t* __restrict__ off, uint32_t* __restrict__ dst)
{
  __shared__ uint32_t tcnt[kHPairParts];
  __shared__ uint32_t tstart[kHPairParts];
  __shared__ unsigned long long cur[kHPairParts];
  __shared__ uint32_t stage_w[kPairTile];
  __shared__ uint16_t stage_q[kPairTile];
  __shared__ uint64_t smem[17];
  constexpr int R  = kPairTile / kHPairParts;  // pairs per thread and tile
  const int64_t b0 = (int64_t)hoffs[(int64_t)blockIdx.x * split];
  const int64_t b1 = (int64_t)hoffs[(int64_t)(blockIdx.x + 1) * split];
  cur[threadIdx.x] = off[(int64_t)threadIdx.x * gridDim.x + blockIdx.x];
  for (int64_t t0 = b0; t0 < b1; t0 += kPairTile) {
    const int tn     = (int)(b1 - t0 < kPairTile ? b1 - t0 : kPairTile);
    tcnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t w[R], part[R], rank[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {  // unconditional (past the tile: its first pair)
      const int i = threadIdx.x + j * kHPairParts;
      w[j]        = words[t0 + (i < tn ? i : 0)];
      part[j]     = parts[t0 + (i < tn ? i : 0)];
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = threadIdx.x + j * kHPairParts;
      rank[j]     = i < tn ? atomicAdd(&tcnt[part[j]], 1u) : 0u;
    }
    __syncthreads();
    uint64_t total;
    tstart[threadIdx.x] = (uint32_t)block_exclusive_scan(tcnt[threadIdx.x], smem, &total);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = threadIdx.x + j * kHPairParts;
      if (i < tn) {
        const uint32_t pos = tstart[part[j]] + rank[j];
        stage_w[pos]       = w[j];
        stage_q[pos]       = (uint16_t)part[j];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < tn; i += kHPairParts) {
      const uint32_t q = stage_q[i];
      dst[cur[q] + (uint64_t)(i - (int)tstart[q])] = stage_w[i];
    }
    __syncthreads();
    cur[threadIdx.x] += tcnt[threadIdx.x];
  }
}

// Inserts the word w (< kWMask + 1) into an LDS set of kPairSet32 words from slot i; 1 = new,
// 0 = present, -1 = no room.
__device__ __forceinline__ int word_set_insert(uint32_t* set, uint32_t w, uint32_t i)
{
  i &= kPairSet32 - 1;
  for (int t = 0; t < kPairSet32; ++t) {
    const uint32_t old = atomicCAS(&set[i], kWEmpty, w);
    if (old == kWEmpty) return 1;
    if (old == w) return 0;
    i = (i + 1) & (kPairSet32 - 1);
  }
  return -1;
}

// One block per coarse hot-pair partition q: split its stored words into kHFine fine partitions
// (within the coarse range of `fine`), then deduplicate each fine partition in an LDS set of
// words (in several passes when it holds more words than the set takes at ~70% load), counting
// new pairs per hot index in LDS (the hot index is recovered from q and the word through
// pair_unmix); each hot key's count is added to its slot once per block.
__global__ __launch_bounds__(1024) void k_gb_hot_distinct(const uint32_t* __restrict__ pairs,
                                                          const uint64_t* __restrict__ hoffsets, int nblk,
                                                          uint32_t* __restrict__ fine, int hot_cap,
                                                          HotList hot, uint32_t* __restrict__ part_cnt,
                                                          int32_t* __restrict__ err)
{
  __shared__ uint32_t set[kPairSet32];
  __shared__ uint32_t cnt[kPairHotMax];
  __shared__ uint32_t fh[kHFine];
  __shared__ uint32_t tcnt[kHFine];
  __shared__ uint32_t tstart[kHFine];
  __shared__ unsigned long long fcur[kHFine];
  __shared__ int64_t fstart[kHFine + 1];
  __shared__ int bad;
  const int q      = blockIdx.x;
  const int64_t b0 = (int64_t)hoffsets[(int64_t)q * nblk];
  const int64_t b1 = (int64_t)hoffsets[(int64_t)(q + 1) * nblk];
  const int nhot   = (int)min(min(*hot.count, (unsigned int)hot_cap), (unsigned int)kPairHotMax);
  for (int i = threadIdx.x; i < nhot; i += blockDim.x) cnt[i] = 0;
  if (threadIdx.x < kHFine) fh[threadIdx.x] = 0;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int64_t r = b0 + threadIdx.x; r < b1; r += (int64_t)kPairBatch * blockDim.x) {
    uint32_t pw[kPairBatch];  // a thread's loads in flight together (past the end: word b0)
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u) {
      const int64_t i = r + (int64_t)u * blockDim.x;
      pw[u]           = pairs[i < b1 ? i : b0];
    }
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u)
      if (r + (int64_t)u * blockDim.x < b1) atomicAdd(&fh[pw[u] >> (32 - kHFineBits)], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int f = 0; f < kHFine; ++f) {
      fstart[f] = run;
      fcur[f]   = (unsigned long long)(b0 + run);
      run += fh[f];
    }
    fstart[kHFine] = run;
  }
  __syncthreads();
  // Tiles of kFineTile words are sorted by fine partition in LDS (in the set's memory, which the
  // deduplication clears first) and each fine partition's run of a tile is stored contiguously.
  constexpr int kFineTile = 8192;              // 256 words per fine partition and tile on average
  constexpr int R         = kFineTile / 1024;  // words per thread and tile
  static_assert(kFineTile <= kPairSet32, "the tile is staged in the set");
  for (int64_t t0 = b0; t0 < b1; t0 += kFineTile) {
    const int tn = (int)(b1 - t0 < kFineTile ? b1 - t0 : kFineTile);
    if (threadIdx.x < kHFine) tcnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t pw[R], fb[R], rank[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {  // unconditional loads (past the tile: its first word), masked
      const int i = threadIdx.x + j * 1024;
      pw[j]       = pairs[t0 + (i < tn ? i : 0)];
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = threadIdx.x + j * 1024;
      fb[j]       = pw[j] >> (32 - kHFineBits);
      rank[j]     = i < tn ? atomicAdd(&tcnt[fb[j]], 1u) : 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (int f = 0; f < kHFine; ++f) {
        tstart[f] = run;
        run += tcnt[f];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (threadIdx.x + j * 1024 < tn) set[tstart[fb[j]] + rank[j]] = pw[j];
    __syncthreads();
    for (int i = threadIdx.x; i < tn; i += 1024) {
      const uint32_t v = set[i];
      const uint32_t f = v >> (32 - kHFineBits);
      fine[fcur[f] + (uint64_t)(i - (int)tstart[f])] = v;
    }
    __syncthreads();
    if (threadIdx.x < kHFine) fcur[threadIdx.x] += tcnt[threadIdx.x];
  }
  __syncthreads();
  for (int f = 0; f < kHFine; ++f) {
    const int64_t r0 = b0 + fstart[f], r1 = b0 + fstart[f + 1];
    const int passes = (int)((r1 - r0) / (int64_t)(0.7 * kPairSet32)) + 1;
    for (int pass = 0; pass < passes; ++pass) {
      for (int i = threadIdx.x; i < kPairSet32; i += blockDim.x) set[i] = kWEmpty;
      __syncthreads();
      constexpr int U = 2;  // a fine partition holds ~2 words per thread: its loads in one batch
      for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)U * blockDim.x) {
        uint32_t pw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t i = r + (int64_t)u * blockDim.x;
          pw[u]           = fine[i < r1 ? i : r0];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (r + (int64_t)u * blockDim.x >= r1) break;
          const uint32_t w = pw[u] & kWMask;
          if (passes > 1 && (int)((w >> 14) % (uint32_t)passes) != pass) continue;
          const int k = word_set_insert(set, w, w);
          if (k > 0) {
            const uint32_t h = (uint32_t)(pair_unmix(((uint64_t)q << 32) | pw[u]) >> 32);
            if (h < (uint32_t)nhot) atomicAdd(&cnt[h], 1u);
            else bad = 1;
          } else if (k < 0) bad = 1;
        }
      }
      __syncthreads();
    }
  }
  // this pair partition's new-pair counts per hot key (summed per key by k_gb_hot_dist_sum)
  for (int i = threadIdx.x; i < nhot; i += blockDim.x) part_cnt[(int64_t)q * hot_cap + i] = cnt[i];
  if (threadIdx.x == 0 && bad) atomicOr(err, 1);
}

// One block per hot key: its distinct count is the sum over the pair partitions.
__global__ __launch_bounds__(256) void k_gb_hot_dist_sum(const uint32_t* __restrict__ part_cnt, int nq, int hot_cap,
                                                         HotList hot, unsigned long long* __restrict__ dist)
{
  __shared__ uint64_t smem[17];
  const int nhot = (int)min(*hot.count, (unsigned int)hot_cap);
  for (int h = blockIdx.x; h < nhot; h += gridDim.x) {
    uint64_t s = 0;
    for (int q = threadIdx.x; q < nq; q += blockDim.x) s += part_cnt[(int64_t)q * hot_cap + h];
    const uint64_t tot = block_reduce_sum(s, smem);
    if (threadIdx.x == 0) dist[hot.gslot[h]] += tot;
  }
}

// One block per cold partition (whose rows hold every row of their keys): the (key, value)
// pairs of its rows with a valid value are split by pair hash into groups of at most ~70% of
// the LDS set (through `fine`, within the partition's range), each group is deduplicated in the
// LDS set, and each new pair counts for its key in an LDS key table that holds all of the
// partition's keys; every key's count is added to its slot at the end.
constexpr int kColdKeySlots = 4096;
constexpr int kColdGroups   = 128;

// the (key, value) pair of partition row r when its value is valid (from the appended pairs when
// k_gb_part3 wrote them: then every index below the partition's pair count holds one)
__device__ __forceinline__ bool cold_pair(const PartBufs& pb, int dcol, int dwidth, int64_t r, unsigned long long& pr)
{
  if (pb.cpairs != nullptr) {
    pr = pb.cpairs[r];
    return true;
  }
  const unsigned char* d = pb.rec + r * (int64_t)pb.stride;
  if (!((d[pb.flags_off] >> dcol) & 1)) return false;
  pr = (load_bits(d + pb.key_off, pb.key_width, 0) << 32) | (uint32_t)load_bits(d + pb.val_off[dcol], dwidth, 0);
  return true;
}

// group of a pair among ng by its hash: 32 hash bits scaled to [0, ng) by a multiply and a shift
// (no 64-bit division)
__device__ __forceinline__ uint32_t cold_group(uint64_t hsh, int ng)
{
  return (uint32_t)((((hsh >> 24) & 0xffffffffull) * (uint64_t)(uint32_t)ng) >> 32);
}

// The partition's keys and their dense slots (written by k_gb_part3) are loaded first; a key that
// spilled to the global hash table is looked up there at the end.
__global__ __launch_bounds__(1024) void k_gb_cold_distinct(PartBufs pb, int dcol, int dwidth, KeyTab t, int S, int tcap,
                                                           unsigned long long* __restrict__ fine,
                                                           unsigned long long* __restrict__ dist,
                                                           int32_t* __restrict__ err)
{
  __shared__ unsigned long long set[kPairSet];
  __shared__ unsigned long long kkeys[kColdKeySlots];
  __shared__ uint32_t kcnt[kColdKeySlots];
  __shared__ int32_t kslot[kColdKeySlots];
  __shared__ uint32_t gh[kColdGroups];
  __shared__ unsigned long long gcur[kColdGroups];
  __shared__ int64_t gstart[kColdGroups + 1];
  __shared__ int bad;
  const int part   = blockIdx.x;
  const int64_t r0 = (int64_t)pb.offsets[(int64_t)part * pb.nblk];
  const int64_t r1 = pb.cpairs != nullptr ? r0 + pb.cpair_count[part] : (int64_t)pb.offsets[(int64_t)(part + 1) * pb.nblk];
  // groups of at most 0.35 * kPairSet pairs when every row has a pair (short probe runs)
  const int ng = (int)((r1 - r0) / (int64_t)(0.35 * kPairSet)) + 1;
  if (ng > kColdGroups) {  // more rows than this kernel plans for: the caller falls back
    if (threadIdx.x == 0) atomicOr(err, 16);
    return;
  }
  for (int i = threadIdx.x; i < kColdKeySlots; i += blockDim.x) {
    kkeys[i] = kGEmpty;
    kcnt[i]  = 0;
    kslot[i] = -1;
  }
  if (threadIdx.x < kColdGroups) gh[threadIdx.x] = 0;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int q = threadIdx.x; q < S * tcap; q += blockDim.x) {
    const int64_t sl           = (int64_t)t.cap + 2 + (int64_t)part * S * tcap + q;
    const unsigned long long k = t.slots[sl];
    if (k == kGEmpty) continue;
    uint32_t i = (uint32_t)(mix64(k) & (kColdKeySlots - 1));
    int tries  = 0;
    for (; tries < kColdKeySlots; ++tries) {
      if (atomicCAS(&kkeys[i], kGEmpty, k) == kGEmpty) {
        kslot[i] = (int32_t)(sl - ((int64_t)t.cap + 2));
        break;
      }
      i = (i + 1) & (kColdKeySlots - 1);
    }
    if (tries == kColdKeySlots) bad |= 4;
  }
  __syncthreads();
  if (pb.cpairs != nullptr) {  // appended pairs: a thread's loads in flight together
    for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)kPairBatch * blockDim.x) {
      unsigned long long pr[kPairBatch];
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        const int64_t i = r + (int64_t)u * blockDim.x;
        pr[u]           = pb.cpairs[i < r1 ? i : r0];
      }
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u)
        if (r + (int64_t)u * blockDim.x < r1) atomicAdd(&gh[cold_group(pair_hash(pr[u]), ng)], 1u);
    }
  } else {
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      unsigned long long pr;
      if (cold_pair(pb, dcol, dwidth, r, pr)) atomicAdd(&gh[cold_group(pair_hash(pr), ng)], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int g = 0; g < ng; ++g) {
      gstart[g] = run;
      gcur[g]   = (unsigned long long)(r0 + run);
      run += gh[g];
    }
    gstart[ng] = run;
  }
  __syncthreads();
  if (pb.cpairs != nullptr) {
    for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)kPairBatch * blockDim.x) {
      unsigned long long pr[kPairBatch];
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        const int64_t i = r + (int64_t)u * blockDim.x;
        pr[u]           = pb.cpairs[i < r1 ? i : r0];
      }
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u)
        if (r + (int64_t)u * blockDim.x < r1) fine[atomicAdd(&gcur[cold_group(pair_hash(pr[u]), ng)], 1ull)] = pr[u];
    }
  } else {
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      unsigned long long pr;
      if (cold_pair(pb, dcol, dwidth, r, pr)) fine[atomicAdd(&gcur[cold_group(pair_hash(pr), ng)], 1ull)] = pr;
    }
  }
  __syncthreads();
  for (int g = 0; g < ng; ++g) {
    for (int i = threadIdx.x; i < kPairSet; i += blockDim.x) set[i] = kGEmpty;
    __syncthreads();
    const int64_t g0 = r0 + gstart[g], g1 = r0 + gstart[g + 1];
    for (int64_t r = g0 + threadIdx.x; r < g1; r += (int64_t)kPairBatch * blockDim.x) {
      unsigned long long prs[kPairBatch];
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        const int64_t i = r + (int64_t)u * blockDim.x;
        prs[u]          = fine[i < g1 ? i : g0];
      }
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        if (r + (int64_t)u * blockDim.x >= g1) continue;
        const unsigned long long pr = prs[u];
        const int k                 = pair_set_insert(set, pr, pair_hash(pr));
        if (k < 0) bad |= 2;
        if (k <= 0) continue;
        const unsigned long long kb = pr >> 32;
        uint32_t i                  = (uint32_t)(mix64(kb) & (kColdKeySlots - 1));
        int found                   = -1;
        for (int tries = 0; tries < kColdKeySlots; ++tries) {
          unsigned long long cur = kkeys[i];
          if (cur == kGEmpty) cur = atomicCAS(&kkeys[i], kGEmpty, kb);
          if (cur == kGEmpty || cur == kb) {
            found = (int)i;
            break;
          }
          i = (i + 1) & (kColdKeySlots - 1);
        }
        if (found < 0) bad |= 4;
        else atomicAdd(&kcnt[found], 1u);
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kColdKeySlots; i += blockDim.x) {
    if (kcnt[i] == 0) continue;
    if (kslot[i] >= 0) {  // this block is the only writer of the key's count
      dist[(int64_t)t.cap + 2 + kslot[i]] = kcnt[i];
      continue;
    }
    const int64_t sl = gb_find_bits(kkeys[i], t);  // a key that spilled to the hash table
    if (sl >= 0) atomicAdd(&dist[sl], (unsigned long long)kcnt[i]);
    else bad |= 8;
  }
  __syncthreads();
  if (threadIdx.x == 0 && bad) atomicOr(err, bad);
}

// The same for partitions whose pairs k_gb_part3 appended (one sub-pass): each pair carries its
// key's accumulator slot instead of the key -- the key's dense slot, or its global slot when it
// spilled -- so a new pair counts in an LDS array over the partition's dense slots (dynamic LDS,
// tcap words; a spilled key's count goes straight to its slot) and no key table is needed.
__global__ __launch_bounds__(1024) void k_gb_cold_distinct_slots(PartBufs pb, KeyTab t, int tcap,
                                                                 unsigned long long* __restrict__ fine,
                                                                 unsigned long long* __restrict__ dist,
                                                                 int32_t* __restrict__ err)
{
  extern __shared__ uint32_t dcnt[];  // new pairs per dense slot of this partition
  __shared__ unsigned long long set[kPairSet];
  __shared__ uint32_t gh[kColdGroups];
  __shared__ unsigned long long gcur[kColdGroups];
  __shared__ int64_t gstart[kColdGroups + 1];
  __shared__ int bad;
  const int part   = blockIdx.x;
  const int64_t r0 = (int64_t)pb.offsets[(int64_t)part * pb.nblk];
  const int64_t r1 = r0 + pb.cpair_count[part];
  const int64_t d0 = (int64_t)t.cap + 2 + (int64_t)part * tcap;
  // groups of at most 35% of the set: an insert's probe run (and so a wave's longest, which the
  // wave waits for) stays short
  const int ng     = (int)((r1 - r0) / (int64_t)(0.35 * kPairSet)) + 1;
  if (ng > kColdGroups) {  // more rows than this kernel plans for: the caller falls back
    if (threadIdx.x == 0) atomicOr(err, 16);
    return;
  }
  for (int i = threadIdx.x; i < tcap; i += blockDim.x) dcnt[i] = 0;
  if (threadIdx.x < kColdGroups) gh[threadIdx.x] = 0;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)kPairBatch * blockDim.x) {
    unsigned long long pr[kPairBatch];
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u) {
      const int64_t i = r + (int64_t)u * blockDim.x;
      pr[u]           = pb.cpairs[i < r1 ? i : r0];
    }
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u)
      if (r + (int64_t)u * blockDim.x < r1) atomicAdd(&gh[cold_group(pair_hash(pr[u]), ng)], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int g = 0; g < ng; ++g) {
      gstart[g] = run;
      gcur[g]   = (unsigned long long)(r0 + run);
      run += gh[g];
    }
    gstart[ng] = run;
  }
  __syncthreads();
  if (ng > 1) {  // one group: deduplicate the appended pairs where they are
    for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)kPairBatch * blockDim.x) {
      unsigned long long pr[kPairBatch];
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u) {
        const int64_t i = r + (int64_t)u * blockDim.x;
        pr[u]           = pb.cpairs[i < r1 ? i : r0];
      }
#pragma unroll
      for (int u = 0; u < kPairBatch; ++u)
        if (r + (int64_t)u * blockDim.x < r1) fine[atomicAdd(&gcur[cold_group(pair_hash(pr[u]), ng)], 1ull)] = pr[u];
    }
    __syncthreads();
  }
  const unsigned long long* src = ng > 1 ? fine : pb.cpairs;
  for (int g = 0; g < ng; ++g) {
    for (int i = threadIdx.x; i < kPairSet; i += blockDim.x) set[i] = kGEmpty;
    __syncthreads();
    const int64_t g0 = r0 + gstart[g], g1 = r0 + gstart[g + 1];
    constexpr int U  = 4;
    for (int64_t r = g0 + threadIdx.x; r < g1; r += (int64_t)U * blockDim.x) {
      unsigned long long prs[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = r + (int64_t)u * blockDim.x;
        prs[u]          = src[i < g1 ? i : g0];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (r + (int64_t)u * blockDim.x >= g1) break;
        const int k = pair_set_insert(set, prs[u], pair_hash(prs[u]));
        if (k < 0) bad |= 2;
        if (k <= 0) continue;
        const int64_t sl = (int64_t)(prs[u] >> 32);
        if (sl >= d0 && sl < d0 + tcap) atomicAdd(&dcnt[sl - d0], 1u);
        else atomicAdd(&dist[sl], 1ull);  // a key that spilled to the hash table
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < tcap; i += blockDim.x)  // this block is the only writer of its dense slots
    if (dcnt[i] != 0) dist[d0 + i] = dcnt[i];
  if (threadIdx.x == 0 && bad) atomicOr(err, bad);
}

template <bool GENERIC>
__global__ __launch_bounds__(256) void k_gb_second_pass(DKeys keys, int64_t n, KeyTab t, ColDev c, double* acc,
                                                        const double* mean)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    if (!valid_at(c.valid, r)) continue;
    int64_t s = gb_find<GENERIC>(keys, r, t);
    if (s < 0) continue;
    double x = bits_as_double(load_bits(c.data, c.width, r), c.type, c.width);
    if (mean) {
      double d = x - mean[s];
      atomicAdd(&acc[s], d * d);
    } else {
      atomicAdd(&acc[s], x);
    }
  }
}

// ---- sampling: table size and hot keys ---------------------------------------------------
template <bool GENERIC>
__global__ void k_gb_sample(DKeys keys, int64_t n, int64_t S, int64_t stride, KeyTab t, unsigned int* cnt)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = i * stride;
    if (r >= n) continue;
    int64_t s = gb_find_or_insert<GENERIC>(keys, r, t);
    if (s >= 0) atomicAdd(&cnt[s], 1u);
  }
}

// The same sample for a single fixed-width key, counted first in each block's LDS table (1024
// sampled rows per block), so the global table takes one insert and one add per distinct key
// per block instead of one per sampled row (the heaviest keys would serialise on them).
constexpr int kSampleBlockRows = 1024;  // 64 blocks for the 65536-row sample
constexpr int kSampleSlots     = 2048;
__global__ __launch_bounds__(1024) void k_gb_sample_lds(DKeys keys, int64_t n, int64_t S, int64_t stride, KeyTab t,
                                                        unsigned int* cnt)
{
  __shared__ unsigned long long tk[kSampleSlots];
  __shared__ uint32_t tc[kSampleSlots];
  __shared__ uint32_t s_null, s_empty;
  for (int i = threadIdx.x; i < kSampleSlots; i += blockDim.x) {
    tk[i] = kGEmpty;
    tc[i] = 0;
  }
  if (threadIdx.x == 0) s_null = s_empty = 0;
  __syncthreads();
  const DCol& c      = keys.c[0];
  const int64_t i0   = (int64_t)blockIdx.x * kSampleBlockRows;
  const int64_t iend = i0 + kSampleBlockRows < S ? i0 + kSampleBlockRows : S;
  constexpr int U = kSampleBlockRows / 1024;  // rows per thread, loaded together
  static_assert(U * 1024 == kSampleBlockRows, "one pass of 1024 threads");
  unsigned long long ks[U];
  uint32_t kv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {  // branch-free loads (rows outside the sample load row 0)
    const int64_t i  = i0 + threadIdx.x + (int64_t)u * 1024;
    const int64_t r  = i * stride;
    const bool in    = i < iend && r < n;
    const int64_t rr = in ? r : 0;
    kv[u]            = valid_bit_nb(c.valid, c.data, rr) | (in ? 0u : 2u);  // bit 1: not a sampled row
    ks[u]            = load_bits_nb(c.data, c.width, rr);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (kv[u] & 2u) continue;
    if (!(kv[u] & 1u)) {
      atomicAdd(&s_null, 1u);
      continue;
    }
    const unsigned long long k = ks[u];
    if (k == kGEmpty) {
      atomicAdd(&s_empty, 1u);
      continue;
    }
    uint32_t j = (uint32_t)(mix64(k) & (kSampleSlots - 1));
    while (true) {  // at most kSampleBlockRows keys in kSampleSlots slots: never full
      unsigned long long cur = tk[j];
      if (cur == kGEmpty) cur = atomicCAS(&tk[j], kGEmpty, k);
      if (cur == kGEmpty || cur == k) break;
      j = (j + 1) & (kSampleSlots - 1);
    }
    atomicAdd(&tc[j], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kSampleSlots; j += blockDim.x) {
    if (tc[j] == 0) continue;
    const int64_t sl = gb_insert_bits<false>(tk[j], true, t);
    if (sl >= 0) atomicAdd(&cnt[sl], tc[j]);
  }
  if (threadIdx.x == 0) {
    if (s_null) {
      const int64_t sl = gb_insert_bits<false>(0, false, t);
      if (sl >= 0) atomicAdd(&cnt[sl], s_null);
    }
    if (s_empty) {
      const int64_t sl = gb_insert_bits<false>(kGEmpty, true, t);
      if (sl >= 0) atomicAdd(&cnt[sl], s_empty);
    }
  }
}

// stats[0] = distinct, [1] = singletons, [2] = doubletons, [3 + c] = keys seen c times (c < 64),
// [67] = ~(smallest key), [68] = largest key, over the table's ordinary slots (i < scap; stats
// starts zeroed, so [67] > ~[68] means no ordinary key)
__global__ __launch_bounds__(256) void k_gb_sample_stats(const unsigned int* cnt, const unsigned long long* slots,
                                                         int64_t scap, int64_t n, unsigned long long* stats)
{
  __shared__ unsigned int loc[3 + 64];
  __shared__ unsigned long long mm[2];
  for (int i = threadIdx.x; i < 3 + 64; i += blockDim.x) loc[i] = 0;
  if (threadIdx.x < 2) mm[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long nmin = 0, kmax = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    unsigned c = cnt[i];
    if (c == 0) continue;
    atomicAdd(&loc[0], 1u);
    if (c == 1) atomicAdd(&loc[1], 1u);
    if (c == 2) atomicAdd(&loc[2], 1u);
    atomicAdd(&loc[3 + (c < 63 ? c : 63)], 1u);
    if (i < scap) {
      const unsigned long long k = slots[i];
      nmin                       = max(nmin, ~k);
      kmax                       = max(kmax, k);
    }
  }
  if (nmin) atomicMax(&mm[0], nmin);
  if (kmax) atomicMax(&mm[1], kmax);
  __syncthreads();
  for (int i = threadIdx.x; i < 3 + 64; i += blockDim.x)
    if (loc[i]) atomicAdd(&stats[i], (unsigned long long)loc[i]);
  if (threadIdx.x < 2 && mm[threadIdx.x]) atomicMax(&stats[3 + 64 + threadIdx.x], mm[threadIdx.x]);
}

__global__ void k_gb_hot_select(const unsigned long long* sslots, const unsigned int* cnt, int64_t scap, unsigned int thr,
                                int cap, HotList hot)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < scap + 2; i += (int64_t)gridDim.x * blockDim.x) {
    if (cnt[i] < thr) continue;
    unsigned int h = atomicAdd(hot.count, 1u);
    if (h >= (unsigned)cap) continue;
    hot.keys[h]    = sslots[i];
    hot.special[h] = i == scap ? 1 : (i == scap + 1 ? 2 : 0);
  }
}

__global__ void k_gb_hot_insert(KeyTab t, HotList hot, int cap)
{
  int nh = (int)min(*hot.count, (unsigned int)cap);
  for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += gridDim.x * blockDim.x) {
    int sp = hot.special[h];
    int64_t s;
    if (sp == 1) {
      s             = (int64_t)t.cap;
      t.slots[s]    = 0;
    } else if (sp == 2) {
      s             = (int64_t)t.cap + 1;
      t.slots[s]    = 0;
    } else {
      unsigned long long key = hot.keys[h];
      uint64_t i             = mix64(key) & t.mask;
      while (true) {
        unsigned long long old = atomicCAS(&t.slots[i], kGEmpty, key);
        if (old == kGEmpty) {
          atomicAdd(t.nkeys, 1ull);
          break;
        }
        if (old == key) break;
        i = (i + 1) & t.mask;
      }
      s = (int64_t)i;
    }
    hot.gslot[h] = (int32_t)s;
  }
}

// ---- finalisation ------------------------------------------------------------------------
__global__ void k_gb_occupied(const unsigned long long* slots, int64_t nslot, uint32_t* occ)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nslot; i += (int64_t)gridDim.x * blockDim.x)
    occ[i] = slots[i] != kGEmpty;
}

__global__ void k_gb_compact(const unsigned long long* slots, int64_t nslot, const uint64_t* pos, bool generic,
                             int32_t* slot_of_group, int32_t* rep)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nslot; i += (int64_t)gridDim.x * blockDim.x) {
    if (slots[i] == kGEmpty) continue;
    uint64_t g       = pos[i];
    slot_of_group[g] = (int32_t)i;
    if (generic) rep[g] = (int32_t)(uint32_t)slots[i];
  }
}

struct OutDev {
  int32_t kind;  // LDF_AGG_* ; -1 = the single fixed-width key column
  int32_t mode;
  int32_t in_type;
  int32_t out_type;
  const unsigned long long* main;  // sum / min / max / product / distinct count
  const unsigned long long* l1;
  const unsigned long long* l2;
  const unsigned long long* vcount;
  const unsigned long long* rows;
  const uint32_t* flag;
  const double* fsum;  // fallback sum or M2
  int32_t use_fsum;
  void* out;
  uint64_t* valid;
  unsigned long long* nulls;
};

struct OutPlan {
  OutDev o[kMaxOut];
  int32_t n;
  uint64_t cap;
  const unsigned long long* slots;
};

// One wave covers 64 consecutive groups: each output column's validity is one ballot.
__global__ __launch_bounds__(256) void k_gb_finalize(OutPlan fp, const int32_t* __restrict__ slot_of_group, int64_t ng)
{
  const int lane = threadIdx.x & 63;
  for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < ng;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g      = base + lane;
    const bool in        = g < ng;
    const int64_t s      = in ? slot_of_group[g] : 0;
    const int64_t nrange = ng - base < 64 ? ng - base : 64;
    for (int k = 0; k < fp.n; ++k) {
      const OutDev& o = fp.o[k];
      bool ok         = true;
      if (in) {
        uint64_t c = o.vcount ? o.vcount[s] : 0;
        switch (o.kind) {
          case -1: {  // key
            uint64_t key = (s == (int64_t)fp.cap) ? kGEmpty : (s == (int64_t)fp.cap + 1 ? 0 : fp.slots[s]);
            ok           = s != (int64_t)fp.cap + 1;
            store_bits(o.out, o.out_type, g, key);
            break;
          }
          case LDF_AGG_COUNT: {
            uint64_t v = o.mode == LDF_COUNT_ALL ? o.rows[s] : (o.mode == LDF_COUNT_ONLY_NULL ? o.rows[s] - c : c);
            static_cast<int64_t*>(o.out)[g] = (int64_t)v;
            break;
          }
          case LDF_AGG_COUNT_ALL: static_cast<int64_t*>(o.out)[g] = (int64_t)o.rows[s]; break;
          case LDF_AGG_COUNT_DISTINCT: {
            bool has_null = o.rows && o.rows[s] > c;
            int64_t v     = (int64_t)o.main[s];
            if (o.mode == LDF_COUNT_ALL) v += has_null ? 1 : 0;
            if (o.mode == LDF_COUNT_ONLY_NULL) v = has_null ? 1 : 0;
            static_cast<int64_t*>(o.out)[g] = v;
            break;
          }
          case LDF_AGG_SUM:
            ok = c > 0;
            if (o.out_type == LDF_TYPE_DOUBLE) {
              double v = !ok ? 0.0 : (o.use_fsum ? o.fsum[s] : fixed192_to_double(o.main[s], o.l1[s], o.l2[s]));
              static_cast<double*>(o.out)[g] = v;
            } else {
              static_cast<uint64_t*>(o.out)[g] = ok ? o.main[s] : 0;
            }
            break;
          case LDF_AGG_MEAN: {
            ok       = c > 0;
            double v = !ok ? 0.0 : (o.use_fsum ? o.fsum[s] : fixed192_to_double(o.main[s], o.l1[s], o.l2[s]));
            static_cast<double*>(o.out)[g] = ok ? v / (double)c : 0.0;
            break;
          }
          case LDF_AGG_VARIANCE:
          case LDF_AGG_STDDEV: {
            ok                             = c > 0;
            double v                       = ok ? o.fsum[s] / (double)c : 0.0;
            static_cast<double*>(o.out)[g] = o.kind == LDF_AGG_STDDEV ? sqrt(v) : v;
            break;
          }
          case LDF_AGG_PRODUCT:
            ok                               = c > 0;
            static_cast<uint64_t*>(o.out)[g] = ok ? o.main[s] : 0;
            break;
          case LDF_AGG_MIN:
          case LDF_AGG_MAX: {
            ok              = c > 0;
            uint64_t ord    = o.main[s];
            uint64_t ident  = o.kind == LDF_AGG_MIN ? ~0ull : 0ull;
            const bool fl   = o.in_type == LDF_TYPE_DOUBLE || o.in_type == LDF_TYPE_FLOAT;
            uint64_t bits;
            if (fl) {
              double d;
              if (ord == ident) d = __longlong_as_double(0x7ff8000000000000ll);  // only NaN seen
              else d = __longlong_as_double((long long)((ord >> 63) ? (ord & 0x7fffffffffffffffull) : ~ord));
              bits = o.in_type == LDF_TYPE_FLOAT ? (uint64_t)__float_as_uint((float)d) : (uint64_t)__double_as_longlong(d);
            } else if (is_unsigned_type(o.in_type)) {
              bits = ord;
            } else {
              bits = ord ^ 0x8000000000000000ull;
            }
            store_bits(o.out, o.out_type, g, ok ? bits : 0);
            break;
          }
          default: break;
        }
      } else {
        ok = false;
      }
      if (o.valid != nullptr) {
        uint64_t b = __ballot(ok);
        if (lane == 0) {
          o.valid[base >> 6] = b;
          uint64_t z         = (uint64_t)nrange - (uint64_t)__popcll(b);
          if (z) atomicAdd(o.nulls, (unsigned long long)z);
        }
      }
    }
  }
}

// segment fill: seg[i] = (offset, length, value) over one u64 buffer
struct FillSeg {
  uint64_t off, len, val;
};
constexpr int kMaxFillSeg = 64;
struct FillPlan {
  FillSeg s[kMaxFillSeg];
  int32_t n;
};

__global__ void k_fill_segments(unsigned long long* buf, FillPlan fp)
{
  const int seg = blockIdx.y;
  if (seg >= fp.n) return;
  const FillSeg f = fp.s[seg];
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < f.len; i += (uint64_t)gridDim.x * blockDim.x)
    buf[f.off + i] = f.val;
}

// ---- host --------------------------------------------------------------------------------
static int32_t agg_out_type(int32_t kind, int32_t in_type)
{
  switch (kind) {
    case LDF_AGG_COUNT:
    case LDF_AGG_COUNT_ALL:
    case LDF_AGG_COUNT_DISTINCT: return LDF_TYPE_INT64;
    case LDF_AGG_MEAN:
    case LDF_AGG_VARIANCE:
    case LDF_AGG_STDDEV: return LDF_TYPE_DOUBLE;
    case LDF_AGG_SUM:
    case LDF_AGG_PRODUCT:
      if (is_float(in_type)) return LDF_TYPE_DOUBLE;
      if (is_unsigned_int(in_type)) return LDF_TYPE_UINT64;
      return LDF_TYPE_INT64;
    case LDF_AGG_MIN:
    case LDF_AGG_MAX: return in_type;
    default: return -1;
  }
}

struct HostAcc {
  int32_t kind;
  int32_t col;  // input column of the table (-1 for ACC_ROWS)
  bool global_rows = true;
};

struct Lowered {
  std::vector<HostAcc> acc;
  struct AggRef {
    int main = -1, vcount = -1, rows = -1;
  };
  std::vector<AggRef> ref;
};

static int find_or_add(std::vector<HostAcc>& v, int32_t kind, int32_t col)
{
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i].kind == kind && v[i].col == col) return (int)i;
  v.push_back({kind, col});
  return (int)v.size() - 1;
}

static Lowered lower(const ldf_table_view* tv, const ldf_agg* aggs, int32_t naggs)
{
  Lowered L;
  L.ref.resize(naggs);
  for (int i = 0; i < naggs; ++i) {
    const ldf_agg& g = aggs[i];
    auto& R          = L.ref[i];
    const int c      = g.in_col;
    switch (g.kind) {
      case LDF_AGG_COUNT_ALL: R.rows = find_or_add(L.acc, ACC_ROWS, -1); break;
      case LDF_AGG_COUNT:
        if (g.count_mode != LDF_COUNT_ONLY_VALID) R.rows = find_or_add(L.acc, ACC_ROWS, -1);
        if (g.count_mode != LDF_COUNT_ALL) R.vcount = find_or_add(L.acc, ACC_VCOUNT, c);
        break;
      case LDF_AGG_SUM:
        R.vcount = find_or_add(L.acc, ACC_VCOUNT, c);
        R.main   = find_or_add(L.acc, is_float(tv->cols[c].type_id) ? ACC_FSUM : ACC_ISUM, c);
        break;
      case LDF_AGG_MEAN:
      case LDF_AGG_VARIANCE:
      case LDF_AGG_STDDEV:
        R.vcount = find_or_add(L.acc, ACC_VCOUNT, c);
        R.main   = find_or_add(L.acc, ACC_FSUM, c);
        break;
      case LDF_AGG_MIN:
      case LDF_AGG_MAX:
        R.vcount = find_or_add(L.acc, ACC_VCOUNT, c);
        R.main   = find_or_add(L.acc, g.kind == LDF_AGG_MIN ? ACC_MIN : ACC_MAX, c);
        break;
      case LDF_AGG_PRODUCT:
        R.vcount = find_or_add(L.acc, ACC_VCOUNT, c);
        R.main   = find_or_add(L.acc, ACC_PROD, c);
        break;
      case LDF_AGG_COUNT_DISTINCT:
        R.main = find_or_add(L.acc, ACC_DISTINCT, c);
        if (g.count_mode != LDF_COUNT_ONLY_VALID) {
          R.rows   = find_or_add(L.acc, ACC_ROWS, -1);
          R.vcount = find_or_add(L.acc, ACC_VCOUNT, c);
        }
        break;
      default: break;
    }
  }
  return L;
}

// A second stream per (host thread, device) with two events, for work that runs beside the
// caller's stream inside one call (default priority, non-blocking; created on first use).
struct SideStream {
  hipStream_t s    = nullptr;
  hipEvent_t fork  = nullptr;
  hipEvent_t join  = nullptr;
};

static SideStream& side_stream()
{
  thread_local std::map<int, SideStream> per_device;
  int dev = 0;
  LDF_HIP_TRY(hipGetDevice(&dev));
  SideStream& ss = per_device[dev];
  if (ss.s == nullptr) {
    LDF_HIP_TRY(hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking));
    LDF_HIP_TRY(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming));
    LDF_HIP_TRY(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming));
  }
  return ss;
}

// k_gb_main3 for the plan's number of value columns (a template argument: the row loads are
// unrolled over exactly the columns read); the LDS limit is raised once per instantiation
template <bool DPART, int NC>
static void launch_main3_nc(dim3 grid, int lds, hipStream_t s, const DKeys& keys, int64_t n, const GbPlan& p,
                            const HotList& hot, const PartBufs& pb, const HotParts& hp, const uint16_t* codes)
{
  static bool attr = [] {
    LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_main3<DPART, NC>, hipFuncAttributeMaxDynamicSharedMemorySize, kMain3Lds));
    return true;
  }();
  (void)attr;
  hipLaunchKernelGGL((k_gb_main3<DPART, NC>), grid, dim3(kGbThreads), lds, s, keys, n, p, hot, pb, hp, codes);
  LDF_HIP_TRY(hipGetLastError());
}

template <bool DPART>
static void launch_main3_d(int nc, dim3 grid, int lds, hipStream_t s, const DKeys& keys, int64_t n, const GbPlan& p,
                           const HotList& hot, const PartBufs& pb, const HotParts& hp, const uint16_t* codes)
{
  switch (nc) {
    case 0: launch_main3_nc<DPART, 0>(grid, lds, s, keys, n, p, hot, pb, hp, codes); break;
    case 1: launch_main3_nc<DPART, 1>(grid, lds, s, keys, n, p, hot, pb, hp, codes); break;
    case 2: launch_main3_nc<DPART, 2>(grid, lds, s, keys, n, p, hot, pb, hp, codes); break;
    case 3: launch_main3_nc<DPART, 3>(grid, lds, s, keys, n, p, hot, pb, hp, codes); break;
    default: launch_main3_nc<DPART, 4>(grid, lds, s, keys, n, p, hot, pb, hp, codes); break;
  }
}

static void launch_main3(bool dpart, int nc, dim3 grid, int lds, hipStream_t s, const DKeys& keys, int64_t n,
                         const GbPlan& p, const HotList& hot, const PartBufs& pb, const HotParts& hp,
                         const uint16_t* codes)
{
  if (dpart) launch_main3_d<true>(nc, grid, lds, s, keys, n, p, hot, pb, hp, codes);
  else launch_main3_d<false>(nc, grid, lds, s, keys, n, p, hot, pb, hp, codes);
}

template <bool GENERIC>
static void launch_main(const DKeys& keys, int64_t n, const KeyTab& t, const GbPlan& p, const HotList& hot, int nblk,
                        hipStream_t s)
{
  static bool attr = false;
  if (!attr) {
    LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_main<GENERIC>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr = true;
  }
  ProfScope prof("gb_aggregate", s);
  hipLaunchKernelGGL(k_gb_main<GENERIC>, dim3(nblk), dim3(kGbThreads), p.lds_bytes, s, keys, n, t, p, hot);
  LDF_HIP_TRY(hipGetLastError());
}

template <bool GENERIC>
static void launch_second(const DKeys& keys, int64_t n, const KeyTab& t, const ColDev& c, double* acc, const double* mean,
                          hipStream_t s)
{
  hipLaunchKernelGGL(k_gb_second_pass<GENERIC>, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, s, keys, n, t, c, acc, mean);
  LDF_HIP_TRY(hipGetLastError());
}

__global__ void k_gb_means(const unsigned long long* a0, const unsigned long long* a1, const unsigned long long* a2,
                           const unsigned long long* cnt, const double* fsum, int use_fsum, int64_t nslot, double* mean)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nslot; i += (int64_t)gridDim.x * blockDim.x) {
    unsigned long long c = cnt[i];
    double s             = use_fsum ? fsum[i] : fixed192_to_double(a0[i], a1[i], a2[i]);
    mean[i]              = c ? s / (double)c : 0.0;
  }
}

static ColDev to_coldev(const ldf_column_view& c)
{
  ColDev d;
  d.data    = c.data;
  d.valid   = c.validity;
  d.offsets = c.offsets;
  d.type    = c.type_id;
  d.width   = type_width(c.type_id);
  return d;
}

void run_groupby(const ldf_table_view* tv, const int32_t* key_idx, int32_t nkeys, const ldf_agg* aggs, int32_t naggs,
                        Alloc& a, ldf_table_out* out)
{
  hipStream_t s = a.stream();
  LDF_EXPECTS(tv != nullptr && out != nullptr && out->cols != nullptr, "groupby: NULL argument");
  LDF_EXPECTS(nkeys >= 1, "must pass at least one key column");
  LDF_EXPECTS(nkeys <= kMaxKeyCols, "groupby: at most 8 key columns");
  LDF_EXPECTS(naggs >= 0, "groupby: negative aggregation count");
  LDF_EXPECTS(out->ncols >= nkeys + naggs, "groupby: output table has too few column slots");
  int64_t n = tv->ncols > 0 ? tv->cols[0].length : 0;
  for (int i = 0; i < tv->ncols; ++i) validate_view(tv->cols[i], n, "groupby");
  LDF_EXPECTS(n < INT32_MAX, "groupby: limited to 2^31-1 rows per call");
  DKeys keys;
  keys.n = nkeys;
  for (int i = 0; i < nkeys; ++i) {
    LDF_EXPECTS(key_idx[i] >= 0 && key_idx[i] < tv->ncols, "groupby: key index out of range");
    keys.c[i] = to_dcol(tv->cols[key_idx[i]]);
  }
  const bool generic = !(nkeys == 1 && keys.c[0].width > 0);
  for (int i = 0; i < naggs; ++i) {
    const ldf_agg& g = aggs[i];
    LDF_EXPECTS(g.kind >= LDF_AGG_SUM && g.kind <= LDF_AGG_COUNT_ALL, "Unsupported groupby aggregation");
    LDF_EXPECTS(g.kind != LDF_AGG_APPROX_MEDIAN, "groupby: approximate_median goes through run_groupby_with_medians");
    if (g.kind == LDF_AGG_COUNT_ALL) continue;
    LDF_EXPECTS(g.in_col >= 0 && g.in_col < tv->ncols, "groupby: aggregation input column out of range");
    if (g.kind == LDF_AGG_COUNT || g.kind == LDF_AGG_COUNT_DISTINCT) {
      LDF_EXPECTS(g.count_mode >= LDF_COUNT_ONLY_VALID && g.count_mode <= LDF_COUNT_ALL, "Unsupported count mode");
      continue;
    }
    if (type_width(tv->cols[g.in_col].type_id) == 0)
      throw Error(LDF_E_UNSUPPORTED, "groupby: numeric aggregation of a string column");
  }
  Lowered low = lower(tv, aggs, naggs);
  const int nacc = (int)low.acc.size();

  // ---- sample: table size estimate and hot keys ----
  const int64_t S  = n < kSampleRows ? n : kSampleRows;
  uint64_t scap    = next_pow2((uint64_t)std::max<int64_t>(2 * S, 64));
  Buf<unsigned long long> sslots(a, scap + 2);
  Buf<unsigned int> scnt(a, scap + 2);
  Buf<unsigned long long> smisc(a, 2 + 3 + 64 + 2);  // nkeys, overflow, stats
  {
    const FillJob jobs[3] = {{sslots.get(), kGEmpty, (int64_t)scap + 2, 8},
                             {scnt.get(), 0, (int64_t)scap + 2, 4},
                             {smisc.get(), 0, 2 + 3 + 64 + 2, 8}};
    fill_many(jobs, 3, s);
  }
  {
    KeyTab st{sslots.get(), scap - 1, scap, smisc.get(), reinterpret_cast<int32_t*>(smisc.get() + 1)};
    if (S > 0) {
      int64_t stride = n / S;
      if (generic)
        hipLaunchKernelGGL(k_gb_sample<true>, dim3(grid_for(S, 256)), dim3(256), 0, s, keys, n, S, stride, st, scnt.get());
      else
        hipLaunchKernelGGL(k_gb_sample_lds, dim3((unsigned)((S + kSampleBlockRows - 1) / kSampleBlockRows)), dim3(1024), 0,
                           s, keys, n, S, stride, st, scnt.get());
      LDF_HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(k_gb_sample_stats, dim3(grid_for(scap + 2, 256)), dim3(256), 0, s, scnt.get(), sslots.get(),
                         (int64_t)scap, (int64_t)scap + 2, smisc.get() + 2);
      LDF_HIP_TRY(hipGetLastError());
    }
  }
  unsigned long long stats[3 + 64 + 2] = {};
  copy_to_host(stats, smisc.get() + 2, sizeof(stats), s);
  const uint64_t smin = ~stats[3 + 64], smax = stats[3 + 64 + 1];  // sampled key range (smin > smax: none)
  uint64_t est;
  {
    double d = (double)stats[0], f1 = (double)stats[1];
    if (S == n) est = (uint64_t)d;
    else {
      double e = (f1 > 0.5 * (double)S) ? d * (double)n / (double)S : std::sqrt((double)n / (double)S) * f1 + (d - f1);
      e        = std::min(std::max(e, d), (double)n);
      est      = (uint64_t)e;
    }
  }

  // ---- LDS plan ----
  GbPlan plan;
  std::memset(&plan, 0, sizeof(plan));
  std::vector<int32_t> colmap;  // plan column -> table column
  for (auto& h : low.acc)
    if (h.col >= 0 && std::find(colmap.begin(), colmap.end(), h.col) == colmap.end()) colmap.push_back(h.col);
  LDF_EXPECTS(nacc <= kMaxAcc && (int)colmap.size() <= kMaxCols,
              "groupby: more than 16 accumulators or 4 distinct input columns per call is not supported");
  int acc_bytes = 8;  // rows + flags
  for (auto& h : low.acc) {
    switch (h.kind) {
      case ACC_VCOUNT:
      case ACC_DISTINCT: acc_bytes += 4; break;
      case ACC_ISUM:
      case ACC_MIN:
      case ACC_MAX: acc_bytes += 8; break;
      case ACC_FSUM: acc_bytes += 48; break;
      default: break;
    }
  }
  int hot_cap = 0;
  if (!generic) {
    const int budget = 124 * 1024;  // leaves room for the partition staging of k_gb_main3
    int most         = 2048;        // LDF_GROUPBY_HOT_CAP lowers the ceiling (a tuning knob)
    if (const char* e = std::getenv("LDF_GROUPBY_HOT_CAP")) most = std::max(64, std::min(2048, std::atoi(e)));
    for (auto& h : low.acc)  // a hot pair (hot index, value) of COUNT_DISTINCT has at most 42 bits
      if (h.kind == ACC_DISTINCT) most = std::min(most, kPairHotMax);
    for (int h = 2048; h >= 64; h >>= 1)
      if (h <= most && h * (acc_bytes + 20) + 64 <= budget) {
        hot_cap = h;
        break;
      }
  }
  unsigned int thr = 2;
  {
    uint64_t above = 0;
    for (int c = 63; c >= 2; --c) {
      above += stats[3 + c];
      if (above > (uint64_t)hot_cap) {
        thr = (unsigned)c + 1;
        break;
      }
    }
  }
  int32_t off = 0;
  auto take   = [&](int bytes) {
    int32_t o = off;
    off += (bytes + 15) & ~15;
    return o;
  };
  if (hot_cap > 0) {
    plan.rows_off  = take(hot_cap * 4);
    plan.flags_off = take(hot_cap * 4);
  }
  int flag_bits = 0;
  std::memset(plan.cacc, -1, sizeof(plan.cacc));
  for (int i = 0; i < nacc; ++i) {
    AccDev& d = plan.a[i];
    d.kind    = low.acc[i].kind;
    d.col     = low.acc[i].col < 0 ? -1 : (int32_t)(std::find(colmap.begin(), colmap.end(), low.acc[i].col) - colmap.begin());
    d.lds_off = -1;
    {
      const int ki = d.kind == ACC_VCOUNT ? 0 : d.kind == ACC_ISUM ? 1 : d.kind == ACC_FSUM ? 2 : d.kind == ACC_MIN ? 3
                   : d.kind == ACC_MAX  ? 4 : -1;
      if (ki >= 0) {
        LDF_EXPECTS(plan.cacc[d.col][ki] < 0, "groupby: two accumulators of one kind on one column");
        plan.cacc[d.col][ki] = (int8_t)i;
      }
    }
    if (hot_cap > 0) {
      switch (d.kind) {
        case ACC_VCOUNT:
        case ACC_DISTINCT: d.lds_off = take(hot_cap * 4); break;
        case ACC_ISUM:
        case ACC_MIN:
        case ACC_MAX: d.lds_off = take(hot_cap * 8); break;
        case ACC_FSUM: d.lds_off = take(hot_cap * 48); break;
        default: break;
      }
    }
    if (d.kind == ACC_MIN || d.kind == ACC_MAX) d.flag_bit = flag_bits++;
  }
  LDF_EXPECTS(flag_bits <= 32, "groupby: too many MIN/MAX aggregations");
  // the hot key table last: k_gb_main3 reads row classes and allocates only the accumulators
  plan.acc_end = off;
  if (hot_cap > 0) {
    plan.tab_off  = take(2 * hot_cap * 8);
    plan.tidx_off = take(2 * hot_cap * 2);
  }
  plan.nacc      = nacc;
  plan.ncol      = (int32_t)colmap.size();
  plan.hot_cap   = hot_cap;
  plan.lds_bytes = off;
  for (size_t c = 0; c < colmap.size(); ++c) plan.col[c] = to_coldev(tv->cols[colmap[c]]);

  // hot keys from the sample
  Buf<unsigned long long> hot_keys(a, hot_cap > 0 ? hot_cap : 1);
  Buf<int32_t> hot_special(a, hot_cap > 0 ? hot_cap : 1);
  Buf<int32_t> hot_gslot(a, hot_cap > 0 ? hot_cap : 1);
  Buf<unsigned int> hot_count(a, 1);
  fill_u32(hot_count.get(), 0, 1, s);
  HotList hot{hot_keys.get(), hot_special.get(), hot_gslot.get(), hot_count.get()};
  if (hot_cap > 0 && S > 0) {
    hipLaunchKernelGGL(k_gb_hot_select, dim3(grid_for(scap + 2, 256)), dim3(256), 0, s, sslots.get(), scnt.get(),
                       (int64_t)scap, thr, hot_cap, hot);
    LDF_HIP_TRY(hipGetLastError());
  }
  sslots.reset();
  scnt.reset();

  uint64_t cap    = next_pow2(std::max<uint64_t>(1024, (uint64_t)(2.5 * (double)est)));
  uint64_t maxcap = next_pow2(std::max<uint64_t>(1024, 2 * (uint64_t)n + 2));
  if (cap > maxcap) cap = maxcap;

  // Partitioned path for the keys outside the hot set (single fixed-width key, large inputs,
  // accumulators it handles; COUNT_DISTINCT needs 32-bit keys and values and no null keys).
  bool use_v3 = !generic && hot_cap > 0 && n >= (int64_t(1) << 20);
  {
    const char* env = std::getenv("LDF_GROUPBY_PARTITIONED");
    if (env != nullptr && env[0] == '0') use_v3 = false;
  }
  // its row passes load each value as the aligned 8-byte word holding it: values must be
  // naturally aligned (Arrow buffers are)
  {
    auto aligned = [](const void* d, int w) { return w <= 0 || (reinterpret_cast<uintptr_t>(d) % (uintptr_t)w) == 0; };
    if (!aligned(keys.c[0].data, keys.c[0].width)) use_v3 = false;
    for (size_t c = 0; c < colmap.size(); ++c)
      if (!aligned(tv->cols[colmap[c]].data, type_width(tv->cols[colmap[c]].type_id))) use_v3 = false;
  }
  int ndistinct = 0;
  for (auto& h : low.acc) {
    if (h.kind == ACC_PROD) use_v3 = false;
    if (h.kind == ACC_DISTINCT) {
      ++ndistinct;
      if (type_width(tv->cols[h.col].type_id) > 4 || type_width(tv->cols[h.col].type_id) == 0) use_v3 = false;
    }
  }
  if (ndistinct > 0 && (keys.c[0].width > 4 || keys.c[0].valid != nullptr || ndistinct > 3)) use_v3 = false;
  for (int i = 0; i < naggs; ++i)
    if (aggs[i].kind == LDF_AGG_VARIANCE || aggs[i].kind == LDF_AGG_STDDEV) use_v3 = false;
  // COUNT_DISTINCT through pair partitions deduplicated in LDS (one such accumulator on the
  // partitioned path); a dedupe table that ever overflows sends the call back to the pair-set path
  bool dpart = use_v3 && ndistinct == 1 && hot_cap <= kPairHotMax;
  {
    const char* env = std::getenv("LDF_GROUPBY_DPART");
    if (env != nullptr && env[0] == '0') dpart = false;
  }
  int dacc = -1;
  for (int i = 0; i < nacc; ++i)
    if (low.acc[i].kind == ACC_DISTINCT) dacc = i;
  const int32_t dist_lds_off = dacc >= 0 ? plan.a[dacc].lds_off : -1;

  // The partitioned path's plan: partition tables (k_gb_part3) sized from the accumulators' LDS
  // bytes per slot, the partition count, and the dense slots that hold the keys outside the hot
  // set (one region of table_slots per (partition, sub-pass) block, after the hash table).
  PartPlan q;
  std::memset(&q, 0, sizeof(q));
  int64_t dense = 0;
  if (use_v3 && n > 0) {
    int per_slot = 8 + 4 + 4;
    for (int i = 0; i < nacc; ++i) {
      switch (plan.a[i].kind) {
        case ACC_VCOUNT:
        case ACC_DISTINCT: per_slot += 4; break;
        case ACC_ISUM:
        case ACC_MIN:
        case ACC_MAX: per_slot += 8; break;
        case ACC_FSUM: per_slot += 48; break;
        default: break;
      }
    }
    q.tcap = 256;
    while (q.tcap * 2 * per_slot + 64 * (nacc + 3) <= 150 * 1024 && q.tcap < 8192) q.tcap *= 2;
    // Partitions: the fewest (2^10 .. 2^12) whose keys fit one table each -- the sample's
    // estimate runs low on skewed keys: plan for 1.25x as many at <= 45% load -- as far as the
    // main pass's LDS cursors allow; keys beyond that take sub-passes over a partition.
    // k_gb_main3's LDS: hot accumulators, the partition cursors, and (when it fits) one
    // partition record per thread for packing
    const int hc_bytes = dpart ? kHPairParts * 4 : 0;  // k_gb_main3's per-block pair histogram
    int rec            = keys.c[0].width + 1;
    for (int c = 0; c < plan.ncol; ++c) rec += plan.col[c].width;
    int stride = 8;
    while (stride < rec) stride *= 2;
    const int stage_b = kGbThreads * stride;
    plan.stage        = plan.acc_end + (1 << kMinPartBits) * 4 + hc_bytes + stage_b <= kMain3Lds ? 1 : 0;
    const int main_fixed = plan.acc_end + hc_bytes + (plan.stage ? stage_b : 0);
    int pbits            = kMinPartBits;
    while (pbits < kMaxPartBits && 1.25 * (double)est / (double)(1 << pbits) > 0.45 * q.tcap &&
           main_fixed + (1 << (pbits + 1)) * 4 <= kMain3Lds)
      ++pbits;
    plan.part_bits             = pbits;
    const int nparts           = 1 << pbits;
    const double keys_per_part = 1.25 * (double)est / nparts;
    q.S                        = std::max(1, (int)std::ceil(keys_per_part / (0.45 * q.tcap)));
    dense                      = (int64_t)nparts * q.S * q.tcap;
  }

  for (;;) {
    LDF_EXPECTS(cap + 2 + dense < (uint64_t)INT32_MAX, "groupby: too many groups");
    const int64_t nslot = (int64_t)cap + 2 + dense;
    Buf<unsigned long long> slots(a, nslot);
    // status: [0] nkeys, [1] overflow (int), [2.. 2+nacc) err flags (int), then null counts
    const int nstatus = 2 + nacc + nkeys + naggs + 1;
    Buf<unsigned long long> status(a, nstatus);
    {
      const FillJob jobs[2] = {{slots.get(), kGEmpty, nslot, 8}, {status.get(), 0, nstatus, 8}};
      fill_many(jobs, 2, s);
    }
    KeyTab t{slots.get(), cap - 1, cap, status.get(), reinterpret_cast<int32_t*>(status.get() + 1)};

    // global accumulator arrays: one buffer, filled with one launch
    std::vector<uint64_t> off0(nacc), off1(nacc, 0), off2(nacc, 0), offp(nacc, 0), offf(nacc, 0);
    FillPlan fp;
    fp.n            = 0;
    uint64_t total  = 0;
    // an array of `len` words whose first `fill` words start at `val` (the dense slots of the
    // partitioned path are written whole by k_gb_part3 and need no initial value)
    auto seg = [&](uint64_t len, uint64_t val, uint64_t fill) {
      uint64_t o = total;
      total += len;
      if (fp.n > 0 && fp.s[fp.n - 1].val == val && fp.s[fp.n - 1].off + fp.s[fp.n - 1].len == o && fill == len)
        fp.s[fp.n - 1].len += fill;
      else if (fill > 0)
        fp.s[fp.n++] = FillSeg{o, fill, val};
      return o;
    };
    const uint64_t hashed = cap + 2;  // hash region + the two special slots
    uint64_t pcap_total = 0;
    std::vector<uint64_t> pcap(nacc, 0);
    for (int i = 0; i < nacc; ++i) {
      const HostAcc& h = low.acc[i];
      const bool fl    = h.col >= 0 && is_float(tv->cols[h.col].type_id);
      uint64_t init    = 0;
      if (h.kind == ACC_MIN) init = ~0ull;
      if (h.kind == ACC_PROD) init = fl ? 0x3ff0000000000000ull : 1ull;
      off0[i] = seg(nslot, init, hashed);
      if (h.kind == ACC_FSUM) {
        off1[i] = seg(nslot, 0, hashed);
        off2[i] = seg(nslot, 0, hashed);
      }
      if (h.kind == ACC_MIN || h.kind == ACC_MAX) offf[i] = seg((nslot + 1) / 2, 0, (nslot + 1) / 2);  // u32 flags
      if (h.kind == ACC_DISTINCT && !use_v3) {
        pcap[i] = next_pow2(std::max<uint64_t>(1024, 2 * (uint64_t)n));
        offp[i] = seg(pcap[i], kGEmpty, pcap[i]);
        pcap_total += pcap[i];
      }
      LDF_EXPECTS(fp.n < kMaxFillSeg, "groupby: too many accumulator segments");
    }
    Buf<unsigned long long> state(a, total > 0 ? total : 1);
    if (total > 0) {
      unsigned gx = (unsigned)std::min<uint64_t>(2048, (total / std::max(1, fp.n) + 255) / 256 + 1);
      hipLaunchKernelGGL(k_fill_segments, dim3(gx, fp.n), dim3(256), 0, s, state.get(), fp);
      LDF_HIP_TRY(hipGetLastError());
    }
    unsigned long long* st = state.get();
    for (int i = 0; i < nacc; ++i) {
      AccDev& d = plan.a[i];
      d.g0      = st + off0[i];
      d.g1      = low.acc[i].kind == ACC_FSUM ? st + off1[i] : nullptr;
      d.g2      = low.acc[i].kind == ACC_FSUM ? st + off2[i] : nullptr;
      d.gflag   = (d.kind == ACC_MIN || d.kind == ACC_MAX) ? reinterpret_cast<uint32_t*>(st + offf[i]) : nullptr;
      d.pset    = (d.kind == ACC_DISTINCT && !use_v3) ? st + offp[i] : nullptr;
      d.pmask   = (d.kind == ACC_DISTINCT && !use_v3) ? pcap[i] - 1 : 0;
      d.pset2   = nullptr;
      d.pmask2  = 0;
      d.err     = reinterpret_cast<int32_t*>(status.get() + 2 + i);
    }

    // ---- the pass over the rows ----
    if (hot_cap > 0) {
      hipLaunchKernelGGL(k_gb_hot_insert, dim3(grid_for(hot_cap, 256)), dim3(256), 0, s, t, hot, hot_cap);
      LDF_HIP_TRY(hipGetLastError());
    }
    // buffers of the partitioned path (freed at the end of the attempt)
    Buf<unsigned long long> psets;
    Buf<uint32_t> pcounts;
    Buf<uint64_t> poffs;
    Buf<unsigned char> precs;
    Buf<unsigned long long> hparts;
    Buf<uint32_t> hcounts;
    Buf<uint16_t> codes;
    Buf<uint16_t> dcodes;  // the key's dense class table (k_gb_hist3)
    Buf<uint64_t> hoffs, hpoffs;  // hot pairs: per histogram block; per (pair partition, main block)
    Buf<uint32_t> hpcnt;          // hot pairs per (pair partition, main block)
    Buf<uint32_t> hpairs;  // hot pairs' stored words by pair partition
    Buf<unsigned long long> hfine, cpairs, cfine;
    Buf<uint32_t> part_cnt;  // hot pairs' new-pair counts per (pair partition, hot key)
    Buf<uint32_t> cpair_count;
    plan.dpart = 0;
    plan.dcol  = 0;
    if (dacc >= 0) plan.a[dacc].lds_off = (use_v3 && dpart) ? -1 : dist_lds_off;
    if (n > 0 && use_v3) {
      if (dpart) {
        plan.dpart = 1;
        plan.dcol  = plan.a[dacc].col;
      }
      const int hc_bytes         = dpart ? kHPairParts * 4 : 0;  // the per-block pair histogram
      const int nparts           = 1 << plan.part_bits;
      int rec_bytes              = keys.c[0].width + 1;
      for (int c = 0; c < plan.ncol; ++c) rec_bytes += plan.col[c].width;
      int rec_stride = 8;
      while (rec_stride < rec_bytes) rec_stride *= 2;
      const int main_lds = plan.acc_end + nparts * 4 + hc_bytes + (plan.stage ? kGbThreads * rec_stride : 0);
      int per_cu            = std::max(1, std::min(2, (160 * 1024) / main_lds));
      int nblk              = 256 * per_cu;
      int64_t rpb           = (n + nblk - 1) / nblk;
      rpb                   = (rpb + kSub3 - 1) / kSub3 * kSub3;
      nblk                  = (int)((n + rpb - 1) / rpb);
      LDF_EXPECTS(rpb < INT32_MAX, "groupby: more than 2^31 rows per block of the main pass");
      plan.rows_per_block   = rpb;
      // 1. histogram of the non-hot rows per (partition, histogram block)
      const int nh = nblk * kHistSplit;
      pcounts      = Buf<uint32_t>(a, (size_t)nparts * nh);
      poffs        = Buf<uint64_t>(a, (size_t)nparts * nh + 1);
      {
        static bool attr = false;
        if (!attr) {
          LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_hist3<true>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
          LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_hist3<false>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
          LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_part3<1>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
          LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_part3<2>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
          LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_part3<4>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
          LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_gb_part3<8>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
          attr = true;
        }
        if (dpart) {
          hcounts = Buf<uint32_t>(a, (size_t)nh);  // hot pairs per histogram block
          hoffs   = Buf<uint64_t>(a, (size_t)nh + 1);
        }
        codes = Buf<uint16_t>(a, (size_t)n);
        ProfScope prof("gb_hist", s);
        GbPlan hplan    = plan;  // the hot key table at the start of LDS, the histograms after it
        const bool pack = keys.c[0].width <= 4;
        hplan.tab_off   = 0;
        hplan.tidx_off  = 2 * plan.hot_cap * 8;
        hplan.lds_bytes = pack ? hot32_slots(plan.hot_cap) * 8 : (hplan.tidx_off + 2 * plan.hot_cap * 2 + 15) & ~15;
        // a key of at most 4 bytes whose sampled values span at most n / 4 values (and 2^22) is
        // classified through a table of codes over that range (LDF_GROUPBY_DENSE_CODES=0 turns the
        // table off, a tuning knob)
        static const bool dense_on = [] {
          const char* e = std::getenv("LDF_GROUPBY_DENSE_CODES");
          return e == nullptr || e[0] != '0';
        }();
        DenseCodes dense{nullptr, 0, 0};
        dcodes = Buf<uint16_t>(a, 2);  // without a table its one entry is read and never used
        if (pack && dense_on && smin <= smax && smax - smin < (uint64_t)kDenseMaxRange &&
            (int64_t)(smax - smin + 1) <= n / 4) {
          dense.kmin  = (uint32_t)smin;
          dense.range = (uint32_t)(smax - smin + 1);
          dcodes      = Buf<uint16_t>(a, dense.range);
          hipLaunchKernelGGL(k_gb_dense_fill, dim3(grid_for(dense.range, 256)), dim3(256), 0, s, dcodes.get(), dense.kmin,
                             dense.range, plan.part_bits);
          hipLaunchKernelGGL(k_gb_dense_hot, dim3(grid_for(std::max(plan.hot_cap, 1), 256)), dim3(256), 0, s, dcodes.get(), dense.kmin,
                             dense.range, hot, plan.hot_cap);
          LDF_HIP_TRY(hipGetLastError());
        }
        dense.code = dcodes.get();
        if (pack)
          hipLaunchKernelGGL(k_gb_hist3<true>, dim3(nh), dim3(kHistThreads), hplan.lds_bytes + nparts * 4, s, keys, n,
                             hplan, hot, pcounts.get(), hcounts.get(), codes.get(), dense);
        else
          hipLaunchKernelGGL(k_gb_hist3<false>, dim3(nh), dim3(kHistThreads), hplan.lds_bytes + nparts * 4, s, keys, n,
                             hplan, hot, pcounts.get(), hcounts.get(), codes.get(), dense);
        LDF_HIP_TRY(hipGetLastError());
      }
      exclusive_scan_u32_to_u64(pcounts.get(), poffs.get(), (int64_t)nparts * nh, a);
      if (dpart) exclusive_scan_u32_to_u64(hcounts.get(), hoffs.get(), (int64_t)nh, a);
      uint64_t totals[2] = {0, 0};
      {  // both totals queued into the pinned stage, one synchronisation
        uint64_t* st = static_cast<uint64_t*>(pinned_stage(16));
        st[1]        = 0;
        LDF_HIP_TRY(hipMemcpyAsync(st, poffs.get() + (size_t)nparts * nh, 8, hipMemcpyDeviceToHost, s));
        if (dpart) LDF_HIP_TRY(hipMemcpyAsync(st + 1, hoffs.get() + (size_t)nh, 8, hipMemcpyDeviceToHost, s));
        LDF_HIP_TRY(hipStreamSynchronize(s));
        totals[0] = st[0];
        totals[1] = st[1];
      }
      static const bool trace = [] {
        const char* e = std::getenv("LDF_GROUPBY_TRACE");
        return e != nullptr && e[0] == '1';
      }();
      const int64_t ncold = (int64_t)totals[0];
      const int64_t nhp   = (int64_t)totals[1];  // hot rows' distinct pairs (dpart)
      {  // LDF_GROUPBY_TRACE=1: one line per call with the partitioned path's plan
        if (trace)
          std::fprintf(stderr,
                       "[ldf groupby] rows=%lld est=%lld hot_cap=%d blocks=%d part_bits=%d sub_passes=%d "
                       "table_slots=%d cold_rows=%lld hot_pairs=%lld dpart=%d\n",
                       (long long)n, (long long)est, plan.hot_cap, nblk, plan.part_bits, q.S, q.tcap,
                       (long long)ncold, (long long)nhp, (int)dpart);
      }
      // 2. sets of (key, value) pairs for COUNT_DISTINCT: hot keys by global slot, others by key
      uint64_t hcap = next_pow2(std::max<uint64_t>(1024, (uint64_t)(1.5 * (double)(n - ncold))));
      uint64_t ccap = next_pow2(std::max<uint64_t>(1024, (uint64_t)(1.5 * (double)ncold)));
      if (ndistinct > 0 && !dpart) {
        psets = Buf<unsigned long long>(a, (size_t)ndistinct * (hcap + ccap));
        fill_u64((uint64_t*)psets.get(), kGEmpty, (int64_t)ndistinct * (hcap + ccap), s);
        int di = 0;
        for (int i = 0; i < nacc; ++i) {
          if (plan.a[i].kind != ACC_DISTINCT) continue;
          plan.a[i].pset   = psets.get() + (size_t)di * (hcap + ccap);
          plan.a[i].pmask  = hcap - 1;
          plan.a[i].pset2  = plan.a[i].pset + hcap;
          plan.a[i].pmask2 = ccap - 1;
          ++di;
        }
      }
      // 3. partition record buffers (SoA)
      PartBufs pb;
      std::memset(&pb, 0, sizeof(pb));
      pb.key_width = keys.c[0].width;
      pb.nblk      = nh;
      pb.offsets   = poffs.get();
      {  // record layout: fields widest first (each at a multiple of its width), flags byte last
        int32_t off = 0;
        for (int wdt = 8; wdt >= 1; wdt >>= 1) {
          if (pb.key_width == wdt) {
            pb.key_off = off;
            off += wdt;
          }
          for (int c = 0; c < plan.ncol; ++c)
            if (plan.col[c].width == wdt) {
              pb.val_off[c] = off;
              off += wdt;
            }
        }
        pb.flags_off = off++;
        pb.stride    = 8;
        while (pb.stride < off) pb.stride *= 2;
        LDF_EXPECTS(pb.stride <= 64, "groupby: partition record over 64 bytes");
      }
      precs  = Buf<unsigned char>(a, (size_t)std::max<int64_t>(1, ncold) * (size_t)pb.stride);
      pb.rec = precs.get();
      if (dpart) {
        hpairs      = Buf<uint32_t>(a, (size_t)std::max<int64_t>(1, nhp));
        hfine       = Buf<unsigned long long>(a, (size_t)std::max<int64_t>(1, std::max(nhp, ncold)));
        // row order (words, then partitions: 6 bytes a pair); split by pair partition into hpairs
        // after the pass, after which the hot dedupe reuses the buffer
        pb.hwords   = reinterpret_cast<uint32_t*>(hfine.get());
        pb.hparts   = reinterpret_cast<uint16_t*>(pb.hwords + std::max<int64_t>(1, nhp));
        pb.hoffsets = hoffs.get();
        hpcnt       = Buf<uint32_t>(a, (size_t)kHPairParts * nblk);
        pb.pcounts  = hpcnt.get();
        if (q.S == 1 && ncold > 0) {
          cpairs         = Buf<unsigned long long>(a, (size_t)ncold);
          cpair_count    = Buf<uint32_t>(a, (size_t)nparts);
          pb.cpairs      = cpairs.get();
          pb.cpair_count = cpair_count.get();
        }
      }
      // 4. per-block partials of the hot slots
      HotParts hp;
      std::memset(&hp, 0, sizeof(hp));
      size_t words = 2 * (size_t)nblk * hot_cap;  // rows + flags as u32 pairs
      std::vector<size_t> aoff(nacc, 0);
      for (int i = 0; i < nacc; ++i) {
        if (plan.a[i].lds_off < 0) continue;
        aoff[i] = words;
        words += (size_t)nblk * hot_cap * (plan.a[i].kind == ACC_FSUM ? 3 : 1);
      }
      hparts   = Buf<unsigned long long>(a, words);
      hp.rows  = reinterpret_cast<uint32_t*>(hparts.get());
      hp.flags = hp.rows + (size_t)nblk * hot_cap;
      for (int i = 0; i < nacc; ++i)
        if (plan.a[i].lds_off >= 0) hp.acc[i] = hparts.get() + aoff[i];
      // 5. the pass over the rows
      {
        ProfScope prof("gb_aggregate", s);
        launch_main3(dpart, plan.ncol, dim3(nblk), main_lds, s, keys, n, plan, hot, pb, hp, (const uint16_t*)codes.get());
      }
      // The hot keys' partials (k_gb_hot_reduce) and the hot pairs' COUNT_DISTINCT chain (split by
      // pair partition, deduplicated, summed per hot key) read only the main pass's hot output and
      // write only the hot keys' slots: they run on a second stream beside the partition pass and
      // the cold pairs' deduplication, which each leave much of the GPU idle (one or two blocks per
      // CU, waits on memory).  Every buffer either chain uses lives until the end of this attempt,
      // after the caller's stream has waited for both.
      SideStream& side = side_stream();
      bool forked      = false;
      if (!(dpart && nhp > 0)) {
        hipLaunchKernelGGL(k_gb_hot_reduce, dim3(hot_cap), dim3(256), 0, s, plan, hot, hp, nblk);
        LDF_HIP_TRY(hipGetLastError());
      }
      if (dpart && nhp > 0) {
        hpoffs   = Buf<uint64_t>(a, (size_t)kHPairParts * nblk + 1);
        part_cnt = Buf<uint32_t>(a, (size_t)kHPairParts * hot_cap);
        LDF_HIP_TRY(hipEventRecord(side.fork, s));
        LDF_HIP_TRY(hipStreamWaitEvent(side.s, side.fork, 0));
        forked = true;
        hipLaunchKernelGGL(k_gb_hot_reduce, dim3(hot_cap), dim3(256), 0, side.s, plan, hot, hp, nblk);
        LDF_HIP_TRY(hipGetLastError());
        {  // the pair partitions' cursors, scanned on the second stream (its scratch there too)
          Alloc aside(a.raw(), side.s);
          exclusive_scan_u32_to_u64(hpcnt.get(), hpoffs.get(), (int64_t)kHPairParts * nblk, aside);
        }
        ProfScope prof("gb_hot_distinct", side.s);
        const AccDev& d = plan.a[dacc];
        hipLaunchKernelGGL(k_gb_pair_part, dim3(nblk), dim3(kHPairParts), 0, side.s, (const uint32_t*)pb.hwords,
                           (const uint16_t*)pb.hparts, (const uint64_t*)hoffs.get(), kHistSplit,
                           (const uint64_t*)hpoffs.get(), hpairs.get());
        LDF_HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_gb_hot_distinct, dim3(kHPairParts), dim3(1024), 0, side.s, (const uint32_t*)hpairs.get(),
                           (const uint64_t*)hpoffs.get(), nblk, reinterpret_cast<uint32_t*>(hfine.get()), hot_cap, hot,
                           part_cnt.get(), d.err);
        LDF_HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_gb_hot_dist_sum, dim3(hot_cap), dim3(256), 0, side.s, (const uint32_t*)part_cnt.get(),
                           kHPairParts, hot_cap, hot, d.g0);
        LDF_HIP_TRY(hipGetLastError());
        LDF_HIP_TRY(hipEventRecord(side.join, side.s));
      }
      // 6. one block per (partition, sub-pass)
      int32_t loff = 0;
      auto ltake   = [&](int bytes) {
        int32_t o = loff;
        loff += (bytes + 15) & ~15;
        return o;
      };
      q.key_off   = ltake(q.tcap * 8);
      q.rows_off  = ltake(q.tcap * 4);
      q.flags_off = ltake(q.tcap * 4);
      for (int i = 0; i < nacc; ++i) {
        q.acc_off[i] = -1;
        switch (plan.a[i].kind) {
          case ACC_VCOUNT:
          case ACC_DISTINCT: q.acc_off[i] = ltake(q.tcap * 4); break;
          case ACC_ISUM:
          case ACC_MIN:
          case ACC_MAX: q.acc_off[i] = ltake(q.tcap * 8); break;
          case ACC_FSUM: q.acc_off[i] = ltake(q.tcap * 48); break;
          default: break;
        }
      }
      q.lds_bytes = loff;
      if (ncold > 0) {
        ProfScope prof("gb_partitions", s);
        const int grid = nparts * q.S;
        switch (pb.stride) {
          case 8: hipLaunchKernelGGL(k_gb_part3<1>, dim3(grid), dim3(kGbThreads), q.lds_bytes, s, plan, q, pb, t); break;
          case 16: hipLaunchKernelGGL(k_gb_part3<2>, dim3(grid), dim3(kGbThreads), q.lds_bytes, s, plan, q, pb, t); break;
          case 32: hipLaunchKernelGGL(k_gb_part3<4>, dim3(grid), dim3(kGbThreads), q.lds_bytes, s, plan, q, pb, t); break;
          default: hipLaunchKernelGGL(k_gb_part3<8>, dim3(grid), dim3(kGbThreads), q.lds_bytes, s, plan, q, pb, t); break;
        }
        LDF_HIP_TRY(hipGetLastError());
      }
      // 7. COUNT_DISTINCT of the keys outside the hot set: their pairs by (partition, sub-pass), after
      //    the partition pass has given every key its slot (own scratch: the hot chain uses hfine)
      if (dpart && ncold > 0) {
        ProfScope prof("gb_distinct", s);
        const AccDev& d = plan.a[dacc];
        cfine           = Buf<unsigned long long>(a, (size_t)ncold);
        if (pb.cpairs != nullptr) {
          hipLaunchKernelGGL(k_gb_cold_distinct_slots, dim3(nparts), dim3(1024), (size_t)q.tcap * 4, s, pb, t, q.tcap,
                             cfine.get(), d.g0, d.err);
          LDF_HIP_TRY(hipGetLastError());
        } else {
          hipLaunchKernelGGL(k_gb_cold_distinct, dim3(nparts), dim3(1024), 0, s, pb, plan.dcol,
                             plan.col[plan.dcol].width, t, q.S, q.tcap, cfine.get(), d.g0, d.err);
          LDF_HIP_TRY(hipGetLastError());
        }
      }
      if (forked) LDF_HIP_TRY(hipStreamWaitEvent(s, side.join, 0));
    } else if (n > 0) {
      // one wave of resident blocks (256 CUs; two 1024-thread blocks per CU when LDS allows)
      int per_cu = plan.lds_bytes > 0 ? std::max(1, std::min(2, (160 * 1024) / plan.lds_bytes)) : 2;
      int nblk   = 256 * per_cu;
      int64_t rpb = (n + nblk - 1) / nblk;
      rpb         = (rpb + 1023) / 1024 * 1024;
      nblk        = (int)((n + rpb - 1) / rpb);
      LDF_EXPECTS(rpb < INT32_MAX, "groupby: more than 2^31 rows per block of the main pass");
      plan.rows_per_block = rpb;
      if (generic) launch_main<true>(keys, n, t, plan, hot, nblk, s);
      else launch_main<false>(keys, n, t, plan, hot, nblk, s);
    }

    // compact occupied slots into groups
    Buf<uint32_t> occ(a, nslot);
    Buf<uint64_t> pos(a, nslot + 1);
    hipLaunchKernelGGL(k_gb_occupied, dim3(grid_for(nslot, 256)), dim3(256), 0, s, slots.get(), nslot, occ.get());
    LDF_HIP_TRY(hipGetLastError());
    exclusive_scan_u32_to_u64(occ.get(), pos.get(), nslot, a);
    LDF_HIP_TRY(hipMemcpyAsync(status.get() + nstatus - 1, pos.get() + nslot, 8, hipMemcpyDeviceToDevice, s));
    std::vector<unsigned long long> hs(nstatus);
    copy_to_host(hs.data(), status.get(), nstatus * 8, s);
    if (use_v3 && dpart && dacc >= 0 && (int32_t)hs[2 + dacc]) {  // a dedupe table overflowed
      dpart = false;
      continue;
    }
    if ((int32_t)hs[1]) {
      LDF_EXPECTS(cap < maxcap, "groupby: key table overflow");
      cap = std::min(maxcap, cap * 4);
      continue;
    }
    const int64_t ng = (int64_t)hs[nstatus - 1];

    // FP64 fallback for sums that left the fixed-point range; VARIANCE/STDDEV second pass
    std::vector<Buf<double>> fsum(nacc);
    std::vector<int> use_fsum(nacc, 0);
    for (int i = 0; i < nacc; ++i) {
      if (low.acc[i].kind != ACC_FSUM || !(int32_t)hs[2 + i]) continue;
      fsum[i] = Buf<double>(a, nslot);
      fill_u64(reinterpret_cast<uint64_t*>(fsum[i].get()), 0, nslot, s);
      ColDev c = to_coldev(tv->cols[low.acc[i].col]);
      if (generic) launch_second<true>(keys, n, t, c, fsum[i].get(), nullptr, s);
      else launch_second<false>(keys, n, t, c, fsum[i].get(), nullptr, s);
      use_fsum[i] = 1;
    }
    std::vector<Buf<double>> m2(naggs);
    for (int i = 0; i < naggs; ++i) {
      if (aggs[i].kind != LDF_AGG_VARIANCE && aggs[i].kind != LDF_AGG_STDDEV) continue;
      const int fi = low.ref[i].main, vi = low.ref[i].vcount;
      Buf<double> mean(a, nslot);
      hipLaunchKernelGGL(k_gb_means, dim3(grid_for(nslot, 256)), dim3(256), 0, s, plan.a[fi].g0, plan.a[fi].g1,
                         plan.a[fi].g2, plan.a[vi].g0, fsum[fi].get(), use_fsum[fi], nslot, mean.get());
      LDF_HIP_TRY(hipGetLastError());
      m2[i] = Buf<double>(a, nslot);
      fill_u64(reinterpret_cast<uint64_t*>(m2[i].get()), 0, nslot, s);
      ColDev c = to_coldev(tv->cols[aggs[i].in_col]);
      if (generic) launch_second<true>(keys, n, t, c, m2[i].get(), mean.get(), s);
      else launch_second<false>(keys, n, t, c, m2[i].get(), mean.get(), s);
    }

    Buf<int32_t> sog(a, ng > 0 ? ng : 1);
    Buf<int32_t> rep;
    if (generic) rep = Buf<int32_t>(a, ng > 0 ? ng : 1);
    hipLaunchKernelGGL(k_gb_compact, dim3(grid_for(nslot, 256)), dim3(256), 0, s, slots.get(), nslot, pos.get(), generic,
                       sog.get(), rep.get());
    LDF_HIP_TRY(hipGetLastError());

    out->ncols = nkeys + naggs;
    int done   = 0;
    try {
      OutPlan op;
      std::memset(&op, 0, sizeof(op));
      op.cap                      = cap;
      op.slots                    = slots.get();
      unsigned long long* nullcnt = status.get() + 2 + nacc;  // one per output column
      auto flush_plan             = [&]() {
        if (op.n > 0 && ng > 0) {
          hipLaunchKernelGGL(k_gb_finalize, dim3(grid_for(ng, 256, 256 * 16)), dim3(256), 0, s, op, sog.get(), ng);
          LDF_HIP_TRY(hipGetLastError());
        }
        op.n = 0;
      };
      auto add_out = [&](const OutDev& o) {
        if (op.n == kMaxOut) flush_plan();
        op.o[op.n++] = o;
      };
      if (generic) {
        std::vector<const ldf_column_view*> src(nkeys);
        for (int i = 0; i < nkeys; ++i) src[i] = &tv->cols[key_idx[i]];
        gather_columns(src.data(), nkeys, rep.get(), ng, false, a, out->cols);
        done = nkeys;
      } else {
        const ldf_column_view& kc = tv->cols[key_idx[0]];
        const bool nullable       = kc.validity != nullptr;
        alloc_column(kc.type_id, ng, nullable, 0, a, &out->cols[0]);
        done = 1;
        OutDev o;
        std::memset(&o, 0, sizeof(o));
        o.kind     = -1;
        o.out_type = kc.type_id;
        o.out      = out->cols[0].data;
        o.valid    = nullable ? reinterpret_cast<uint64_t*>(out->cols[0].validity) : nullptr;
        o.nulls    = nullcnt + 0;
        add_out(o);
      }
      for (int i = 0; i < naggs; ++i, ++done) {
        const ldf_agg& g   = aggs[i];
        const int32_t in_t = g.kind == LDF_AGG_COUNT_ALL ? LDF_TYPE_INT64 : tv->cols[g.in_col].type_id;
        const int32_t ot   = agg_out_type(g.kind, in_t);
        const bool nullable = !(g.kind == LDF_AGG_COUNT_ALL || g.kind == LDF_AGG_COUNT_DISTINCT);
        alloc_column(ot, ng, nullable, 0, a, &out->cols[done]);
        const auto& R = low.ref[i];
        OutDev o;
        std::memset(&o, 0, sizeof(o));
        o.kind     = g.kind;
        o.mode     = g.count_mode;
        o.in_type  = in_t;
        o.out_type = ot;
        if (R.main >= 0) {
          o.main = plan.a[R.main].g0;
          o.l1   = plan.a[R.main].g1;
          o.l2   = plan.a[R.main].g2;
          o.flag = plan.a[R.main].gflag;
          if (low.acc[R.main].kind == ACC_FSUM) {
            o.use_fsum = use_fsum[R.main];
            o.fsum     = fsum[R.main].get();
          }
        }
        if (g.kind == LDF_AGG_VARIANCE || g.kind == LDF_AGG_STDDEV) o.fsum = m2[i].get();
        o.vcount = R.vcount >= 0 ? plan.a[R.vcount].g0 : nullptr;
        o.rows   = R.rows >= 0 ? plan.a[R.rows].g0 : nullptr;
        o.out    = out->cols[done].data;
        o.valid  = nullable ? reinterpret_cast<uint64_t*>(out->cols[done].validity) : nullptr;
        o.nulls  = nullcnt + done;
        add_out(o);
      }
      flush_plan();
      std::vector<unsigned long long> nc(nkeys + naggs);
      copy_to_host(nc.data(), nullcnt, nc.size() * 8, s);
      for (int i = generic ? nkeys : 0; i < nkeys + naggs; ++i)
        out->cols[i].null_count = out->cols[i].validity ? (int64_t)nc[i] : 0;
    } catch (...) {
      for (int i = 0; i < done; ++i) free_column(&out->cols[i], a);
      out->ncols = 0;
      throw;
    }
    return;
  }
}

}  // namespace ldf

using namespace ldf;

namespace ldf {
void run_groupby_with_medians(const ldf_table_view* tv, const int32_t* key_idx, int32_t nkeys, const ldf_agg* aggs,
                              int32_t naggs, Alloc& a, ldf_table_out* out);  // median.hip
}

extern "C" int ldf_groupby(const ldf_table_view* table, const int32_t* keys, int32_t nkeys, const ldf_agg* aggs,
                           int32_t naggs, const ldf_allocator* alloc, void* stream, ldf_table_out* out)
{
  return guarded([&] {
    Alloc a(alloc, (hipStream_t)stream);
    bool median = false;
    for (int i = 0; i < naggs && aggs != nullptr; ++i) median |= aggs[i].kind == LDF_AGG_APPROX_MEDIAN;
    if (median) run_groupby_with_medians(table, keys, nkeys, aggs, naggs, a, out);
    else run_groupby(table, keys, nkeys, aggs, naggs, a, out);
  });
}
