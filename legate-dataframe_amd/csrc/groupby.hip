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
// Structure (one pass over the rows, everything keyed by hash-table slot):
//   key table : open addressing, 8-byte slots installed by 64-bit atomicCAS.  A single
//               fixed-width key stores its bits (two extra slots: the key equal to the empty
//               marker, and the null key); composite / string keys store
//               (fingerprint32 << 32 | representative row).
//   state     : per aggregation, arrays indexed by slot.  Floating SUM / MEAN / VARIANCE
//               accumulate x * 2^96 exactly in a 192-bit two's-complement integer spread over
//               three 64-bit words; the carry out of each word is read from the returned old
//               value of the atomic add, so the sum is exact and independent of the order of
//               the atomics and the final double is the correctly rounded exact sum.  Values
//               with |x| >= 2^62 or non-finite values divert that aggregation to a FP64 atomic
//               fallback pass.  Bits below 2^-96 are truncated (|error| <= n * 2^-96).
//   distinct  : COUNT_DISTINCT inserts (slot, value) into a second open-addressing set; a
//               successful CAS counts a new distinct value for the slot.
// The table size comes from a GEE distinct-value estimate over a 65536-row sample; if the
// table still fills beyond 3/4 the call is re-run with four times the capacity.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ldf_common.hpp"
#include "ldf_device.hpp"

namespace ldf {

constexpr unsigned long long kGEmpty = ~0ull;
constexpr int kMaxAggsPerLaunch       = 12;

struct AggDesc {
  int32_t kind;
  int32_t in_type;
  int32_t width;  // 0 = string
  int32_t mode;
  const void* data;
  const uint8_t* valid;
  const int64_t* offsets;
  unsigned long long* a0;   // SUM/PRODUCT/MIN/MAX accumulator or limb 0
  unsigned long long* a1;   // limb 1
  unsigned long long* a2;   // limb 2
  unsigned long long* cnt;  // valid-value / row / distinct count
  uint8_t* flag;            // NaN seen (MIN/MAX) or null seen (COUNT_DISTINCT)
  unsigned long long* pset; // COUNT_DISTINCT pair set
  uint64_t pmask;
  double* fsum;             // fallback FP64 sum / second-pass accumulator
  const double* mean;       // VARIANCE second pass
  int32_t* err;             // set when the exact path cannot represent a value
};

struct AggBatch {
  AggDesc d[kMaxAggsPerLaunch];
  int32_t n;
};

struct KeyTab {
  unsigned long long* slots;
  uint64_t mask;
  uint64_t cap;
  unsigned long long* nkeys;  // distinct keys inserted
  int32_t* overflow;
};

// ---- key table --------------------------------------------------------------------------
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
    for (uint64_t probes = 0; probes <= t.mask; ++probes) {
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
    for (uint64_t probes = 0; probes <= t.mask; ++probes) {
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

// ---- exact fixed-point FP64 accumulation ------------------------------------------------
// x -> round-toward-zero(x * 2^96) as a 192-bit two's-complement integer (l0 low .. l2 high).
// Returns false for values the fixed-point range cannot hold (non-finite or |x| >= 2^62).
__device__ __forceinline__ bool to_fixed192(double x, uint64_t& l0, uint64_t& l1, uint64_t& l2)
{
  uint64_t b   = (uint64_t)__double_as_longlong(x);
  int ex       = (int)((b >> 52) & 0x7ff);
  l0 = l1 = l2 = 0;
  if (ex == 0x7ff || ex >= 1023 + 62) return false;
  if (ex == 0) return true;  // zero or subnormal: below 2^-96, contributes 0
  uint64_t m = (b & ((1ull << 52) - 1)) | (1ull << 52);
  int shift  = ex - 1075 + 96;  // position of m's lowest bit in the fixed-point integer
  if (shift >= 0) {
    if (shift < 64) {
      l0 = m << shift;
      l1 = shift ? (m >> (64 - shift)) : 0;
    } else {  // shift < 106
      int s2 = shift - 64;
      l1     = m << s2;
      l2     = s2 ? (m >> (64 - s2)) : 0;
    }
  } else {
    l0 = (-shift < 64) ? (m >> (-shift)) : 0;
  }
  if (b >> 63) {  // negate
    l0 = ~l0 + 1;
    uint64_t c = (l0 == 0);
    l1         = ~l1 + c;
    c          = c && (l1 == 0);
    l2         = ~l2 + c;
  }
  return true;
}

__device__ __forceinline__ void add_fixed192(unsigned long long* a0, unsigned long long* a1, unsigned long long* a2,
                                             uint64_t l0, uint64_t l1, uint64_t l2)
{
  unsigned long long o0 = atomicAdd(a0, (unsigned long long)l0);
  uint64_t c0           = (o0 + l0) < o0;
  uint64_t y1           = l1 + c0;
  uint64_t c1a          = (y1 < l1);
  unsigned long long o1 = atomicAdd(a1, (unsigned long long)y1);
  uint64_t c1           = c1a | (uint64_t)((o1 + y1) < o1);
  uint64_t y2           = l2 + c1;
  if (y2 != 0) atomicAdd(a2, (unsigned long long)y2);
}

// Correctly rounded double of a 192-bit two's-complement integer times 2^-96.
__device__ inline double fixed192_to_double(uint64_t l0, uint64_t l1, uint64_t l2)
{
  bool neg = (l2 >> 63) != 0;
  if (neg) {
    l0 = ~l0 + 1;
    uint64_t c = (l0 == 0);
    l1         = ~l1 + c;
    c          = c && (l1 == 0);
    l2         = ~l2 + c;
  }
  int lz;
  if (l2) lz = __clzll(l2);
  else if (l1) lz = 64 + __clzll(l1);
  else if (l0) lz = 128 + __clzll(l0);
  else return 0.0;
  // normalise so the top set bit is bit 191
  uint64_t t2, t1, t0;
  if (lz >= 128) {
    t2 = l0 << (lz - 128);
    t1 = 0;
    t0 = 0;
  } else if (lz >= 64) {
    int s = lz - 64;
    t2    = s ? (l1 << s) | (l0 >> (64 - s)) : l1;
    t1    = s ? (l0 << s) : l0;
    t0    = 0;
  } else {
    int s = lz;
    t2    = s ? (l2 << s) | (l1 >> (64 - s)) : l2;
    t1    = s ? (l1 << s) | (l0 >> (64 - s)) : l1;
    t0    = s ? (l0 << s) : l0;
  }
  int p         = 191 - lz;  // exponent of the top bit (value = t * 2^(p-191) * 2^-96)
  uint64_t q    = t2 >> 11;
  uint64_t rem  = t2 & 0x7ff;
  bool sticky   = (t1 | t0) != 0;
  if (rem > 0x400 || (rem == 0x400 && (sticky || (q & 1)))) {
    ++q;
    if (q == (1ull << 53)) {
      q >>= 1;
      ++p;
    }
  }
  double v = ldexp((double)q, p - 52 - 96);
  return neg ? -v : v;
}

// ---- value helpers ----------------------------------------------------------------------
__device__ __forceinline__ double load_as_double(const AggDesc& a, int64_t r)
{
  switch (a.in_type) {
    case LDF_TYPE_DOUBLE: return static_cast<const double*>(a.data)[r];
    case LDF_TYPE_FLOAT: return (double)static_cast<const float*>(a.data)[r];
    case LDF_TYPE_BOOL:
    case LDF_TYPE_UINT8:
    case LDF_TYPE_UINT16:
    case LDF_TYPE_UINT32:
    case LDF_TYPE_UINT64: return (double)load_bits(a.data, a.width, r);
    default: return (double)load_signed(a.data, a.width, r);
  }
}

__device__ __forceinline__ bool is_float_type(int t) { return t == LDF_TYPE_DOUBLE || t == LDF_TYPE_FLOAT; }
__device__ __forceinline__ bool is_unsigned_type(int t)
{
  return t == LDF_TYPE_BOOL || t == LDF_TYPE_UINT8 || t == LDF_TYPE_UINT16 || t == LDF_TYPE_UINT32 || t == LDF_TYPE_UINT64;
}

// order-preserving map to uint64 (NaN excluded by the caller)
__device__ __forceinline__ uint64_t ordered_bits(const AggDesc& a, int64_t r)
{
  if (is_float_type(a.in_type)) {
    double d   = load_as_double(a, r);
    uint64_t b = (uint64_t)__double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
  }
  if (is_unsigned_type(a.in_type)) return load_bits(a.data, a.width, r);
  return (uint64_t)load_signed(a.data, a.width, r) ^ 0x8000000000000000ull;
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

// value bits used by COUNT_DISTINCT (floating point compared bitwise, as Arrow does)
__device__ __forceinline__ uint64_t distinct_bits(const AggDesc& a, int64_t r) { return load_bits(a.data, a.width, r); }

template <bool GENERIC>
__device__ inline void distinct_insert(const AggDesc& a, const DKeys& keys, int64_t r, int64_t slot)
{
  if (a.width > 0 && a.width <= 4) {
    unsigned long long pk = ((unsigned long long)(uint32_t)slot << 32) | (uint32_t)distinct_bits(a, r);
    uint64_t i            = mix64(pk) & a.pmask;
    while (true) {
      unsigned long long cur = atomic_load_relaxed(&a.pset[i]);
      if (cur == pk) return;
      if (cur == kGEmpty) {
        unsigned long long old = atomicCAS(&a.pset[i], kGEmpty, pk);
        if (old == kGEmpty) {
          atomicAdd(&a.cnt[slot], 1ull);
          return;
        }
        if (old == pk) return;
      }
      i = (i + 1) & a.pmask;
    }
  } else {
    // (fingerprint, representative row); equality re-checks the key tuple and the value
    uint64_t vh;
    if (a.width == 0) {
      int64_t b = a.offsets[r], e = a.offsets[r + 1];
      vh        = string_hash64(static_cast<const uint8_t*>(a.data) + b, e - b);
    } else {
      vh = mix64(distinct_bits(a, r) ^ 0x51afd7ed558ccdffull);
    }
    uint64_t h                = mix64(vh ^ mix64((uint64_t)slot + 0x9e3779b97f4a7c15ull));
    uint32_t fp               = (uint32_t)(h >> 32);
    unsigned long long packed = ((unsigned long long)fp << 32) | (uint32_t)r;
    uint64_t i                = h & a.pmask;
    while (true) {
      unsigned long long cur = atomic_load_relaxed(&a.pset[i]);
      if (cur == kGEmpty) {
        unsigned long long old = atomicCAS(&a.pset[i], kGEmpty, packed);
        if (old == kGEmpty) {
          atomicAdd(&a.cnt[slot], 1ull);
          return;
        }
        cur = old;
      }
      if ((uint32_t)(cur >> 32) == fp) {
        int64_t o = (int64_t)(uint32_t)cur;
        bool same = rows_equal(keys, r, keys, o);
        if (same) {
          if (a.width == 0) {
            DCol c;
            c.data    = a.data;
            c.offsets = a.offsets;
            c.valid   = nullptr;
            c.width   = 0;
            same      = strings_equal(c, r, c, o);
          } else {
            same = distinct_bits(a, r) == distinct_bits(a, o);
          }
        }
        if (same) return;
      }
      i = (i + 1) & a.pmask;
    }
  }
}

template <bool GENERIC>
__device__ __forceinline__ void apply_agg(const AggDesc& a, const DKeys& keys, int64_t r, int64_t s)
{
  const bool v = valid_at(a.valid, r);
  switch (a.kind) {
    case LDF_AGG_COUNT_ALL: atomicAdd(&a.cnt[s], 1ull); break;
    case LDF_AGG_COUNT: {
      bool c = (a.mode == LDF_COUNT_ALL) || (a.mode == LDF_COUNT_ONLY_NULL ? !v : v);
      if (c) atomicAdd(&a.cnt[s], 1ull);
      break;
    }
    case LDF_AGG_SUM:
    case LDF_AGG_MEAN:
    case LDF_AGG_VARIANCE:
    case LDF_AGG_STDDEV: {
      if (!v) break;
      atomicAdd(&a.cnt[s], 1ull);
      if (a.kind == LDF_AGG_SUM && !is_float_type(a.in_type)) {
        uint64_t x = is_unsigned_type(a.in_type) ? load_bits(a.data, a.width, r) : (uint64_t)load_signed(a.data, a.width, r);
        atomicAdd(&a.a0[s], (unsigned long long)x);
      } else {
        double x = load_as_double(a, r);
        uint64_t l0, l1, l2;
        if (to_fixed192(x, l0, l1, l2)) add_fixed192(&a.a0[s], &a.a1[s], &a.a2[s], l0, l1, l2);
        else *a.err = 1;
      }
      break;
    }
    case LDF_AGG_PRODUCT: {
      if (!v) break;
      atomicAdd(&a.cnt[s], 1ull);
      if (is_float_type(a.in_type)) atomic_mul_f64(&a.a0[s], load_as_double(a, r));
      else {
        uint64_t x = is_unsigned_type(a.in_type) ? load_bits(a.data, a.width, r) : (uint64_t)load_signed(a.data, a.width, r);
        atomic_mul_u64(&a.a0[s], x);
      }
      break;
    }
    case LDF_AGG_MIN:
    case LDF_AGG_MAX: {
      if (!v) break;
      atomicAdd(&a.cnt[s], 1ull);
      if (is_float_type(a.in_type) && isnan(load_as_double(a, r))) {
        a.flag[s] = 1;
        break;
      }
      uint64_t o = ordered_bits(a, r);
      if (a.kind == LDF_AGG_MIN) atomicMin(&a.a0[s], (unsigned long long)o);
      else atomicMax(&a.a0[s], (unsigned long long)o);
      break;
    }
    case LDF_AGG_COUNT_DISTINCT: {
      if (!v) {
        if (a.mode != LDF_COUNT_ONLY_VALID) a.flag[s] = 1;
        break;
      }
      if (a.mode == LDF_COUNT_ONLY_NULL) break;
      distinct_insert<GENERIC>(a, keys, r, s);
      break;
    }
    default: break;
  }
}

template <bool GENERIC>
__global__ __launch_bounds__(256) void k_gb_aggregate(DKeys keys, int64_t n, KeyTab t, AggBatch b)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = gb_find_or_insert<GENERIC>(keys, r, t);
    if (s < 0) continue;
    for (int i = 0; i < b.n; ++i) apply_agg<GENERIC>(b.d[i], keys, r, s);
  }
}

// FP64 fallback (sums that left the fixed-point range) and VARIANCE second pass.
template <bool GENERIC>
__global__ __launch_bounds__(256) void k_gb_second_pass(DKeys keys, int64_t n, KeyTab t, AggDesc a, int second)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    if (!valid_at(a.valid, r)) continue;
    int64_t s = gb_find<GENERIC>(keys, r, t);
    if (s < 0) continue;
    double x = load_as_double(a, r);
    if (second) {
      double d = x - a.mean[s];
      atomicAdd(&a.fsum[s], d * d);
    } else {
      atomicAdd(&a.fsum[s], x);
    }
  }
}

// ---- sampling: distinct-count estimate ---------------------------------------------------
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

__global__ void k_gb_sample_stats(const unsigned int* cnt, int64_t n, unsigned long long* out3)
{
  unsigned long long d = 0, f1 = 0, f2 = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    unsigned c = cnt[i];
    d += c > 0;
    f1 += c == 1;
    f2 += c == 2;
  }
  if (d) atomicAdd(&out3[0], d);
  if (f1) atomicAdd(&out3[1], f1);
  if (f2) atomicAdd(&out3[2], f2);
}

// ---- finalisation ------------------------------------------------------------------------
__global__ void k_gb_occupied(const unsigned long long* slots, int64_t nslot, uint32_t* occ)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nslot; i += (int64_t)gridDim.x * blockDim.x)
    occ[i] = slots[i] != kGEmpty;
}

// output position of every occupied slot; rep[g] = representative row (generic keys)
__global__ void k_gb_compact(const unsigned long long* slots, int64_t nslot, uint64_t cap, const uint64_t* pos, bool generic,
                             int32_t* slot_of_group, int32_t* rep)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nslot; i += (int64_t)gridDim.x * blockDim.x) {
    if (slots[i] == kGEmpty) continue;
    uint64_t g        = pos[i];
    slot_of_group[g]  = (int32_t)i;
    if (generic) rep[g] = (int32_t)(uint32_t)slots[i];
  }
}

// single fixed-width key: write the key bits of every group
__global__ void k_gb_write_key(const unsigned long long* slots, const int32_t* slot_of_group, int64_t ng, uint64_t cap,
                               int width, void* out, uint8_t* valid)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = slot_of_group[g];
    uint64_t k = (s == (int64_t)cap) ? kGEmpty : (s == (int64_t)cap + 1 ? 0 : slots[s]);
    bool ok    = s != (int64_t)cap + 1;
    switch (width) {
      case 1: static_cast<uint8_t*>(out)[g] = (uint8_t)k; break;
      case 2: static_cast<uint16_t*>(out)[g] = (uint16_t)k; break;
      case 4: static_cast<uint32_t*>(out)[g] = (uint32_t)k; break;
      default: static_cast<uint64_t*>(out)[g] = k; break;
    }
    if (valid) valid[g] = ok;
  }
}

struct FinalDesc {
  int32_t kind;
  int32_t in_type;
  int32_t out_type;
  int32_t mode;
  const unsigned long long* a0;
  const unsigned long long* a1;
  const unsigned long long* a2;
  const unsigned long long* cnt;
  const uint8_t* flag;
  const double* fsum;
  int32_t use_fsum;
};

__device__ __forceinline__ double fixed_sum(const FinalDesc& f, int64_t s)
{
  if (f.use_fsum) return f.fsum[s];
  return fixed192_to_double(f.a0[s], f.a1[s], f.a2[s]);
}

__global__ void k_gb_finalize(FinalDesc f, const int32_t* slot_of_group, int64_t ng, void* out, uint8_t* valid)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    int64_t s          = slot_of_group[g];
    unsigned long long c = f.cnt ? f.cnt[s] : 0;
    bool ok            = true;
    switch (f.kind) {
      case LDF_AGG_COUNT:
      case LDF_AGG_COUNT_ALL: static_cast<int64_t*>(out)[g] = (int64_t)c; break;
      case LDF_AGG_COUNT_DISTINCT: {
        int64_t v = (int64_t)c;
        if (f.mode == LDF_COUNT_ALL) v += f.flag[s] ? 1 : 0;
        if (f.mode == LDF_COUNT_ONLY_NULL) v = f.flag[s] ? 1 : 0;
        static_cast<int64_t*>(out)[g] = v;
        break;
      }
      case LDF_AGG_SUM:
        ok = c > 0;
        if (f.out_type == LDF_TYPE_DOUBLE) static_cast<double*>(out)[g] = ok ? fixed_sum(f, s) : 0.0;
        else static_cast<uint64_t*>(out)[g] = ok ? f.a0[s] : 0;
        break;
      case LDF_AGG_MEAN:
        ok                             = c > 0;
        static_cast<double*>(out)[g] = ok ? fixed_sum(f, s) / (double)c : 0.0;
        break;
      case LDF_AGG_VARIANCE:
      case LDF_AGG_STDDEV: {
        ok       = c > 0;
        double v = ok ? f.fsum[s] / (double)c : 0.0;
        static_cast<double*>(out)[g] = (f.kind == LDF_AGG_STDDEV) ? sqrt(v) : v;
        break;
      }
      case LDF_AGG_PRODUCT:
        ok = c > 0;
        static_cast<uint64_t*>(out)[g] = ok ? f.a0[s] : 0;
        break;
      case LDF_AGG_MIN:
      case LDF_AGG_MAX: {
        ok         = c > 0;
        uint64_t o = f.a0[s];
        uint64_t ident = (f.kind == LDF_AGG_MIN) ? ~0ull : 0ull;
        bool fl    = f.in_type == LDF_TYPE_DOUBLE || f.in_type == LDF_TYPE_FLOAT;
        uint64_t bits;
        if (fl) {
          double d;
          if (o == ident) {
            d = __longlong_as_double(0x7ff8000000000000ll);  // only NaN seen
          } else {
            uint64_t b = (o >> 63) ? (o & 0x7fffffffffffffffull) : ~o;
            d          = __longlong_as_double((long long)b);
          }
          if (f.in_type == LDF_TYPE_FLOAT) {
            float x = (float)d;
            bits    = (uint32_t)__float_as_uint(x);
          } else {
            bits = (uint64_t)__double_as_longlong(d);
          }
        } else if (f.in_type == LDF_TYPE_BOOL || f.in_type == LDF_TYPE_UINT8 || f.in_type == LDF_TYPE_UINT16 ||
                   f.in_type == LDF_TYPE_UINT32 || f.in_type == LDF_TYPE_UINT64) {
          bits = o;
        } else {
          bits = o ^ 0x8000000000000000ull;
        }
        if (!ok) bits = 0;
        switch (f.out_type) {
          case LDF_TYPE_BOOL:
          case LDF_TYPE_UINT8:
          case LDF_TYPE_INT8: static_cast<uint8_t*>(out)[g] = (uint8_t)bits; break;
          case LDF_TYPE_UINT16:
          case LDF_TYPE_INT16: static_cast<uint16_t*>(out)[g] = (uint16_t)bits; break;
          case LDF_TYPE_UINT32:
          case LDF_TYPE_INT32:
          case LDF_TYPE_FLOAT:
          case LDF_TYPE_DATE32: static_cast<uint32_t*>(out)[g] = (uint32_t)bits; break;
          default: static_cast<uint64_t*>(out)[g] = bits; break;
        }
        break;
      }
      default: break;
    }
    if (valid) valid[g] = ok;
  }
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

__global__ void k_pack_bools(const uint8_t* __restrict__ b, int64_t n, uint8_t* __restrict__ mask, unsigned long long* nulls)
{
  uint64_t z     = 0;
  int64_t nbytes = (n + 7) / 8;
  for (int64_t byte = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; byte < nbytes; byte += (int64_t)gridDim.x * blockDim.x) {
    uint8_t m = 0;
    for (int k = 0; k < 8; ++k) {
      int64_t i = byte * 8 + k;
      if (i < n) {
        if (b[i]) m |= (uint8_t)(1u << k);
        else ++z;
      }
    }
    mask[byte] = m;
  }
  if (z) atomicAdd(nulls, (unsigned long long)z);
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

struct AggState {
  Buf<unsigned long long> a0, a1, a2, cnt, pset;
  Buf<uint8_t> flag;
  Buf<double> fsum, mean;
  Buf<int32_t> err;
  uint64_t pmask = 0;
  int use_fsum   = 0;
};

template <bool GENERIC>
static void launch_aggregate(const DKeys& keys, int64_t n, const KeyTab& t, const AggBatch& b, hipStream_t s)
{
  ProfScope prof("gb_aggregate", s);
  hipLaunchKernelGGL(k_gb_aggregate<GENERIC>, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, s, keys, n, t, b);
  LDF_HIP_TRY(hipGetLastError());
}

template <bool GENERIC>
static void launch_second(const DKeys& keys, int64_t n, const KeyTab& t, const AggDesc& d, int second, hipStream_t s)
{
  hipLaunchKernelGGL(k_gb_second_pass<GENERIC>, dim3(grid_for(n, 256, 256 * 32)), dim3(256), 0, s, keys, n, t, d, second);
  LDF_HIP_TRY(hipGetLastError());
}

static uint64_t estimate_groups(const DKeys& keys, int64_t n, bool generic, Alloc& a)
{
  hipStream_t s = a.stream();
  const int64_t S = n < 65536 ? n : 65536;
  if (S == 0) return 1;
  int64_t stride = n / S;
  uint64_t scap  = next_pow2((uint64_t)S * 2);
  Buf<unsigned long long> slots(a, scap + 2);
  Buf<unsigned int> cnt(a, scap + 2);
  Buf<unsigned long long> misc(a, 5);  // nkeys, overflow(int), d, f1, f2
  fill_u64((uint64_t*)slots.get(), kGEmpty, scap + 2, s);
  fill_u32(cnt.get(), 0, scap + 2, s);
  fill_u64((uint64_t*)misc.get(), 0, 5, s);
  KeyTab t{slots.get(), scap - 1, scap, misc.get(), reinterpret_cast<int32_t*>(misc.get() + 1)};
  if (generic)
    hipLaunchKernelGGL(k_gb_sample<true>, dim3(grid_for(S, 256)), dim3(256), 0, s, keys, n, S, stride, t, cnt.get());
  else
    hipLaunchKernelGGL(k_gb_sample<false>, dim3(grid_for(S, 256)), dim3(256), 0, s, keys, n, S, stride, t, cnt.get());
  LDF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_gb_sample_stats, dim3(grid_for(scap + 2, 256)), dim3(256), 0, s, cnt.get(), (int64_t)scap + 2,
                     misc.get() + 2);
  LDF_HIP_TRY(hipGetLastError());
  unsigned long long h[3];
  LDF_HIP_TRY(hipMemcpyAsync(h, misc.get() + 2, sizeof(h), hipMemcpyDeviceToHost, s));
  LDF_HIP_TRY(hipStreamSynchronize(s));
  double d = (double)h[0], f1 = (double)h[1];
  if (S == n) return (uint64_t)d;
  double est;
  if (f1 > 0.5 * (double)S) est = d * (double)n / (double)S;  // mostly singletons: near-unique keys
  else est = std::sqrt((double)n / (double)S) * f1 + (d - f1);  // GEE
  if (est > (double)n) est = (double)n;
  if (est < d) est = d;
  return (uint64_t)est;
}

static void run_groupby(const ldf_table_view* tv, const int32_t* key_idx, int32_t nkeys, const ldf_agg* aggs, int32_t naggs,
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

  // validate aggregations
  for (int i = 0; i < naggs; ++i) {
    const ldf_agg& g = aggs[i];
    LDF_EXPECTS(g.kind >= LDF_AGG_SUM && g.kind <= LDF_AGG_COUNT_ALL, "Unsupported groupby aggregation");
    if (g.kind == LDF_AGG_APPROX_MEDIAN) throw Error(LDF_E_UNSUPPORTED, "approximate_median is not implemented");
    if (g.kind == LDF_AGG_COUNT_ALL) continue;
    LDF_EXPECTS(g.in_col >= 0 && g.in_col < tv->ncols, "groupby: aggregation input column out of range");
    int32_t t = tv->cols[g.in_col].type_id;
    if (g.kind == LDF_AGG_COUNT || g.kind == LDF_AGG_COUNT_DISTINCT) {
      LDF_EXPECTS(g.count_mode >= LDF_COUNT_ONLY_VALID && g.count_mode <= LDF_COUNT_ALL, "Unsupported count mode");
      continue;
    }
    if (type_width(t) == 0) throw Error(LDF_E_UNSUPPORTED, "groupby: numeric aggregation of a string column");
  }

  uint64_t est = estimate_groups(keys, n, generic, a);
  uint64_t cap = next_pow2(std::max<uint64_t>(1024, (uint64_t)(2.5 * (double)est)));
  uint64_t maxcap = next_pow2(std::max<uint64_t>(1024, 2 * (uint64_t)n + 2));
  if (cap > maxcap) cap = maxcap;

  for (int attempt = 0;; ++attempt) {
    LDF_EXPECTS(cap + 2 < (uint64_t)INT32_MAX, "groupby: too many groups");
    const int64_t nslot = (int64_t)cap + 2;
    Buf<unsigned long long> slots(a, nslot);
    Buf<unsigned long long> misc(a, 2);
    fill_u64((uint64_t*)slots.get(), kGEmpty, nslot, s);
    fill_u64((uint64_t*)misc.get(), 0, 2, s);
    KeyTab t{slots.get(), cap - 1, cap, misc.get(), reinterpret_cast<int32_t*>(misc.get() + 1)};

    std::vector<AggState> st(naggs);
    std::vector<AggDesc> desc(naggs);
    Buf<int32_t> errs(a, naggs > 0 ? naggs : 1);
    fill_u32(reinterpret_cast<uint32_t*>(errs.get()), 0, naggs > 0 ? naggs : 1, s);
    for (int i = 0; i < naggs; ++i) {
      const ldf_agg& g = aggs[i];
      AggDesc& d       = desc[i];
      std::memset(&d, 0, sizeof(d));
      d.kind = g.kind;
      d.mode = g.count_mode;
      if (g.kind != LDF_AGG_COUNT_ALL) {
        const ldf_column_view& c = tv->cols[g.in_col];
        d.in_type                = c.type_id;
        d.width                  = type_width(c.type_id);
        d.data                   = c.data;
        d.valid                  = c.validity;
        d.offsets                = c.offsets;
      }
      d.err = errs.get() + i;
      AggState& S = st[i];
      S.cnt       = Buf<unsigned long long>(a, nslot);
      fill_u64((uint64_t*)S.cnt.get(), 0, nslot, s);
      d.cnt = S.cnt.get();
      switch (g.kind) {
        case LDF_AGG_SUM:
        case LDF_AGG_MEAN:
        case LDF_AGG_VARIANCE:
        case LDF_AGG_STDDEV:
          S.a0 = Buf<unsigned long long>(a, nslot);
          fill_u64((uint64_t*)S.a0.get(), 0, nslot, s);
          d.a0 = S.a0.get();
          if (!(g.kind == LDF_AGG_SUM && !is_float(d.in_type))) {
            S.a1 = Buf<unsigned long long>(a, nslot);
            S.a2 = Buf<unsigned long long>(a, nslot);
            fill_u64((uint64_t*)S.a1.get(), 0, nslot, s);
            fill_u64((uint64_t*)S.a2.get(), 0, nslot, s);
            d.a1 = S.a1.get();
            d.a2 = S.a2.get();
          }
          break;
        case LDF_AGG_PRODUCT:
          S.a0 = Buf<unsigned long long>(a, nslot);
          fill_u64((uint64_t*)S.a0.get(), is_float(d.in_type) ? 0x3ff0000000000000ull : 1ull, nslot, s);
          d.a0 = S.a0.get();
          break;
        case LDF_AGG_MIN:
        case LDF_AGG_MAX:
          S.a0 = Buf<unsigned long long>(a, nslot);
          fill_u64((uint64_t*)S.a0.get(), g.kind == LDF_AGG_MIN ? ~0ull : 0ull, nslot, s);
          d.a0   = S.a0.get();
          S.flag = Buf<uint8_t>(a, nslot);
          fill_u8(S.flag.get(), 0, nslot, s);
          d.flag = S.flag.get();
          break;
        case LDF_AGG_COUNT_DISTINCT: {
          S.flag = Buf<uint8_t>(a, nslot);
          fill_u8(S.flag.get(), 0, nslot, s);
          d.flag        = S.flag.get();
          uint64_t pcap = next_pow2(std::max<uint64_t>(1024, 2 * (uint64_t)n));
          S.pset        = Buf<unsigned long long>(a, pcap);
          fill_u64((uint64_t*)S.pset.get(), kGEmpty, pcap, s);
          d.pset  = S.pset.get();
          d.pmask = pcap - 1;
          S.pmask = pcap - 1;
          break;
        }
        default: break;
      }
    }

    // one pass over the rows (aggregations in batches of kMaxAggsPerLaunch)
    if (n > 0) {
      if (naggs == 0) {
        AggBatch b;
        b.n = 0;
        if (generic) launch_aggregate<true>(keys, n, t, b, s);
        else launch_aggregate<false>(keys, n, t, b, s);
      }
      for (int i0 = 0; i0 < naggs; i0 += kMaxAggsPerLaunch) {
        AggBatch b;
        b.n = std::min(kMaxAggsPerLaunch, naggs - i0);
        for (int j = 0; j < b.n; ++j) b.d[j] = desc[i0 + j];
        if (generic) launch_aggregate<true>(keys, n, t, b, s);
        else launch_aggregate<false>(keys, n, t, b, s);
      }
    }
    int32_t ov = 0;
    LDF_HIP_TRY(hipMemcpyAsync(&ov, misc.get() + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    std::vector<int32_t> herr(naggs > 0 ? naggs : 1);
    LDF_HIP_TRY(hipMemcpyAsync(herr.data(), errs.get(), herr.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LDF_HIP_TRY(hipStreamSynchronize(s));
    if (ov) {
      LDF_EXPECTS(cap < maxcap, "groupby: key table overflow");
      cap = std::min(maxcap, cap * 4);
      continue;
    }

    // fallback FP64 sums and VARIANCE/STDDEV second passes
    for (int i = 0; i < naggs; ++i) {
      const ldf_agg& g = aggs[i];
      AggDesc& d       = desc[i];
      AggState& S      = st[i];
      bool fpsum = (g.kind == LDF_AGG_SUM && is_float(d.in_type)) || g.kind == LDF_AGG_MEAN || g.kind == LDF_AGG_VARIANCE ||
                   g.kind == LDF_AGG_STDDEV;
      if (fpsum && herr[i]) {
        S.fsum = Buf<double>(a, nslot);
        fill_u64(reinterpret_cast<uint64_t*>(S.fsum.get()), 0, nslot, s);
        d.fsum = S.fsum.get();
        if (generic) launch_second<true>(keys, n, t, d, 0, s);
        else launch_second<false>(keys, n, t, d, 0, s);
        S.use_fsum = 1;
      }
      if (g.kind == LDF_AGG_VARIANCE || g.kind == LDF_AGG_STDDEV) {
        S.mean = Buf<double>(a, nslot);
        hipLaunchKernelGGL(k_gb_means, dim3(grid_for(nslot, 256)), dim3(256), 0, s, d.a0, d.a1, d.a2, d.cnt, d.fsum,
                           S.use_fsum, nslot, S.mean.get());
        LDF_HIP_TRY(hipGetLastError());
        Buf<double> m2(a, nslot);
        fill_u64(reinterpret_cast<uint64_t*>(m2.get()), 0, nslot, s);
        d.fsum = m2.get();
        d.mean = S.mean.get();
        if (generic) launch_second<true>(keys, n, t, d, 1, s);
        else launch_second<false>(keys, n, t, d, 1, s);
        S.fsum = std::move(m2);
      }
    }

    // compact occupied slots into groups
    Buf<uint32_t> occ(a, nslot);
    Buf<uint64_t> pos(a, nslot + 1);
    hipLaunchKernelGGL(k_gb_occupied, dim3(grid_for(nslot, 256)), dim3(256), 0, s, slots.get(), nslot, occ.get());
    LDF_HIP_TRY(hipGetLastError());
    exclusive_scan_u32_to_u64(occ.get(), pos.get(), nslot, a);
    int64_t ng = (int64_t)read_u64(pos.get() + nslot, s);
    Buf<int32_t> sog(a, ng > 0 ? ng : 1);
    Buf<int32_t> rep;
    if (generic) rep = Buf<int32_t>(a, ng > 0 ? ng : 1);
    hipLaunchKernelGGL(k_gb_compact, dim3(grid_for(nslot, 256)), dim3(256), 0, s, slots.get(), nslot, cap, pos.get(),
                       generic, sog.get(), rep.get());
    LDF_HIP_TRY(hipGetLastError());

    out->ncols = nkeys + naggs;
    int done   = 0;
    try {
      Buf<uint8_t> vbytes(a, ng > 0 ? ng : 1);
      Buf<unsigned long long> nulls(a, 1);
      auto pack_validity = [&](ldf_column* c) {
        fill_u64(reinterpret_cast<uint64_t*>(nulls.get()), 0, 1, s);
        if (ng > 0) {
          hipLaunchKernelGGL(k_pack_bools, dim3(grid_for((ng + 7) / 8, 256)), dim3(256), 0, s, vbytes.get(), ng, c->validity,
                             nulls.get());
          LDF_HIP_TRY(hipGetLastError());
        }
        c->null_count = (int64_t)read_u64(reinterpret_cast<const uint64_t*>(nulls.get()), s);
      };
      // keys
      if (generic) {
        for (int i = 0; i < nkeys; ++i, ++done)
          gather_column(tv->cols[key_idx[i]], rep.get(), ng, false, a, &out->cols[done]);
      } else {
        const ldf_column_view& kc = tv->cols[key_idx[0]];
        bool nullable             = kc.validity != nullptr;
        alloc_column(kc.type_id, ng, nullable, 0, a, &out->cols[0]);
        done = 1;
        if (ng > 0) {
          hipLaunchKernelGGL(k_gb_write_key, dim3(grid_for(ng, 256)), dim3(256), 0, s, slots.get(), sog.get(), ng, cap,
                             type_width(kc.type_id), out->cols[0].data, nullable ? vbytes.get() : (uint8_t*)nullptr);
          LDF_HIP_TRY(hipGetLastError());
        }
        if (nullable) pack_validity(&out->cols[0]);
      }
      // aggregations
      for (int i = 0; i < naggs; ++i, ++done) {
        const ldf_agg& g = aggs[i];
        int32_t in_type  = g.kind == LDF_AGG_COUNT_ALL ? LDF_TYPE_INT64 : tv->cols[g.in_col].type_id;
        int32_t ot       = agg_out_type(g.kind, in_type);
        bool nullable    = !(g.kind == LDF_AGG_COUNT_ALL || g.kind == LDF_AGG_COUNT_DISTINCT);
        alloc_column(ot, ng, nullable, 0, a, &out->cols[done]);
        FinalDesc f;
        std::memset(&f, 0, sizeof(f));
        f.kind     = g.kind;
        f.in_type  = in_type;
        f.out_type = ot;
        f.mode     = g.count_mode;
        f.a0       = desc[i].a0;
        f.a1       = desc[i].a1;
        f.a2       = desc[i].a2;
        f.cnt      = desc[i].cnt;
        f.flag     = desc[i].flag;
        f.fsum     = st[i].fsum.get();
        f.use_fsum = st[i].use_fsum;
        if (ng > 0) {
          hipLaunchKernelGGL(k_gb_finalize, dim3(grid_for(ng, 256)), dim3(256), 0, s, f, sog.get(), ng,
                             out->cols[done].data, nullable ? vbytes.get() : (uint8_t*)nullptr);
          LDF_HIP_TRY(hipGetLastError());
        }
        if (nullable) pack_validity(&out->cols[done]);
      }
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

extern "C" int ldf_groupby(const ldf_table_view* table, const int32_t* keys, int32_t nkeys, const ldf_agg* aggs,
                           int32_t naggs, const ldf_allocator* alloc, void* stream, ldf_table_out* out)
{
  return guarded([&] {
    Alloc a(alloc, (hipStream_t)stream);
    run_groupby(table, keys, nkeys, aggs, naggs, a, out);
  });
}
