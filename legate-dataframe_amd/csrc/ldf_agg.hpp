// Device helpers shared by the groupby kernels (plain and batched): exact FP64 summation in a
// 192-bit fixed-point integer, value-bit conversions and typed stores.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/ldf_hip.h"

namespace ldf {

// ---- exact fixed-point FP64 accumulation ------------------------------------------------
// x -> round-toward-zero(x * 2^96) as a 192-bit two's-complement integer (l0 low .. l2 high).
// Returns false for values the fixed-point range cannot hold (non-finite or |x| >= 2^62).
__device__ __forceinline__ bool to_fixed192(double x, uint64_t& l0, uint64_t& l1, uint64_t& l2)
{
  uint64_t b = (uint64_t)__double_as_longlong(x);
  int ex     = (int)((b >> 52) & 0x7ff);
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
    l0         = ~l0 + 1;
    uint64_t c = (l0 == 0);
    l1         = ~l1 + c;
    c          = c && (l1 == 0);
    l2         = ~l2 + c;
  }
  return true;
}

// 192-bit add of (l0, l1, l2) into three words; works on LDS and global pointers alike.
__device__ __forceinline__ void add_fixed192(unsigned long long* a0, unsigned long long* a1, unsigned long long* a2,
                                             uint64_t l0, uint64_t l1, uint64_t l2)
{
  unsigned long long o0 = atomicAdd(a0, (unsigned long long)l0);
  uint64_t c0           = (o0 + l0) < o0;
  uint64_t y1           = l1 + c0;
  uint64_t c1a          = (y1 < l1);
  uint64_t c1           = c1a;
  if (y1 != 0) {
    unsigned long long o1 = atomicAdd(a1, (unsigned long long)y1);
    c1 |= (uint64_t)((o1 + y1) < o1);
  }
  uint64_t y2 = l2 + c1;
  if (y2 != 0) atomicAdd(a2, (unsigned long long)y2);
}

// LDS form of an exact sum: six signed 64-bit words, word k accumulating +-(32-bit chunk k of
// round-toward-zero(|x| * 2^96)).  A value's 53-bit mantissa touches at most three consecutive
// chunks, so an add is at most three LDS atomics without return value, whatever the sign.  A word
// stays exact while it receives fewer than 2^31 addends (callers keep the rows per table below
// that); lds_fsum_value folds the words into the 192-bit two's-complement sum once, at flush time.
// Returns false for values the fixed-point range cannot hold (non-finite or |x| >= 2^62).
__device__ __forceinline__ bool lds_fsum_add(unsigned long long* w, double x)
{
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const int ex     = (int)((b >> 52) & 0x7ff);
  if (ex == 0x7ff || ex >= 1023 + 62) return false;
  if (ex == 0) return true;  // zero or subnormal: below 2^-96, contributes 0
  uint64_t m = (b & ((1ull << 52) - 1)) | (1ull << 52);
  int shift  = ex - 1075 + 96;  // position of m's lowest bit in the fixed-point integer
  if (shift < 0) {
    m     = (-shift < 64) ? (m >> (-shift)) : 0;
    shift = 0;
  }
  const int k0      = shift >> 5;  // <= 3: shift <= 105
  const int r       = shift & 31;
  const uint64_t lo = m << r;
  const uint32_t hi = r ? (uint32_t)(m >> (64 - r)) : 0u;  // m << r < 2^84
  const bool neg    = (b >> 63) != 0;
  const uint32_t c0 = (uint32_t)lo, c1 = (uint32_t)(lo >> 32);
  if (c0) atomicAdd(&w[k0], neg ? 0ull - (unsigned long long)c0 : (unsigned long long)c0);
  if (c1) atomicAdd(&w[k0 + 1], neg ? 0ull - (unsigned long long)c1 : (unsigned long long)c1);
  if (hi) atomicAdd(&w[k0 + 2], neg ? 0ull - (unsigned long long)hi : (unsigned long long)hi);
  return true;
}

__device__ __forceinline__ void add192(uint64_t& r0, uint64_t& r1, uint64_t& r2, uint64_t a0, uint64_t a1, uint64_t a2)
{
  uint64_t t0 = r0 + a0, c0 = t0 < r0;
  uint64_t t1 = r1 + a1, c1 = t1 < r1;
  uint64_t t1c = t1 + c0;
  c1 |= t1c < t1;
  r0 = t0;
  r1 = t1c;
  r2 = r2 + a2 + c1;
}

__device__ __forceinline__ void lds_fsum_value(const unsigned long long* w, uint64_t& r0, uint64_t& r1, uint64_t& r2)
{
  r0 = r1 = r2 = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {  // r += sign_extend(w[k]) << 32k  (mod 2^192)
    const uint64_t x = w[k];
    const uint64_t e = (x >> 63) ? ~0ull : 0ull;
    switch (k) {
      case 0: add192(r0, r1, r2, x, e, e); break;
      case 1: add192(r0, r1, r2, x << 32, (x >> 32) | (e << 32), e); break;
      case 2: add192(r0, r1, r2, 0, x, e); break;
      case 3: add192(r0, r1, r2, 0, x << 32, (x >> 32) | (e << 32)); break;
      case 4: add192(r0, r1, r2, 0, 0, x); break;
      default: add192(r0, r1, r2, 0, 0, x << 32); break;
    }
  }
}

// Correctly rounded double of a 192-bit two's-complement integer times 2^-96.
__device__ inline double fixed192_to_double(uint64_t l0, uint64_t l1, uint64_t l2)
{
  bool neg = (l2 >> 63) != 0;
  if (neg) {
    l0         = ~l0 + 1;
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
  int p        = 191 - lz;  // exponent of the top bit
  uint64_t q   = t2 >> 11;
  uint64_t rem = t2 & 0x7ff;
  bool sticky  = (t1 | t0) != 0;
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
__device__ __forceinline__ bool is_float_type(int t) { return t == LDF_TYPE_DOUBLE || t == LDF_TYPE_FLOAT; }
__device__ __forceinline__ bool is_unsigned_type(int t)
{
  return t == LDF_TYPE_BOOL || t == LDF_TYPE_UINT8 || t == LDF_TYPE_UINT16 || t == LDF_TYPE_UINT32 || t == LDF_TYPE_UINT64;
}

// value bits -> double / wrapping integer / order-preserving uint64 (NaN excluded by callers)
__device__ __forceinline__ double bits_as_double(uint64_t b, int type, int width)
{
  if (type == LDF_TYPE_DOUBLE) return __longlong_as_double((long long)b);
  if (type == LDF_TYPE_FLOAT) return (double)__uint_as_float((uint32_t)b);
  if (is_unsigned_type(type)) return (double)b;
  switch (width) {
    case 1: return (double)(int8_t)b;
    case 2: return (double)(int16_t)b;
    case 4: return (double)(int32_t)b;
    default: return (double)(int64_t)b;
  }
}

__device__ __forceinline__ uint64_t bits_as_int(uint64_t b, int type, int width)
{
  if (is_unsigned_type(type)) return b;
  switch (width) {
    case 1: return (uint64_t)(int64_t)(int8_t)b;
    case 2: return (uint64_t)(int64_t)(int16_t)b;
    case 4: return (uint64_t)(int64_t)(int32_t)b;
    default: return b;
  }
}

__device__ __forceinline__ uint64_t ordered_of(uint64_t b, int type, int width)
{
  if (is_float_type(type)) {
    double d    = bits_as_double(b, type, width);
    uint64_t db = (uint64_t)__double_as_longlong(d);
    return (db >> 63) ? ~db : (db | 0x8000000000000000ull);
  }
  if (is_unsigned_type(type)) return b;
  return bits_as_int(b, type, width) ^ 0x8000000000000000ull;
}

__device__ __forceinline__ void store_bits(void* out, int type, int64_t g, uint64_t bits)
{
  switch (type) {
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
}

}  // namespace ldf
