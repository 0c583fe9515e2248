// Device helpers shared by the join, groupby and hash-partition kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ldf_common.hpp"

namespace ldf {

constexpr int kWave = 64;

__device__ __forceinline__ bool valid_at(const uint8_t* v, int64_t i)
{
  return v == nullptr || ((v[i >> 3] >> (i & 7)) & 1);
}

// Raw value bits of a fixed-width column, zero-extended to 64 bits.  Keys compare by these bits:
// Arrow's hash join and grouper compare floating-point keys bitwise (0.0 != -0.0, NaN == NaN).
__device__ __forceinline__ uint64_t load_bits(const void* data, int width, int64_t row)
{
  switch (width) {
    case 1: return static_cast<const uint8_t*>(data)[row];
    case 2: return static_cast<const uint16_t*>(data)[row];
    case 4: return static_cast<const uint32_t*>(data)[row];
    default: return static_cast<const uint64_t*>(data)[row];
  }
}

// Branch-free forms of valid_at / load_bits for loops that load several rows and columns before
// using any of them.  With a branch around each load (the null bitmap test, the switch on the
// width) the compiler waits for every load before issuing the next one.  These issue plain loads
// whose results are combined only by masks (no selects, which the compiler would turn back into
// branches around the loads), so a thread's loads are all in flight together.  `i` / `row` must
// be a row of the column (callers clamp rows past the end).

// Validity bit of row i; a column without a bitmap reads byte i / 8 of `any` instead (any buffer
// of at least that many bytes: the column's values) and reports 1.
__device__ __forceinline__ uint32_t valid_bit_nb(const uint8_t* v, const void* any, int64_t i)
{
  const uint8_t* p = (v != nullptr ? v : static_cast<const uint8_t*>(any)) + (i >> 3);
  return (((uint32_t)*p >> (uint32_t)(i & 7)) & 1u) | (uint32_t)(v == nullptr);
}

// width 8: two dwords; width 1, 2, 4: the dword holding the value (in the same page as the
// value, so never out of bounds) shifted and masked, and that dword loaded a second time
__device__ __forceinline__ uint64_t load_bits_nb(const void* data, int width, int64_t row)
{
  const uint8_t* b     = static_cast<const uint8_t*>(data) + row * (int64_t)width;
  const uint32_t mis   = (uint32_t)reinterpret_cast<uintptr_t>(b) & 3u;  // pointer arithmetic only:
  const uint32_t* w    = reinterpret_cast<const uint32_t*>(b - mis);    // the loads stay global loads
  const uint32_t wide  = (uint32_t)width >> 3;                           // 1 for width 8
  const uint32_t lo    = w[0];
  const uint32_t hi    = w[wide];
  const uint32_t m     = width >= 4 ? 0xffffffffu : ((1u << (8 * width)) - 1u);
  const uint64_t hpart = (uint64_t)(hi & (0u - wide)) << 32;
  return hpart | (uint64_t)((lo >> (8u * mis)) & m);
}

// A fixed-width column seen from a block that reads rows [begin, begin + 2^31): row begin + o is
// addressed by the 32-bit offset o.  A value is one aligned 8-byte load -- the aligned word that
// holds a naturally aligned value of width <= 8 (column data must be aligned to its width) lies in
// the value's page -- shifted and masked: one load instruction and no branch per value.
struct BlockCol {
  const uint8_t* data;   // the value of row `begin`
  const uint8_t* valid;  // the bitmap byte of row begin & ~7 (the values when there is no bitmap)
  uint64_t mask;         // all ones for width 8
  uint32_t width;
  uint32_t vshift;  // begin & 7
  uint32_t nonull;  // 1: no bitmap (every row valid)
};

__device__ __forceinline__ BlockCol block_col(const void* data, const uint8_t* valid, int width, int64_t begin)
{
  BlockCol c;
  c.data   = static_cast<const uint8_t*>(data) + begin * (int64_t)width;
  c.valid  = valid != nullptr ? valid + (begin >> 3) : c.data;
  c.mask   = width >= 8 ? ~0ull : ((1ull << (8 * width)) - 1ull);
  c.width  = (uint32_t)width;
  c.vshift = (uint32_t)(begin & 7);
  c.nonull = valid == nullptr ? 1u : 0u;
  return c;
}

__device__ __forceinline__ uint64_t bc_bits(const BlockCol& c, uint32_t o)
{
  const uint8_t* a   = c.data + (uint64_t)o * c.width;
  const uint32_t mis = (uint32_t)reinterpret_cast<uintptr_t>(a) & 7u;
  const uint64_t w   = *reinterpret_cast<const uint64_t*>(a - mis);
  return (w >> (8u * mis)) & c.mask;
}

__device__ __forceinline__ uint32_t bc_valid(const BlockCol& c, uint32_t o)
{
  const uint32_t i = o + c.vshift;
  return (((uint32_t)c.valid[i >> 3] >> (i & 7u)) & 1u) | c.nonull;
}

// Sign-extended integer value of a fixed-width integer column.
__device__ __forceinline__ int64_t load_signed(const void* data, int width, int64_t row)
{
  switch (width) {
    case 1: return static_cast<const int8_t*>(data)[row];
    case 2: return static_cast<const int16_t*>(data)[row];
    case 4: return static_cast<const int32_t*>(data)[row];
    default: return static_cast<const int64_t*>(data)[row];
  }
}

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ uint32_t fmix32(uint32_t h)
{
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ uint32_t murmur3_mix_block(uint32_t h, uint32_t k)
{
  k *= 0xcc9e2d51u;
  k = rotl32(k, 15);
  k *= 0x1b873593u;
  h ^= k;
  h = rotl32(h, 13);
  return h * 5u + 0xe6546b64u;
}

// MurmurHash3_x86_32 (Austin Appleby, public domain) of a fixed-width value's bytes.
__device__ __forceinline__ uint32_t murmur3_fixed(uint64_t bits, int width, uint32_t seed)
{
  uint32_t h = seed;
  if (width >= 4) {
    h = murmur3_mix_block(h, (uint32_t)bits);
    if (width == 8) h = murmur3_mix_block(h, (uint32_t)(bits >> 32));
  } else {
    uint32_t k1 = (uint32_t)bits;  // tail bytes, little endian
    k1 *= 0xcc9e2d51u;
    k1 = rotl32(k1, 15);
    k1 *= 0x1b873593u;
    h ^= k1;
  }
  h ^= (uint32_t)width;
  return fmix32(h);
}

// Value bits as libcudf's MurmurHash3_x86_32 hashes a floating-point element: every NaN becomes
// the canonical quiet NaN and -0.0 becomes 0.0 first (its normalize_nans_and_zeros), so values
// that compare equal land in the same partition.  `fwidth` is 4 (float), 8 (double) or 0 (not a
// floating-point column: bits unchanged).
__device__ __forceinline__ uint64_t murmur_input_bits(uint64_t bits, int fwidth)
{
  if (fwidth == 4) {
    const float f = __uint_as_float((uint32_t)bits);
    return f != f ? 0x7fc00000ull : (f == 0.0f ? 0ull : bits);
  }
  if (fwidth == 8) {
    const double d = __longlong_as_double((long long)bits);
    return d != d ? 0x7ff8000000000000ull : (d == 0.0 ? 0ull : bits);
  }
  return bits;
}

// MurmurHash3_x86_32 over an arbitrary byte string.
__device__ inline uint32_t murmur3_bytes(const uint8_t* p, int64_t len, uint32_t seed)
{
  uint32_t h       = seed;
  int64_t nblocks  = len / 4;
  for (int64_t i = 0; i < nblocks; ++i) {
    uint32_t k = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
                 ((uint32_t)p[4 * i + 3] << 24);
    h = murmur3_mix_block(h, k);
  }
  const uint8_t* tail = p + nblocks * 4;
  uint32_t k1         = 0;
  switch (len & 3) {
    case 3: k1 ^= (uint32_t)tail[2] << 16; [[fallthrough]];
    case 2: k1 ^= (uint32_t)tail[1] << 8; [[fallthrough]];
    case 1:
      k1 ^= tail[0];
      k1 *= 0xcc9e2d51u;
      k1 = rotl32(k1, 15);
      k1 *= 0x1b873593u;
      h ^= k1;
  }
  h ^= (uint32_t)len;
  return fmix32(h);
}

// a mod d for 32-bit a and a divisor fixed per launch, by multiplication (Lemire, Kaser and
// Kurz, "Faster Remainder by Direct Computation", 2019): with m = floor((2^64 - 1) / d) + 1,
// a mod d = high 64 bits of ((m * a mod 2^64) * d); exact for every a and every d >= 1.
struct FastMod {
  uint64_t m;
  uint32_t d;
  uint32_t pow2;  // d is a power of two: a mod d = a & (d - 1)
};
inline FastMod fastmod_make(uint32_t d) { return FastMod{~0ull / d + 1, d, (d & (d - 1)) == 0 ? 1u : 0u}; }
__device__ __forceinline__ uint32_t fastmod(uint32_t a, const FastMod& f)
{
  if (f.pow2) return a & (f.d - 1);
  const uint64_t low = f.m * (uint64_t)a;
  return (uint32_t)__umul64hi(low, (uint64_t)f.d);
}

// boost-style combine used by libcudf's row hasher.
__device__ __forceinline__ uint32_t hash_combine32(uint32_t lhs, uint32_t rhs)
{
  return lhs ^ (rhs + 0x9e3779b9u + (lhs << 6) + (lhs >> 2));
}

// 64-bit finaliser (MurmurHash3 fmix64) for the internal hash tables.
__device__ __forceinline__ uint64_t mix64(uint64_t k)
{
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__device__ inline uint64_t string_hash64(const uint8_t* p, int64_t len)
{
  uint64_t h = 0x9e3779b97f4a7c15ull ^ (uint64_t)len;
  int64_t i  = 0;
  for (; i + 8 <= len; i += 8) {
    uint64_t w = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) w |= (uint64_t)p[i + b] << (8 * b);
    h = mix64(h ^ w) + 0x632be59bd9b4e019ull;
  }
  uint64_t w = 0;
  for (int b = 0; i + b < len; ++b) w |= (uint64_t)p[i + b] << (8 * b);
  return mix64(h ^ w ^ 0x5851f42d4c957f2dull);
}

// Internal 64-bit hash of one key element (not the partition hash).
__device__ __forceinline__ uint64_t element_hash64(const DCol& c, int64_t row)
{
  if (!valid_at(c.valid, row)) return 0x2545f4914f6cdd1dull;
  if (c.width == 0) {
    const int64_t* off = c.offsets;
    int64_t b = off[row], e = off[row + 1];
    return string_hash64(static_cast<const uint8_t*>(c.data) + b, e - b);
  }
  return mix64(load_bits(c.data, c.width, row) + 0x9e3779b97f4a7c15ull);
}

__device__ __forceinline__ uint64_t row_hash64(const DKeys& k, int64_t row)
{
  uint64_t h = 0;
  for (int i = 0; i < k.n; ++i) {
    uint64_t e = element_hash64(k.c[i], row);
    h          = mix64(h ^ (e + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2)));
  }
  return h;
}

__device__ inline bool strings_equal(const DCol& a, int64_t ra, const DCol& b, int64_t rb)
{
  int64_t ab = a.offsets[ra], ae = a.offsets[ra + 1];
  int64_t bb = b.offsets[rb], be = b.offsets[rb + 1];
  if (ae - ab != be - bb) return false;
  const uint8_t* pa = static_cast<const uint8_t*>(a.data) + ab;
  const uint8_t* pb = static_cast<const uint8_t*>(b.data) + bb;
  for (int64_t i = 0; i < ae - ab; ++i)
    if (pa[i] != pb[i]) return false;
  return true;
}

// Key-tuple equality; a null equals a null (only reached when nulls compare equal).
__device__ inline bool rows_equal(const DKeys& a, int64_t ra, const DKeys& b, int64_t rb)
{
  for (int i = 0; i < a.n; ++i) {
    bool va = valid_at(a.c[i].valid, ra), vb = valid_at(b.c[i].valid, rb);
    if (va != vb) return false;
    if (!va) continue;
    if (a.c[i].width == 0) {
      if (!strings_equal(a.c[i], ra, b.c[i], rb)) return false;
    } else if (load_bits(a.c[i].data, a.c[i].width, ra) != load_bits(b.c[i].data, b.c[i].width, rb)) {
      return false;
    }
  }
  return true;
}

__device__ __forceinline__ bool any_key_null(const DKeys& k, int64_t row)
{
  for (int i = 0; i < k.n; ++i)
    if (!valid_at(k.c[i].valid, row)) return true;
  return false;
}

// Relaxed device-scope atomic load (bypasses the non-coherent vector L1).
template <typename T>
__device__ __forceinline__ T atomic_load_relaxed(const T* p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------------------------------
// block-wide scans / reductions (blockDim.x a multiple of 64, <= 1024)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t wave_inclusive_scan(uint64_t v)
{
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan across the block; returns the thread's exclusive prefix, *total = block sum.
// `smem` needs 17 uint64 entries.
__device__ inline uint64_t block_exclusive_scan(uint64_t v, uint64_t* smem, uint64_t* total)
{
  const int lane  = threadIdx.x & (kWave - 1);
  const int wave  = threadIdx.x / kWave;
  const int waves = blockDim.x / kWave;
  uint64_t inc    = wave_inclusive_scan(v);
  if (lane == kWave - 1) smem[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t run = 0;
    for (int w = 0; w < waves; ++w) {
      uint64_t t = smem[w];
      smem[w]    = run;
      run += t;
    }
    smem[16] = run;
  }
  __syncthreads();
  uint64_t res = smem[wave] + inc - v;
  *total       = smem[16];
  __syncthreads();
  return res;
}

__device__ inline uint64_t block_reduce_sum(uint64_t v, uint64_t* smem)
{
  uint64_t total;
  (void)block_exclusive_scan(v, smem, &total);
  return total;
}

}  // namespace ldf
