// Fixed-size records of a few fields, built and read in registers: a row's fields packed into
// one power-of-two record (8..64 bytes) so a scattered write or a random read of the row is one
// aligned vector access instead of one access per field.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ldf {

// A record being built in registers: eight named words (no private array, which a dynamically
// indexed access would place in scratch memory).  `off` is the same for every lane, so put()
// branches on it with scalar branches; values come zero-extended from their column.
struct RecWords {
  uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0, w6 = 0, w7 = 0;
  __device__ __forceinline__ void put(int off, uint64_t bits)
  {
    const uint64_t sv = bits << ((off & 7) * 8);
    switch (off >> 3) {
      case 0: w0 |= sv; break;
      case 1: w1 |= sv; break;
      case 2: w2 |= sv; break;
      case 3: w3 |= sv; break;
      case 4: w4 |= sv; break;
      case 5: w5 |= sv; break;
      case 6: w6 |= sv; break;
      default: w7 |= sv; break;
    }
  }
  __device__ __forceinline__ void store(unsigned char* base, int stride, int64_t i) const
  {
    unsigned char* d = base + i * (int64_t)stride;
    if (stride == 8) {
      *reinterpret_cast<uint64_t*>(d) = w0;
      return;
    }
    uint4* q = reinterpret_cast<uint4*>(d);
    q[0]     = make_uint4((uint32_t)w0, (uint32_t)(w0 >> 32), (uint32_t)w1, (uint32_t)(w1 >> 32));
    if (stride >= 32) q[1] = make_uint4((uint32_t)w2, (uint32_t)(w2 >> 32), (uint32_t)w3, (uint32_t)(w3 >> 32));
    if (stride == 64) {
      q[2] = make_uint4((uint32_t)w4, (uint32_t)(w4 >> 32), (uint32_t)w5, (uint32_t)(w5 >> 32));
      q[3] = make_uint4((uint32_t)w6, (uint32_t)(w6 >> 32), (uint32_t)w7, (uint32_t)(w7 >> 32));
    }
  }
};

// A field of a loaded record.  `off` and `width` are the same for every lane (launch
// arguments), so the word is chosen by a scalar branch rather than by selecting over every word.
template <int W>
__device__ __forceinline__ uint64_t rec_get(const uint64_t (&w)[W], int off, int width)
{
  const int wi = __builtin_amdgcn_readfirstlane(off) >> 3;
  uint64_t x;
  switch (wi) {
    case 0: x = w[0]; break;
    case 1: x = w[W > 1 ? 1 : 0]; break;
    case 2: x = w[W > 2 ? 2 : 0]; break;
    case 3: x = w[W > 3 ? 3 : 0]; break;
    case 4: x = w[W > 4 ? 4 : 0]; break;
    case 5: x = w[W > 5 ? 5 : 0]; break;
    case 6: x = w[W > 6 ? 6 : 0]; break;
    default: x = w[W > 7 ? 7 : 0]; break;
  }
  x >>= (off & 7) * 8;
  return width >= 8 ? x : x & ((1ull << (8 * width)) - 1);
}

template <int W>
__device__ __forceinline__ void rec_load(const unsigned char* base, int64_t i, uint64_t (&w)[W])
{
  const unsigned char* d = base + i * (int64_t)(8 * W);
  if (W == 1) {
    w[0] = *reinterpret_cast<const uint64_t*>(d);
    return;
  }
  const uint4* q = reinterpret_cast<const uint4*>(d);
#pragma unroll
  for (int k = 0; k < W / 2; ++k) {
    const uint4 x = q[k];
    w[2 * k]      = (uint64_t)x.x | ((uint64_t)x.y << 32);
    w[2 * k + 1]  = (uint64_t)x.z | ((uint64_t)x.w << 32);
  }
}

// Lays out fields of the given byte widths (1, 2, 4 or 8), widest first, each at a multiple of
// its width; fills offs[i] and returns the record stride (a power of two, >= 8), or -1 when the
// fields take more than 64 bytes.
inline int rec_layout(const int* widths, int n, int* offs)
{
  int off = 0;
  for (int w = 8; w >= 1; w >>= 1)
    for (int i = 0; i < n; ++i)
      if (widths[i] == w) {
        offs[i] = off;
        off += w;
      }
  int stride = 8;
  while (stride < off) stride *= 2;
  return stride <= 64 ? stride : -1;
}

}  // namespace ldf
