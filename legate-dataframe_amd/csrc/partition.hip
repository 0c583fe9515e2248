// Local murmur3 hash partition: one GPU's share of task::repartition_by_hash.
//
// Reference: cpp/src/core/repartition_by_hash.cu:296-310 calls
//   cudf::hash_partition(table, cols, nparts, HASH_MURMUR3, DEFAULT_HASH_SEED)
// and splits the result at the returned offsets.  The row hash restated here is libcudf's:
// h = seed; for each hashed column h = hash_combine(h, MurmurHash3_x86_32(value, seed)),
// a null element hashing to 0xFFFFFFFF, a float/double element hashed after its NaNs are made
// the canonical quiet NaN and -0.0 is made 0.0; the partition is h mod nparts.
//
// Kernels (all HBM-streaming, no MFMA):
//   k_hp_hist    block b counts the partition ids of its contiguous row range in LDS and
//                writes counts[p * nblk + b] (partition-major, so one exclusive scan of the
//                whole array yields every (partition, block) write cursor);
//   k_hp_scatter the same rows again, 8192 at a time: each wave ranks its 512 rows by
//                partition (stable: rank = rows of the same partition earlier in the wave's
//                chunk), the ranks are combined across the 16 waves in LDS, rows are staged
//                in partition order in LDS and then written out as contiguous runs, so rows
//                keep their input order inside a partition.  A group of 64 rows finds the
//                lanes of each partition by having every lane OR its bit into a 64-bit LDS
//                word per (wave, partition) and read it back: three LDS operations instead
//                of one ballot and its mask arithmetic per partition bit (the ballots are kept
//                for partition counts whose words do not fit the idle staging area).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ldf_common.hpp"
#include "ldf_device.hpp"

namespace ldf {

constexpr int kHpThreads = 1024;                   // 16 waves
constexpr int kHpWaves   = kHpThreads / kWave;
constexpr int kHpSub     = 8192;                   // rows staged in LDS per step
constexpr int kHpPerWave = kHpSub / kHpWaves;      // 512 contiguous rows per wave
constexpr int kHpGroups  = kHpPerWave / kWave;     // 8 groups of 64
constexpr int kHpMaxParts = 2048;

struct HashCols {
  DCol c[kMaxKeyCols];
  int32_t n;
  uint32_t seed;
};

__device__ __forceinline__ uint32_t partition_row_hash(const HashCols& hc, int64_t row)
{
  uint32_t h = hc.seed;
  for (int i = 0; i < hc.n; ++i) {
    const DCol& c = hc.c[i];
    uint32_t e;
    if (!valid_at(c.valid, row)) {
      e = 0xFFFFFFFFu;
    } else if (c.width == 0) {
      int64_t b = c.offsets[row], en = c.offsets[row + 1];
      e = murmur3_bytes(static_cast<const uint8_t*>(c.data) + b, en - b, hc.seed);
    } else {
      const int fw = c.type == LDF_TYPE_FLOAT ? 4 : (c.type == LDF_TYPE_DOUBLE ? 8 : 0);
      e = murmur3_fixed(murmur_input_bits(load_bits(c.data, c.width, row), fw), c.width, hc.seed);
    }
    h = hash_combine32(h, e);
  }
  return h;
}

__global__ void k_hp_hash(HashCols hc, int64_t n, uint32_t* __restrict__ out)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = partition_row_hash(hc, i);
}

__global__ __launch_bounds__(kHpThreads) void k_hp_hist(HashCols hc, int64_t n, int64_t rows_per_block, uint32_t nparts,
                                                        FastMod fm, uint32_t* __restrict__ counts)
{
  extern __shared__ uint32_t hist[];
  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) hist[p] = 0;
  __syncthreads();
  int64_t begin = blockIdx.x * rows_per_block;
  int64_t end   = begin + rows_per_block < n ? begin + rows_per_block : n;
  for (int64_t r = begin + threadIdx.x; r < end; r += blockDim.x) {
    uint32_t p = fastmod(partition_row_hash(hc, r), fm);
    atomicAdd(&hist[p], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) counts[(int64_t)p * gridDim.x + blockIdx.x] = hist[p];
}

// The histogram of one fixed-width non-null column hashed on itself: eight rows per thread
// have their loads in flight together.  FW = 4 / 8 for a float / double column (its bits are
// normalised before hashing), else 0.
template <typename T, int FW>
__global__ __launch_bounds__(kHpThreads) void k_hp_hist_v(const T* __restrict__ src, int64_t n, uint32_t seed,
                                                          int64_t rows_per_block, uint32_t nparts, FastMod fm,
                                                          uint32_t* __restrict__ counts)
{
  extern __shared__ uint32_t hist[];
  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) hist[p] = 0;
  __syncthreads();
  const int64_t begin = blockIdx.x * rows_per_block;
  const int64_t end   = begin + rows_per_block < n ? begin + rows_per_block : n;
  constexpr int R     = 8;
  for (int64_t r0 = begin + threadIdx.x; r0 < end; r0 += (int64_t)R * blockDim.x) {
    T v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int64_t r = r0 + (int64_t)j * blockDim.x;
      v[j]            = r < end ? src[r] : T(0);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (r0 + (int64_t)j * blockDim.x >= end) break;
      const uint32_t p =
        fastmod(hash_combine32(seed, murmur3_fixed(murmur_input_bits((uint64_t)v[j], FW), (int)sizeof(T), seed)), fm);
      atomicAdd(&hist[p], 1u);
    }
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) counts[(int64_t)p * gridDim.x + blockIdx.x] = hist[p];
}

struct ScatterCols {
  const void* src[kMaxKeyCols];
  void* dst[kMaxKeyCols];
  int32_t width[kMaxKeyCols];
  int32_t n;
};

// LDS layout (dynamic): cursor u64[P] | wtab u16[W][P] | total u32[P] | start u32[P] |
//                       stage_row u16[kHpSub] | stage_part u16[kHpSub]
template <bool ORM>
__global__ __launch_bounds__(kHpThreads) void k_hp_scatter(HashCols hc, int64_t n, int64_t rows_per_block, uint32_t nparts,
                                                           FastMod fm, int pbits, const uint64_t* __restrict__ offsets,
                                                           ScatterCols sc, int64_t* __restrict__ perm)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* cursor   = reinterpret_cast<uint64_t*>(smem_raw);
  uint16_t* wtab     = reinterpret_cast<uint16_t*>(cursor + nparts);
  uint32_t* total    = reinterpret_cast<uint32_t*>(wtab + kHpWaves * nparts + ((kHpWaves * nparts) & 1));
  uint32_t* start    = total + nparts;
  uint16_t* stage_r  = reinterpret_cast<uint16_t*>(start + nparts);
  uint16_t* stage_p  = stage_r + kHpSub;
  __shared__ uint64_t scan_smem[17];

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) cursor[p] = offsets[(int64_t)p * gridDim.x + blockIdx.x];

  int64_t begin = blockIdx.x * rows_per_block;
  int64_t end   = begin + rows_per_block < n ? begin + rows_per_block : n;

  for (int64_t sub = begin; sub < end; sub += kHpSub) {
    const int sub_n = (int)((end - sub) < kHpSub ? (end - sub) : kHpSub);
    uint64_t* wmask = reinterpret_cast<uint64_t*>(stage_r);  // ORM: lane masks, in the idle staging area
    for (uint32_t i = threadIdx.x; i < kHpWaves * nparts; i += blockDim.x) {
      wtab[i] = 0;
      if (ORM) wmask[i] = 0;
    }
    __syncthreads();

    // 1. per-wave stable ranks over the wave's contiguous 512-row chunk
    uint16_t my_part[kHpGroups];
    uint16_t my_rank[kHpGroups];
#pragma unroll
    for (int g = 0; g < kHpGroups; ++g) {
      const int local = wave * kHpPerWave + g * kWave + lane;
      const bool active = local < sub_n;
      uint32_t p = 0;
      if (active) p = fastmod(partition_row_hash(hc, sub + local), fm);
      if (ORM) {  // as in k_hp_scatter_v
        uint32_t rank_in_group = 0, old = 0;
        if (active) {
          uint64_t* w = wmask + wave * nparts + p;
          __hip_atomic_fetch_or(w, 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const uint64_t same  = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          old                  = wtab[wave * nparts + p];
          const uint64_t below = same & lanes_below;
          rank_in_group        = (uint32_t)__popcll(below);
          if (below == 0) wtab[wave * nparts + p] = (uint16_t)(old + (uint32_t)__popcll(same));
          __hip_atomic_store(w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        my_part[g] = (uint16_t)p;
        my_rank[g] = (uint16_t)(old + rank_in_group);
        continue;
      }
      uint64_t same = __ballot(active);
      for (int b = 0; b < pbits; ++b) {
        uint64_t bb = __ballot(active && ((p >> b) & 1));
        same &= ((p >> b) & 1) ? bb : ~bb;
      }
      uint32_t rank_in_group = (uint32_t)__popcll(same & lanes_below);
      uint32_t cnt           = (uint32_t)__popcll(same);
      int leader             = __ffsll((unsigned long long)same) - 1;
      uint32_t old           = 0;
      if (active && leader == lane) {
        old = wtab[wave * nparts + p];
        wtab[wave * nparts + p] = (uint16_t)(old + cnt);
      }
      old = __shfl(old, leader < 0 ? 0 : leader, kWave);
      my_part[g] = (uint16_t)p;
      my_rank[g] = (uint16_t)(old + rank_in_group);
    }
    __syncthreads();

    // 2. prefix over waves for every partition; per-partition totals of this step
    for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) {
      uint32_t run = 0;
      for (int w = 0; w < kHpWaves; ++w) {
        uint32_t c = wtab[w * nparts + p];
        wtab[w * nparts + p] = (uint16_t)run;
        run += c;
      }
      total[p] = run;
    }
    __syncthreads();

    // 3. exclusive scan of the totals -> start of each partition in the staging area
    {
      uint32_t per = (nparts + blockDim.x - 1) / blockDim.x;
      uint32_t p0  = threadIdx.x * per;
      uint64_t s   = 0;
      for (uint32_t k = 0; k < per; ++k)
        if (p0 + k < nparts) s += total[p0 + k];
      uint64_t tot;
      uint64_t run = block_exclusive_scan(s, scan_smem, &tot);
      for (uint32_t k = 0; k < per; ++k)
        if (p0 + k < nparts) {
          start[p0 + k] = (uint32_t)run;
          run += total[p0 + k];
        }
    }
    __syncthreads();

    // 4. stage row ids in partition order
#pragma unroll
    for (int g = 0; g < kHpGroups; ++g) {
      const int local = wave * kHpPerWave + g * kWave + lane;
      if (local < sub_n) {
        uint32_t p   = my_part[g];
        uint32_t pos = start[p] + wtab[wave * nparts + p] + my_rank[g];
        stage_r[pos] = (uint16_t)local;
        stage_p[pos] = (uint16_t)p;
      }
    }
    __syncthreads();

    // 5. write contiguous runs
    for (int s = threadIdx.x; s < sub_n; s += blockDim.x) {
      uint32_t p     = stage_p[s];
      int64_t src    = sub + stage_r[s];
      int64_t dst    = (int64_t)cursor[p] + (s - (int64_t)start[p]);
      for (int c = 0; c < sc.n; ++c) {
        switch (sc.width[c]) {
          case 1: static_cast<uint8_t*>(sc.dst[c])[dst] = static_cast<const uint8_t*>(sc.src[c])[src]; break;
          case 2: static_cast<uint16_t*>(sc.dst[c])[dst] = static_cast<const uint16_t*>(sc.src[c])[src]; break;
          case 4: static_cast<uint32_t*>(sc.dst[c])[dst] = static_cast<const uint32_t*>(sc.src[c])[src]; break;
          default: static_cast<uint64_t*>(sc.dst[c])[dst] = static_cast<const uint64_t*>(sc.src[c])[src]; break;
        }
      }
      if (perm != nullptr) perm[dst] = src;
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) cursor[p] += total[p];
    __syncthreads();
  }
}

// One fixed-width, non-null column: the same stable scatter with 512-thread blocks over
// 7168-row steps, but each thread keeps the values it loaded for hashing in registers and the
// values themselves are staged in LDS, so the column is read once here and written once as
// contiguous runs.  Two blocks share a CU (an int64 step stages 70 KB), so one block's ranking
// overlaps the other's memory traffic; the step is as long as that allows, because longer steps
// write longer runs per partition (28 rows of 256 partitions at a time; 4096-row steps took 6.05 ms
// at 1B keys, 6144 5.37 ms, 7168 5.32 ms; larger steps leave one block per CU or spill registers).
// LDS layout (dynamic): cursor u64[P] | wtab u16[W][P] | total u32[P] | start u32[P] |
//                       stage_val T[kVSub] | stage_part u16[kVSub]
constexpr int kVThreads = 512;
constexpr int kVWaves   = kVThreads / kWave;
constexpr int kVSub     = 7168;
constexpr int kVPerWave = kVSub / kVWaves;  // 896 contiguous rows per wave
constexpr int kVGroups  = kVPerWave / kWave;

template <typename T, int FW, bool ORM>
__global__ __launch_bounds__(kVThreads, 4) void k_hp_scatter_v(const T* __restrict__ src, int64_t n, uint32_t seed,
                                                            int64_t rows_per_block, uint32_t nparts, FastMod fm, int pbits,
                                                            const uint64_t* __restrict__ offsets, T* __restrict__ dst)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* cursor  = reinterpret_cast<uint64_t*>(smem_raw);
  uint16_t* wtab    = reinterpret_cast<uint16_t*>(cursor + nparts);
  uint32_t* total   = reinterpret_cast<uint32_t*>(wtab + kVWaves * nparts + ((kVWaves * nparts) & 1));
  uint32_t* start   = total + nparts;
  T* stage_v        = reinterpret_cast<T*>(reinterpret_cast<unsigned char*>(start + nparts) +
                                    ((16 - (((uintptr_t)(start + nparts)) & 15)) & 15));
  uint16_t* stage_p = reinterpret_cast<uint16_t*>(stage_v + kVSub);
  __shared__ uint64_t scan_smem[17];

  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) cursor[p] = offsets[(int64_t)p * gridDim.x + blockIdx.x];

  const int64_t begin = blockIdx.x * rows_per_block;
  const int64_t end   = begin + rows_per_block < n ? begin + rows_per_block : n;

  for (int64_t sub = begin; sub < end; sub += kVSub) {
    const int sub_n = (int)((end - sub) < kVSub ? (end - sub) : kVSub);
    T my_val[kVGroups];
#pragma unroll
    for (int g = 0; g < kVGroups; ++g) {
      const int64_t r = sub + wave * kVPerWave + g * kWave + lane;
      my_val[g]       = r < end ? src[r] : T(0);
    }
    uint64_t* wmask = reinterpret_cast<uint64_t*>(stage_v);  // ORM: lane masks, in the idle staging area
    for (uint32_t i = threadIdx.x; i < kVWaves * nparts; i += blockDim.x) {
      wtab[i] = 0;
      if (ORM) wmask[i] = 0;
    }
    __syncthreads();

    // 1. partition of each row; stable ranks within the wave's 512 rows
    uint32_t my_pr[kVGroups];  // partition << 16 | rank, one register per row
#pragma unroll
    for (int g = 0; g < kVGroups; ++g) {
      const int local   = wave * kVPerWave + g * kWave + lane;
      const bool active = local < sub_n;
      uint32_t p        = 0;
      if (active)
        p = fastmod(hash_combine32(seed, murmur3_fixed(murmur_input_bits((uint64_t)my_val[g], FW), (int)sizeof(T), seed)),
                    fm);
      if (ORM) {
        // the lanes of partition p: each lane ORs its bit into the wave's word for p and reads
        // the word back (LDS instructions of one wave complete in issue order, so the read sees
        // every lane's bit), then the word is cleared for the next group
        uint32_t rank_in_group = 0, old = 0;
        if (active) {
          uint64_t* w = wmask + wave * nparts + p;
          __hip_atomic_fetch_or(w, 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const uint64_t same = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          old                 = wtab[wave * nparts + p];
          const uint64_t below = same & lanes_below;
          rank_in_group        = (uint32_t)__popcll(below);
          if (below == 0) wtab[wave * nparts + p] = (uint16_t)(old + (uint32_t)__popcll(same));
          __hip_atomic_store(w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        my_pr[g] = (p << 16) | (old + rank_in_group);
        continue;
      }
      uint64_t same = __ballot(active);
      for (int b = 0; b < pbits; ++b) {
        uint64_t bb = __ballot(active && ((p >> b) & 1));
        same &= ((p >> b) & 1) ? bb : ~bb;
      }
      uint32_t rank_in_group = (uint32_t)__popcll(same & lanes_below);
      uint32_t cnt           = (uint32_t)__popcll(same);
      int leader             = __ffsll((unsigned long long)same) - 1;
      uint32_t old           = 0;
      if (active && leader == lane) {
        old                     = wtab[wave * nparts + p];
        wtab[wave * nparts + p] = (uint16_t)(old + cnt);
      }
      old        = __shfl(old, leader < 0 ? 0 : leader, kWave);
      my_pr[g] = (p << 16) | (old + rank_in_group);
    }
    __syncthreads();

    // 2. prefix over waves for every partition; per-partition totals of this step
    for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) {
      uint32_t run = 0;
#pragma unroll
      for (int w = 0; w < kVWaves; ++w) {
        uint32_t c           = wtab[w * nparts + p];
        wtab[w * nparts + p] = (uint16_t)run;
        run += c;
      }
      total[p] = run;
    }
    __syncthreads();

    // 3. exclusive scan of the totals -> start of each partition in the staging area
    {
      uint32_t per = (nparts + blockDim.x - 1) / blockDim.x;
      uint32_t p0  = threadIdx.x * per;
      uint64_t s   = 0;
      for (uint32_t k = 0; k < per; ++k)
        if (p0 + k < nparts) s += total[p0 + k];
      uint64_t tot;
      uint64_t run = block_exclusive_scan(s, scan_smem, &tot);
      for (uint32_t k = 0; k < per; ++k)
        if (p0 + k < nparts) {
          start[p0 + k] = (uint32_t)run;
          run += total[p0 + k];
        }
    }
    __syncthreads();

    // 4. stage values in partition order
#pragma unroll
    for (int g = 0; g < kVGroups; ++g) {
      const int local = wave * kVPerWave + g * kWave + lane;
      if (local < sub_n) {
        const uint32_t p   = my_pr[g] >> 16;
        const uint32_t pos = start[p] + wtab[wave * nparts + p] + (my_pr[g] & 0xFFFFu);
        stage_v[pos]       = my_val[g];
        stage_p[pos]       = (uint16_t)p;
      }
    }
    __syncthreads();

    // 5. write contiguous runs
    for (int s = threadIdx.x; s < sub_n; s += blockDim.x) {
      const uint32_t p = stage_p[s];
      dst[(int64_t)cursor[p] + (s - (int64_t)start[p])] = stage_v[s];
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) cursor[p] += total[p];
    __syncthreads();
  }
}

__global__ void k_hp_pick_offsets(const uint64_t* __restrict__ scan, uint32_t nparts, uint32_t nblk,
                                  uint64_t* __restrict__ out)
{
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= nparts; p += gridDim.x * blockDim.x)
    out[p] = scan[(int64_t)p * nblk];
}

__global__ void k_i64_to_i32(const int64_t* __restrict__ in, int64_t n, int32_t* __restrict__ out)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (int32_t)in[i];
}

static HashCols make_hash_cols(const ldf_table_view* t, const int32_t* cols, int32_t ncols, uint32_t seed)
{
  LDF_EXPECTS(ncols >= 1 && ncols <= kMaxKeyCols, "hash partition: between 1 and 8 hashed columns");
  HashCols hc;
  hc.n    = ncols;
  hc.seed = seed;
  for (int i = 0; i < ncols; ++i) {
    LDF_EXPECTS(cols[i] >= 0 && cols[i] < t->ncols, "hash partition: column index out of range");
    hc.c[i] = to_dcol(t->cols[cols[i]]);
  }
  return hc;
}

static size_t scatter_v_lds_bytes(uint32_t nparts, int width)
{
  size_t b = nparts * 8 + (size_t)kVWaves * nparts * 2 + (((size_t)kVWaves * nparts) & 1) * 2 + nparts * 8;
  b        = (b + 15) & ~size_t(15);
  return ((b + (size_t)kVSub * (width + 2)) + 15) & ~size_t(15);
}
constexpr size_t kHpVLdsMax = 150 * 1024;

template <typename T, int FW>
static void launch_scatter_v(const ldf_column_view& c, int64_t n, uint32_t seed, int64_t rpb, int64_t nblk,
                             uint32_t nparts, int pbits, const uint64_t* scan, void* dst, hipStream_t s)
{
  static bool attr = false;
  if (!attr) {
    LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_hp_scatter_v<T, FW, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kHpVLdsMax));
    LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_hp_scatter_v<T, FW, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kHpVLdsMax));
    attr = true;
  }
  const size_t lds = scatter_v_lds_bytes(nparts, sizeof(T));
  const bool orm   = (size_t)kVWaves * nparts * 8 <= (size_t)kVSub * sizeof(T);  // the lane masks fit the staging area
  if (orm)
    hipLaunchKernelGGL((k_hp_scatter_v<T, FW, true>), dim3(nblk), dim3(kVThreads), lds, s,
                       static_cast<const T*>(c.data), n, seed, rpb, nparts, fastmod_make(nparts), pbits, scan,
                       static_cast<T*>(dst));
  else
    hipLaunchKernelGGL((k_hp_scatter_v<T, FW, false>), dim3(nblk), dim3(kVThreads), lds, s,
                       static_cast<const T*>(c.data), n, seed, rpb, nparts, fastmod_make(nparts), pbits, scan,
                       static_cast<T*>(dst));
  LDF_HIP_TRY(hipGetLastError());
}

static size_t scatter_lds_bytes(uint32_t nparts)
{
  size_t b = nparts * 8 + (size_t)kHpWaves * nparts * 2 + (((size_t)kHpWaves * nparts) & 1) * 2 + nparts * 8 + kHpSub * 4;
  return (b + 15) & ~size_t(15);
}

// The generic scatter; with at most 256 partitions the lane masks of its 16 waves (8 bytes per
// wave and partition) fit the row staging area and rank the rows.
static void launch_scatter(int64_t nblk, uint32_t nparts, hipStream_t s, const HashCols& hc, int64_t n, int64_t rpb,
                           const FastMod& fm, int pbits, const uint64_t* scan, const ScatterCols& sc, int64_t* perm)
{
  static bool done = false;
  if (!done) {
    LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_hp_scatter<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)scatter_lds_bytes(kHpMaxParts)));
    LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_hp_scatter<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)scatter_lds_bytes(kHpMaxParts)));
    done = true;
  }
  if ((size_t)kHpWaves * nparts * 8 <= (size_t)kHpSub * 4)  // stage_row and stage_part, contiguous
    hipLaunchKernelGGL(k_hp_scatter<true>, dim3(nblk), dim3(kHpThreads), scatter_lds_bytes(nparts), s, hc, n, rpb,
                       nparts, fm, pbits, scan, sc, perm);
  else
    hipLaunchKernelGGL(k_hp_scatter<false>, dim3(nblk), dim3(kHpThreads), scatter_lds_bytes(nparts), s, hc, n, rpb,
                       nparts, fm, pbits, scan, sc, perm);
  LDF_HIP_TRY(hipGetLastError());
}

}  // namespace ldf

using namespace ldf;

extern "C" int ldf_murmur3_hash(const ldf_table_view* table, const int32_t* cols, int32_t ncols, uint32_t seed,
                                uint32_t* out_hashes, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(table != nullptr && out_hashes != nullptr, "NULL argument");
    int64_t n = table->ncols > 0 ? table->cols[0].length : 0;
    for (int i = 0; i < table->ncols; ++i) validate_view(table->cols[i], n, "murmur3_hash");
    HashCols hc = make_hash_cols(table, cols, ncols, seed);
    if (n == 0) return;
    hipLaunchKernelGGL(k_hp_hash, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, hc, n, out_hashes);
    LDF_HIP_TRY(hipGetLastError());
  });
}

namespace ldf {

// Blocks of the partition kernels for n rows: rows_per_block is a multiple of the staging step.
static void partition_blocks(int64_t n, int64_t max_blocks, int64_t* nblk, int64_t* rpb)
{
  int64_t b = (n + kHpSub - 1) / kHpSub;
  if (b > max_blocks) b = max_blocks;
  if (b < 1) b = 1;
  int64_t r = (n + b - 1) / b;
  r         = (r + kHpSub - 1) / kHpSub * kHpSub;
  *rpb      = r;
  *nblk     = (n + r - 1) / r;
}

// Histogram, scan and scatter of a table whose every column is fixed width and non-nullable
// (`dst[i]` receives column i, rows grouped by partition), with the scratch given: counts
// u32[nparts * nblk], scan u64[nparts * nblk + 1], ws u64[scan_workspace_elems(nparts * nblk)],
// offs u64[nparts + 1] (the partition start rows, on the device).
static void partition_fixed(const ldf_table_view* table, const HashCols& hc, const int32_t* cols, int32_t ncols,
                            int64_t n, int32_t nparts, uint32_t seed, int64_t nblk, int64_t rpb, void* const* dst,
                            uint32_t* counts, uint64_t* scan, uint64_t* ws, uint64_t* offs, bool profile, hipStream_t s)
{
  int pbits = 0;
  while ((1u << pbits) < (uint32_t)nparts) ++pbits;
  const int w0      = type_width(table->cols[0].type_id);
  const int fw0     = table->cols[0].type_id == LDF_TYPE_FLOAT ? 4 : (table->cols[0].type_id == LDF_TYPE_DOUBLE ? 8 : 0);
  // one fixed-width column hashed on itself: the value-staged kernels
  const bool single = table->ncols == 1 && ncols == 1 && cols[0] == 0 &&
                      scatter_v_lds_bytes((uint32_t)nparts, w0) <= kHpVLdsMax;
  const FastMod fm  = fastmod_make((uint32_t)nparts);
  {
    ProfScope prof(profile ? "hp_hist" : nullptr, s);
    const size_t hl = nparts * sizeof(uint32_t);
    const void* d0  = table->cols[0].data;
    if (!single)
      hipLaunchKernelGGL(k_hp_hist, dim3(nblk), dim3(kHpThreads), hl, s, hc, n, rpb, (uint32_t)nparts, fm, counts);
    else if (w0 == 1)
      hipLaunchKernelGGL((k_hp_hist_v<uint8_t, 0>), dim3(nblk), dim3(kHpThreads), hl, s, (const uint8_t*)d0, n, seed,
                         rpb, (uint32_t)nparts, fm, counts);
    else if (w0 == 2)
      hipLaunchKernelGGL((k_hp_hist_v<uint16_t, 0>), dim3(nblk), dim3(kHpThreads), hl, s, (const uint16_t*)d0, n, seed,
                         rpb, (uint32_t)nparts, fm, counts);
    else if (w0 == 4 && fw0 == 4)
      hipLaunchKernelGGL((k_hp_hist_v<uint32_t, 4>), dim3(nblk), dim3(kHpThreads), hl, s, (const uint32_t*)d0, n, seed,
                         rpb, (uint32_t)nparts, fm, counts);
    else if (w0 == 4)
      hipLaunchKernelGGL((k_hp_hist_v<uint32_t, 0>), dim3(nblk), dim3(kHpThreads), hl, s, (const uint32_t*)d0, n, seed,
                         rpb, (uint32_t)nparts, fm, counts);
    else if (fw0 == 8)
      hipLaunchKernelGGL((k_hp_hist_v<uint64_t, 8>), dim3(nblk), dim3(kHpThreads), hl, s, (const uint64_t*)d0, n, seed,
                         rpb, (uint32_t)nparts, fm, counts);
    else
      hipLaunchKernelGGL((k_hp_hist_v<uint64_t, 0>), dim3(nblk), dim3(kHpThreads), hl, s, (const uint64_t*)d0, n, seed,
                         rpb, (uint32_t)nparts, fm, counts);
    LDF_HIP_TRY(hipGetLastError());
  }
  exclusive_scan_u32_to_u64_ws(counts, scan, (int64_t)nparts * nblk, ws, s);

  ProfScope prof(profile ? "hp_scatter" : nullptr, s);
  if (single) {
    const ldf_column_view& c0 = table->cols[0];
    if (w0 == 1) launch_scatter_v<uint8_t, 0>(c0, n, seed, rpb, nblk, nparts, pbits, scan, dst[0], s);
    else if (w0 == 2) launch_scatter_v<uint16_t, 0>(c0, n, seed, rpb, nblk, nparts, pbits, scan, dst[0], s);
    else if (w0 == 4 && fw0 == 4) launch_scatter_v<uint32_t, 4>(c0, n, seed, rpb, nblk, nparts, pbits, scan, dst[0], s);
    else if (w0 == 4) launch_scatter_v<uint32_t, 0>(c0, n, seed, rpb, nblk, nparts, pbits, scan, dst[0], s);
    else if (fw0 == 8) launch_scatter_v<uint64_t, 8>(c0, n, seed, rpb, nblk, nparts, pbits, scan, dst[0], s);
    else launch_scatter_v<uint64_t, 0>(c0, n, seed, rpb, nblk, nparts, pbits, scan, dst[0], s);
  } else {
    ScatterCols sc;
    sc.n = table->ncols;
    for (int i = 0; i < table->ncols; ++i) {
      sc.src[i]   = table->cols[i].data;
      sc.dst[i]   = dst[i];
      sc.width[i] = type_width(table->cols[i].type_id);
    }
    launch_scatter(nblk, (uint32_t)nparts, s, hc, n, rpb, fm, pbits, scan, sc, nullptr);
  }
  hipLaunchKernelGGL(k_hp_pick_offsets, dim3(grid_for(nparts + 1, 256)), dim3(256), 0, s, (const uint64_t*)scan,
                     (uint32_t)nparts, (uint32_t)nblk, offs);
  LDF_HIP_TRY(hipGetLastError());
}

// ---- allocation-free variant: scratch in a fixed device array --------------------------------
// Each GPU the library runs on holds its own copy of `scratch` until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static unsigned char* scratch_for_current_device()
{
  int dev = 0;
  LDF_HIP_TRY(hipGetDevice(&dev));
  LDF_EXPECTS(dev >= 0 && dev < kMaxDevices, "hash partition (no alloc): device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr     = nullptr;
    hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || ptr == nullptr) {
      (void)hipGetLastError();
      throw Error(LDF_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    }
    g_scratch_ptr[dev] = ptr;
  }
  return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

}  // namespace ldf

using namespace ldf;

extern "C" int ldf_hash_partition(const ldf_table_view* table, const int32_t* cols, int32_t ncols, int32_t nparts,
                                  uint32_t seed, const ldf_allocator* alloc, void* stream, ldf_table_out* out,
                                  int64_t* offsets)
{
  return guarded([&] {
    LDF_EXPECTS(table != nullptr && out != nullptr && offsets != nullptr, "NULL argument");
    LDF_EXPECTS(nparts >= 1, "hash partition: nparts must be >= 1");
    if (nparts > kHpMaxParts)
      throw Error(LDF_E_UNSUPPORTED, "hash partition: at most 2048 partitions are supported");
    LDF_EXPECTS(out->ncols >= table->ncols, "hash partition: output table has too few column slots");
    hipStream_t s = (hipStream_t)stream;
    Alloc a(alloc, s);
    int64_t n = table->ncols > 0 ? table->cols[0].length : 0;
    for (int i = 0; i < table->ncols; ++i) validate_view(table->cols[i], n, "hash_partition");
    HashCols hc = make_hash_cols(table, cols, ncols, seed);
    LDF_EXPECTS(n < (int64_t(1) << 40), "hash partition: too many rows");

    int pbits = 0;
    while ((1u << pbits) < (uint32_t)nparts) ++pbits;

    // every column fixed width and non-nullable -> rows move in the scatter kernel itself
    bool direct = table->ncols <= kMaxKeyCols;
    for (int i = 0; i < table->ncols; ++i)
      if (type_width(table->cols[i].type_id) == 0 || table->cols[i].validity != nullptr) direct = false;

    out->ncols = table->ncols;
    if (n == 0) {
      for (int i = 0; i < table->ncols; ++i) {
        alloc_column(table->cols[i].type_id, 0, table->cols[i].validity != nullptr, 0, a, &out->cols[i]);
        if (type_width(table->cols[i].type_id) == 0) fill_u64((uint64_t*)out->cols[i].offsets, 0, 1, s);
      }
      for (int p = 0; p <= nparts; ++p) offsets[p] = 0;
      return;
    }

    int64_t nblk, rpb;
    partition_blocks(n, 1536, &nblk, &rpb);
    const int64_t ncounts = (int64_t)nparts * nblk;
    Buf<uint32_t> counts(a, (size_t)ncounts);
    Buf<uint64_t> scan(a, (size_t)ncounts + 1);
    Buf<uint64_t> ws(a, (size_t)scan_workspace_elems(ncounts));
    Buf<uint64_t> offs(a, (size_t)nparts + 1);
    if (direct) {
      void* dst[kMaxKeyCols];
      for (int i = 0; i < table->ncols; ++i) {
        alloc_column(table->cols[i].type_id, n, false, 0, a, &out->cols[i]);
        dst[i] = out->cols[i].data;
      }
      partition_fixed(table, hc, cols, ncols, n, nparts, seed, nblk, rpb, dst, counts.get(), scan.get(), ws.get(),
                      offs.get(), true, s);
    } else {
      LDF_EXPECTS(n < (int64_t(1) << 31), "hash partition: nullable or string columns limited to 2^31-1 rows");
      const FastMod fm = fastmod_make((uint32_t)nparts);
      {
        ProfScope prof("hp_hist", s);
        hipLaunchKernelGGL(k_hp_hist, dim3(nblk), dim3(kHpThreads), nparts * sizeof(uint32_t), s, hc, n, rpb,
                           (uint32_t)nparts, fm, counts.get());
        LDF_HIP_TRY(hipGetLastError());
      }
      exclusive_scan_u32_to_u64_ws(counts.get(), scan.get(), ncounts, ws.get(), s);
      Buf<int64_t> perm(a, n);
      ScatterCols sc;
      sc.n = 0;
      {
        ProfScope prof("hp_scatter", s);
        launch_scatter(nblk, (uint32_t)nparts, s, hc, n, rpb, fm, pbits, scan.get(), sc, perm.get());
      }
      Buf<int32_t> map(a, n);
      hipLaunchKernelGGL(k_i64_to_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, perm.get(), n, map.get());
      LDF_HIP_TRY(hipGetLastError());
      std::vector<const ldf_column_view*> src(table->ncols);
      for (int i = 0; i < table->ncols; ++i) src[i] = &table->cols[i];
      gather_columns(src.data(), table->ncols, map.get(), n, false, a, out->cols, /*random_map=*/true);
      hipLaunchKernelGGL(k_hp_pick_offsets, dim3(grid_for(nparts + 1, 256)), dim3(256), 0, s,
                         (const uint64_t*)scan.get(), (uint32_t)nparts, (uint32_t)nblk, offs.get());
      LDF_HIP_TRY(hipGetLastError());
    }
    std::vector<uint64_t> h(nparts + 1);
    copy_to_host(h.data(), offs.get(), (nparts + 1) * sizeof(uint64_t), s);
    for (int p = 0; p <= nparts; ++p) offsets[p] = (int64_t)h[p];
  });
}

extern "C" int ldf_hash_partition_noalloc(const ldf_table_view* table, const int32_t* cols, int32_t ncols,
                                          int32_t nparts, uint32_t seed, void* stream, ldf_table_out* out,
                                          int64_t* offsets)
{
  return guarded([&] {
    LDF_EXPECTS(table != nullptr && out != nullptr && offsets != nullptr && out->cols != nullptr, "NULL argument");
    LDF_EXPECTS(nparts >= 1, "hash partition (no alloc): nparts must be >= 1");
    if (nparts > kHpMaxParts)
      throw Error(LDF_E_UNSUPPORTED, "hash partition (no alloc): at most 2048 partitions are supported");
    LDF_EXPECTS(table->ncols >= 1 && table->ncols <= kMaxKeyCols,
                "hash partition (no alloc): between 1 and 8 columns");
    LDF_EXPECTS(out->ncols >= table->ncols, "hash partition (no alloc): output table has too few column slots");
    hipStream_t s   = (hipStream_t)stream;
    const int64_t n = table->cols[0].length;
    for (int i = 0; i < table->ncols; ++i) {
      validate_view(table->cols[i], n, "hash_partition_noalloc");
      LDF_EXPECTS(type_width(table->cols[i].type_id) > 0 && table->cols[i].validity == nullptr,
                  "hash partition (no alloc): every column must be fixed width and non-nullable");
      LDF_EXPECTS(out->cols[i].data != nullptr || n == 0,
                  "hash partition (no alloc): out->cols[i].data must hold the partitioned column");
      LDF_EXPECTS(out->cols[i].data_bytes >= n * type_width(table->cols[i].type_id),
                  "hash partition (no alloc): out->cols[i].data_bytes is smaller than the column");
    }
    HashCols hc = make_hash_cols(table, cols, ncols, seed);
    LDF_EXPECTS(n < (int64_t(1) << 40), "hash partition (no alloc): too many rows");
    unsigned char* base = scratch_for_current_device();  // once per device: the array's device address
    for (int i = 0; i < table->ncols; ++i) {
      out->cols[i].type_id    = table->cols[i].type_id;
      out->cols[i].length     = n;
      out->cols[i].validity   = nullptr;
      out->cols[i].offsets    = nullptr;
      out->cols[i].null_count = 0;
      out->cols[i].validity_bytes = 0;
    }
    out->ncols = table->ncols;
    if (n == 0) {
      for (int p = 0; p <= nparts; ++p) offsets[p] = 0;
      return;
    }
    // the scratch holds counts u32[nparts * nblk], scan u64[nparts * nblk + 1], the scan's
    // workspace and offs u64[nparts + 1]: fewer blocks for more partitions
    int64_t max_blocks = 1536;
    auto need = [&](int64_t nb) {
      const int64_t c = (int64_t)nparts * nb;
      return (size_t)(((c * 4 + 255) / 256) * 256 + ((c + 1) * 8 + 255) / 256 * 256 +
                      (scan_workspace_elems(c) * 8 + 255) / 256 * 256 + (nparts + 1) * 8);
    };
    while (max_blocks > 1 && need(max_blocks) > kNoAllocScratchBytes) max_blocks /= 2;
    int64_t nblk, rpb;
    partition_blocks(n, max_blocks, &nblk, &rpb);
    const int64_t c = (int64_t)nparts * nblk;
    if (need(nblk) > kNoAllocScratchBytes)
      throw Error(LDF_E_INVALID, "hash partition (no alloc): the partition counts exceed the " +
                                   std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch");
    uint32_t* counts = reinterpret_cast<uint32_t*>(base);
    uint64_t* scan   = reinterpret_cast<uint64_t*>(base + (c * 4 + 255) / 256 * 256);
    uint64_t* ws     = reinterpret_cast<uint64_t*>(reinterpret_cast<unsigned char*>(scan) + ((c + 1) * 8 + 255) / 256 * 256);
    uint64_t* offs   = reinterpret_cast<uint64_t*>(reinterpret_cast<unsigned char*>(ws) +
                                                 (scan_workspace_elems(c) * 8 + 255) / 256 * 256);
    void* dst[kMaxKeyCols];
    for (int i = 0; i < table->ncols; ++i) dst[i] = out->cols[i].data;
    // every scratch word the kernels read is written first in this call (the histogram stores
    // every count, the scan every prefix), so nothing needs clearing between calls
    partition_fixed(table, hc, cols, ncols, n, nparts, seed, nblk, rpb, dst, counts, scan, ws, offs, false, s);
    static_assert(sizeof(int64_t) == sizeof(uint64_t), "offsets are copied as 64-bit words");
    LDF_HIP_TRY(hipMemcpyAsync(offsets, offs, (size_t)(nparts + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    LDF_HIP_TRY(hipStreamSynchronize(s));
  });
}
