// Batched form: many independent joins or groupbys of one schema executed together.
//
// The reference runs one Legate point task per partition -- JoinTask::gpu_variant
// (cpp/src/join.cu:187-244) or GroupByAggregationTask::gpu_variant
// (cpp/src/groupby_aggregation.cu:90-167) on each partition's PhysicalTable -- so thousands of
// small, uneven partitions each pay a launch sequence, host synchronisations and allocations.
// Here a batch is planned once (sizes, hash tables, work lists and outputs, with one host
// synchronisation for the output sizes) and every run enqueues the whole computation with no
// host synchronisation and no allocation, so the run can be captured in a hipGraph and replayed.
//
// Work queue: every phase that walks instances is one persistent kernel launched with a
// resident grid; its blocks pop (instance, tile) items from a device cursor (atomicAdd) until
// the list is exhausted, so uneven instances balance dynamically, and every block leaves once
// the cursor passes the end.  Items are listed largest instance first.  Phases are separated by
// kernel boundaries: no block waits for another inside a kernel.
//
// Join batch (INNER / LEFT / SEMI / ANTI, one fixed-width key): the same algorithm as ldf_join
// -- a CSR multimap per instance (16-byte slots {key, cnt, start}, 64-bit CAS insert, build rows
// grouped by slot), a count pass, a device-wide scan of per-tile output counts, an emit pass
// expanding each tile over the block, and a gather of all output columns -- over one slot array
// that holds every instance's table in its own region.
//
// Groupby batch (SUM, MEAN, COUNT, COUNT_ALL, MIN, MAX; one fixed-width key): per instance the
// rows are scattered into P_i hash partitions (P_i from a sampled distinct-count estimate, so a
// partition holds a few hundred keys), then one block aggregates each partition in an LDS hash
// table -- all accumulation in LDS, exact FP64 sums as 192-bit fixed point, no global atomics
// on accumulators -- and writes its finished groups to a staging area; a final pass gathers the
// staged groups of all partitions into the concatenated output.  A partition whose keys overflow
// the LDS table is aggregated again in 2, 4, ... hash sub-passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ldf_agg.hpp"
#include "ldf_common.hpp"
#include "ldf_device.hpp"

namespace ldf {
namespace {

constexpr unsigned long long kBE = ~0ull;
constexpr int kQT                = 256;  // threads of the join work-queue kernels
constexpr int kGT                = 512;  // threads of the groupby work-queue kernels
constexpr int kMaxBOut           = 16;   // join output columns
constexpr int kMaxVal            = 4;    // groupby value columns
constexpr int kMaxBAcc           = 16;
constexpr int kMaxBAgg           = 16;
constexpr int kMaxParts          = 4096;  // partitions per groupby instance
constexpr int kStageFactor       = 1;     // staged groups per partition = kStageFactor * LDS slots
constexpr int kMaxSubPass        = 8;

struct Item {
  int32_t inst;
  int32_t tile;
};

// Block-wide pop from a work queue; every block leaves once the cursor passes nitems.
__device__ __forceinline__ int pop_item(unsigned int* cursor, int* sh)
{
  __syncthreads();
  if (threadIdx.x == 0) *sh = (int)atomicAdd(cursor, 1u);
  __syncthreads();
  return *sh;
}

// =========================================================================================
// join batch kernels
// =========================================================================================
struct BSlot {
  unsigned long long key;
  uint32_t cnt;
  uint32_t start;
};

struct JInst {
  const void* pkey;
  const uint8_t* pvalid;
  int64_t np;
  const void* bkey;
  const uint8_t* bvalid;
  int64_t nb;
  int64_t tab;  // first slot of this instance's table (cap + 2 slots)
  uint64_t mask;
  uint64_t cap;
  int64_t brow;   // first build row in the concatenated build arrays
  int64_t prow;   // first probe row in the concatenated probe arrays
  int64_t ptile;  // first probe tile in the concatenated tile arrays
  int64_t utile;  // FULL: first of the tiles that hold its unmatched build rows (one per build tile)
  BSlot* tabp;            // the instance's table (cap + 2 slots)
  const uint32_t* rowsp;  // build rows grouped by slot; entries are brow + local row
};

__global__ void k_bj_fill(BSlot* __restrict__ t, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    t[i].key   = kBE;
    t[i].cnt   = 0;
    t[i].start = 0;
  }
}

__device__ __forceinline__ int64_t bj_insert(const JInst& I, int w, int64_t r, BSlot* tab, bool ne)
{
  if (!valid_at(I.bvalid, r)) return ne ? I.tab + (int64_t)I.cap + 1 : -1;
  const unsigned long long k = load_bits(I.bkey, w, r);
  if (k == kBE) return I.tab + (int64_t)I.cap;
  BSlot* t   = tab + I.tab;
  uint64_t i = mix64(k) & I.mask;
  while (true) {  // the table has more slots than build rows: an empty slot exists
    unsigned long long cur = atomic_load_relaxed(&t[i].key);
    if (cur == k) return I.tab + (int64_t)i;
    if (cur == kBE) {
      unsigned long long old = atomicCAS(&t[i].key, kBE, k);
      if (old == kBE || old == k) return I.tab + (int64_t)i;
    }
    i = (i + 1) & I.mask;
  }
}

// slot of the probe row's key in the instance's table `t` (its HBM table, or a copy in LDS;
// local index), -1 = no build row has it
__device__ __forceinline__ int64_t bj_find(const JInst& I, const BSlot* t, int w, int64_t r, bool ne)
{
  if (!valid_at(I.pvalid, r)) return ne ? (int64_t)I.cap + 1 : -1;
  const unsigned long long k = load_bits(I.pkey, w, r);
  if (k == kBE) return (int64_t)I.cap;
  uint64_t i     = mix64(k) & I.mask;
  while (true) {
    unsigned long long cur = t[i].key;
    if (cur == k) return (int64_t)i;
    if (cur == kBE) return -1;
    i = (i + 1) & I.mask;
  }
}

__device__ __forceinline__ uint64_t bj_out_rows(uint32_t cnt, int jt)
{
  switch (jt) {
    case LDF_JOIN_INNER: return cnt;
    case LDF_JOIN_LEFT:
    case LDF_JOIN_FULL: return cnt > 0 ? cnt : 1;
    case LDF_JOIN_SEMI: return cnt > 0 ? 1 : 0;
    default: return cnt > 0 ? 0 : 1;  // ANTI
  }
}

__global__ __launch_bounds__(kQT) void k_bj_build(const JInst* __restrict__ inst, const Item* __restrict__ items,
                                                  int nitems, unsigned int* cursor, int tile, int w, bool ne,
                                                  BSlot* __restrict__ tab, int64_t* __restrict__ slot_of_row)
{
  __shared__ int sh;
  for (;;) {
    const int it = pop_item(cursor, &sh);
    if (it >= nitems) break;
    const Item m     = items[it];
    const JInst I    = inst[m.inst];
    const int64_t r0 = (int64_t)m.tile * tile;
    const int64_t r1 = r0 + tile < I.nb ? r0 + tile : I.nb;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kQT) {
      const int64_t s = bj_insert(I, w, r, tab, ne);
      if (s >= 0) atomicAdd(&tab[s].cnt, 1u);
      slot_of_row[I.brow + r] = s;
    }
  }
}

__global__ void k_bj_counts(const BSlot* __restrict__ tab, int64_t n, uint32_t* __restrict__ cnt)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    cnt[i] = tab[i].cnt;
}

__global__ void k_bj_starts(BSlot* __restrict__ tab, int64_t n, const uint64_t* __restrict__ start)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    tab[i].start = (uint32_t)start[i];
}

// rows[] holds concatenated build row numbers grouped by slot
__global__ void k_bj_group(const int64_t* __restrict__ slot_of_row, int64_t nbt, const BSlot* __restrict__ tab,
                           uint32_t* __restrict__ fill, uint32_t* __restrict__ rows)
{
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nbt; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = slot_of_row[g];
    if (s < 0) continue;
    rows[tab[s].start + atomicAdd(&fill[s], 1u)] = (uint32_t)g;
  }
}

// info[prow + r] = start | cnt << 32; btot[ptile + tile] = output rows of the tile
// FULL: hit[slot] = 1 for every slot a probe row matches (plain stores of the same value).
// With lds_max > 0 an instance whose table has at most lds_max slots is probed in an LDS copy of
// it (the block loads the table once per work item: placement chosen by the autotuner).
__global__ __launch_bounds__(kQT) void k_bj_count(const JInst* __restrict__ inst, const Item* __restrict__ items,
                                                  int nitems, unsigned int* cursor, int tile, int w, bool ne, int jt,
                                                  uint64_t* __restrict__ info, uint64_t* __restrict__ btot,
                                                  uint8_t* __restrict__ hit, int lds_max)
{
  extern __shared__ __align__(16) unsigned char dyn[];
  BSlot* ltab = reinterpret_cast<BSlot*>(dyn);
  __shared__ int sh;
  __shared__ uint64_t smem[17];
  for (;;) {
    const int it = pop_item(cursor, &sh);
    if (it >= nitems) break;
    const Item m     = items[it];
    const JInst I    = inst[m.inst];
    const int64_t r0 = (int64_t)m.tile * tile;
    const int64_t r1 = r0 + tile < I.np ? r0 + tile : I.np;
    const bool in_lds = lds_max > 0 && I.cap <= (uint64_t)lds_max;  // the same for the whole block
    if (in_lds) {
      for (int64_t i = threadIdx.x; i < (int64_t)I.cap + 2; i += kQT) ltab[i] = I.tabp[i];
      __syncthreads();
    }
    const BSlot* T = in_lds ? ltab : I.tabp;
    uint64_t sum   = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kQT) {
      const int64_t s = bj_find(I, T, w, r, ne);
      uint32_t cnt = 0, st = 0;
      if (s >= 0) {
        cnt = T[s].cnt;
        st  = T[s].start;
        if (hit != nullptr && cnt > 0 && hit[I.tab + s] == 0) hit[I.tab + s] = 1;
      }
      info[I.prow + r] = (uint64_t)st | ((uint64_t)cnt << 32);
      sum += bj_out_rows(cnt, jt);
    }
    const uint64_t t = block_reduce_sum(sum, smem);
    if (threadIdx.x == 0) btot[I.ptile + m.tile] = t;
  }
}

// Expands each probe tile's output rows over the whole block (binary search over the tile's
// row offsets in LDS), writing instance-local row numbers (-1 = no rhs row).
__global__ __launch_bounds__(kQT) void k_bj_emit(const JInst* __restrict__ inst, const Item* __restrict__ items,
                                                 int nitems, unsigned int* cursor, int tile, int jt,
                                                 const uint64_t* __restrict__ info,
                                                 const uint64_t* __restrict__ boff, int64_t out_cap,
                                                 int32_t* __restrict__ lmap, int32_t* __restrict__ rmap,
                                                 int32_t* __restrict__ flags)
{
  extern __shared__ __align__(16) unsigned char dyn[];
  uint64_t* off = reinterpret_cast<uint64_t*>(dyn);  // tile + 1
  uint32_t* st  = reinterpret_cast<uint32_t*>(off + tile + 1);
  uint32_t* mc  = st + tile;
  __shared__ int sh;
  __shared__ uint64_t smem[17];
  const int per = tile / kQT;
  for (;;) {
    const int it = pop_item(cursor, &sh);
    if (it >= nitems) break;
    const Item m       = items[it];
    const JInst I      = inst[m.inst];
    const int64_t base = (int64_t)m.tile * tile;
    const int nrows    = (int)(I.np - base < tile ? I.np - base : tile);
    uint64_t s         = 0;
    for (int j = 0; j < per; ++j) {
      const int li = threadIdx.x * per + j;
      if (li < nrows) {
        const uint64_t x = info[I.prow + base + li];
        const uint32_t c = (uint32_t)(x >> 32);
        st[li]           = (uint32_t)x;
        mc[li]           = c;
        s += bj_out_rows(c, jt);
      }
    }
    uint64_t total;
    uint64_t run = block_exclusive_scan(s, smem, &total);
    for (int j = 0; j < per; ++j) {
      const int li = threadIdx.x * per + j;
      if (li < nrows) {
        off[li] = run;
        run += bj_out_rows(mc[li], jt);
      }
    }
    if (threadIdx.x == 0) off[nrows] = total;
    __syncthreads();
    const uint64_t gbase = boff[I.ptile + m.tile];
    if (gbase + total > (uint64_t)out_cap) {
      if (threadIdx.x == 0) flags[0] = 1;  // more output than planned: the inputs changed
      continue;
    }
    for (uint64_t q = threadIdx.x; q < total; q += kQT) {
      int lo = 0, hi = nrows - 1;  // last row with off <= q
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= q) lo = mid;
        else hi = mid - 1;
      }
      const uint64_t gq = gbase + q;
      lmap[gq]          = (int32_t)(base + lo);
      if (rmap != nullptr)
        rmap[gq] = mc[lo] == 0 ? -1 : (int32_t)((int64_t)I.rowsp[st[lo] + (q - off[lo])] - I.brow);
    }
  }
}

// FULL: the build rows of each (instance, build tile) whose slot no probe row matched --
// btot[utile + tile] = their number (count pass); the emit pass writes them, in build-row order,
// as output rows with no lhs row (lmap = -1) after the instance's probe rows.
__global__ __launch_bounds__(kQT) void k_bj_unmatched(const JInst* __restrict__ inst, const Item* __restrict__ items,
                                                      int nitems, unsigned int* cursor, int btile,
                                                      const int64_t* __restrict__ slot_row,
                                                      const uint8_t* __restrict__ hit, uint64_t* __restrict__ btot,
                                                      const uint64_t* __restrict__ boff, int64_t out_cap,
                                                      int32_t* __restrict__ lmap, int32_t* __restrict__ rmap,
                                                      int32_t* __restrict__ flags)
{
  __shared__ int sh;
  __shared__ uint64_t smem[17];
  const int per = btile / kQT;
  for (;;) {
    const int it = pop_item(cursor, &sh);
    if (it >= nitems) break;
    const Item m     = items[it];
    const JInst I    = inst[m.inst];
    const int64_t r0 = (int64_t)m.tile * btile + (int64_t)threadIdx.x * per;  // this thread's rows, in order
    uint32_t c       = 0;
    for (int j = 0; j < per; ++j) {
      const int64_t r = r0 + j;
      if (r < I.nb && hit[slot_row[I.brow + r]] == 0) ++c;
    }
    uint64_t total;
    uint64_t pos = block_exclusive_scan(c, smem, &total);
    if (boff == nullptr) {  // count pass
      if (threadIdx.x == 0) btot[I.utile + m.tile] = total;
      continue;
    }
    const uint64_t gbase = boff[I.utile + m.tile];
    if (gbase + total > (uint64_t)out_cap) {
      if (threadIdx.x == 0) flags[0] = 1;
      continue;
    }
    for (int j = 0; j < per; ++j) {
      const int64_t r = r0 + j;
      if (r < I.nb && hit[slot_row[I.brow + r]] == 0) {
        lmap[gbase + pos] = -1;
        rmap[gbase + pos] = (int32_t)r;
        ++pos;
      }
    }
  }
}

// ooff[i] = first output row of instance i; ooff[ninst] = total output rows
__global__ void k_bj_ooff(const JInst* __restrict__ inst, int ninst, const uint64_t* __restrict__ boff, int64_t ntiles,
                          int64_t* __restrict__ ooff)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= ninst; i += gridDim.x * blockDim.x)
    ooff[i] = (int64_t)(i < ninst ? boff[inst[i].ptile] : boff[ntiles]);
}

struct BGather {
  int ncols;
  int width[kMaxBOut];
  int side[kMaxBOut];  // 0: lhs row (lmap), 1: rhs row (rmap)
  void* dst[kMaxBOut];
  uint64_t* dval[kMaxBOut];
};

template <typename T>
__device__ __forceinline__ void bgather_store(const void* src, void* dst, int32_t m, int64_t i)
{
  static_cast<T*>(dst)[i] = m >= 0 ? static_cast<const T*>(src)[m] : T(0);
}

// Output row i belongs to the instance whose [ooff[k], ooff[k+1]) holds it; sources are the
// instance's own columns.  One wave covers 64 consecutive output rows (one validity word).
__global__ __launch_bounds__(256) void k_bj_gather(BGather g, const void* const* __restrict__ src,
                                                   const uint8_t* const* __restrict__ sval, int ninst,
                                                   const int64_t* __restrict__ ooff, int64_t cap,
                                                   const int32_t* __restrict__ lmap, const int32_t* __restrict__ rmap,
                                                   unsigned long long* __restrict__ nulls)
{
  const int lane   = threadIdx.x & 63;
  const int64_t n0 = ooff[ninst];
  const int64_t n  = n0 < cap ? n0 : cap;
  uint64_t nl[kMaxBOut];
#pragma unroll
  for (int c = 0; c < kMaxBOut; ++c) nl[c] = 0;
  for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + lane;
    int k           = 0;
    if (lane == 0) {  // last instance whose first row is <= base
      int lo = 0, hi = ninst - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ooff[mid] <= base) lo = mid;
        else hi = mid - 1;
      }
      k = lo;
    }
    k = __shfl(k, 0);
    if (i < n)
      while (ooff[k + 1] <= i) ++k;
    const int32_t lm = i < n ? lmap[i] : -1;
    const int32_t rm = (i < n && rmap != nullptr) ? rmap[i] : -1;
#pragma unroll
    for (int c = 0; c < kMaxBOut; ++c) {
      if (c >= g.ncols) break;
      const int32_t mr = g.side[c] ? rm : lm;
      const int idx    = k * g.ncols + c;
      const bool ok    = i < n && mr >= 0 && valid_at(sval[idx], mr);
      if (i < n) {
        switch (g.width[c]) {
          case 1: bgather_store<uint8_t>(src[idx], g.dst[c], mr, i); break;
          case 2: bgather_store<uint16_t>(src[idx], g.dst[c], mr, i); break;
          case 4: bgather_store<uint32_t>(src[idx], g.dst[c], mr, i); break;
          default: bgather_store<uint64_t>(src[idx], g.dst[c], mr, i); break;
        }
      }
      if (g.dval[c] != nullptr) {
        const uint64_t b = __ballot(ok);
        if (lane == 0) {
          g.dval[c][base >> 6]   = b;
          const int64_t in_range = n - base < 64 ? n - base : 64;
          nl[c] += (uint64_t)in_range - (uint64_t)__popcll(b);
        }
      }
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < kMaxBOut; ++c)
      if (c < g.ncols && nl[c] != 0) atomicAdd(&nulls[c], (unsigned long long)nl[c]);
  }
}

// =========================================================================================
// groupby batch kernels
// =========================================================================================
enum BAcc : int32_t { B_ROWS = 0, B_VCOUNT = 1, B_ISUM = 2, B_FSUM = 3, B_MIN = 4, B_MAX = 5 };

struct GInst {
  const void* key;
  const uint8_t* kvalid;
  int64_t n;
  const void* val[kMaxVal];
  const uint8_t* vvalid[kMaxVal];
  int64_t pbase;  // first partition of this instance
  int32_t pbits;  // log2 of the instance's partition count
};

struct BAccDev {
  int32_t kind;
  int32_t col;
  int32_t off;  // byte offset of this accumulator's array in the LDS table
};

struct BOutDev {
  int32_t kind;      // LDF_AGG_*; -1 = key
  int32_t mode;      // CountOptions mode
  int32_t in_type;   // value column type
  int32_t out_type;  // output type
  int32_t width;     // output width
  int32_t main;      // accumulator index (SUM / MEAN / MIN / MAX)
  int32_t vcount;    // accumulator index of the valid count, -1 = the row count
  int32_t nullable;  // output has a validity bitmap
  void* stage;       // staged values (one per staged group)
  uint8_t* stage_valid;
  void* out;
  uint64_t* out_valid;
};

struct GPlanB {
  int nacc;
  BAccDev a[kMaxBAcc];
  int nout;
  BOutDev o[kMaxBAgg + 1];
  int ncol;
  int ctype[kMaxVal];
  int cwidth[kMaxVal];
  int ktype;
  int kwidth;
  int lslots;     // LDS table slots (power of two); two more hold the special keys
  int lds_bytes;  // LDS table bytes
  int rows_off;   // byte offset of the u32 row counters
  int stage_cap;  // staged groups per partition
  // partition buffers (SoA), one entry per input row
  unsigned long long* pkey;
  unsigned long long* pval[kMaxVal];
  uint8_t* pflag;  // bit 0: key valid; bit 1 + c: value c valid
  uint64_t* pstart;    // partition start rows (scan of the partition counts)
  unsigned long long* pcursor;
  uint32_t* pcount;
  uint32_t* gcount;    // groups per partition
  int32_t* flags;      // [0] FP64 sum range error, [1] staging overflow
};

__device__ __forceinline__ uint32_t part_of(uint64_t kb, bool kv, int pbits)
{
  if (!kv || pbits == 0) return 0;
  return (uint32_t)(mix64(kb) >> (64 - pbits));
}

// Rows a thread of the groupby scatter kernels loads before using any.
constexpr int kBgRows = 4;

// The partition counts of rows [r0, r1) of one instance into `hist` (branch-free loads, all of a
// thread's rows in flight together; r1 > r0).
__device__ __forceinline__ void bg_tile_hist(const GInst& I, const GPlanB& p, int64_t r0, int64_t r1, uint32_t* hist)
{
  for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)kBgRows * kGT) {
    uint64_t kb[kBgRows];
    uint32_t kv[kBgRows];
#pragma unroll
    for (int u = 0; u < kBgRows; ++u) {
      const int64_t rr = r + (int64_t)u * kGT;
      const int64_t rc = rr < r1 ? rr : r1 - 1;
      kb[u]            = load_bits_nb(I.key, p.kwidth, rc);
      kv[u]            = valid_bit_nb(I.kvalid, I.key, rc);
    }
#pragma unroll
    for (int u = 0; u < kBgRows; ++u)
      if (r + (int64_t)u * kGT < r1) atomicAdd(&hist[part_of(kb[u], kv[u] != 0, I.pbits)], 1u);
  }
}

// LDS hash partitions per tile: counts (first pass) or reserved positions (second pass)
__global__ __launch_bounds__(kGT) void k_bg_hist(const GInst* __restrict__ inst, const Item* __restrict__ items,
                                                 int nitems, unsigned int* cursor, int tile, GPlanB p)
{
  __shared__ int sh;
  __shared__ uint32_t hist[kMaxParts];
  for (;;) {
    const int it = pop_item(cursor, &sh);
    if (it >= nitems) break;
    const Item m     = items[it];
    const GInst I    = inst[m.inst];
    const int np     = 1 << I.pbits;
    for (int q = threadIdx.x; q < np; q += kGT) hist[q] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)m.tile * tile;
    const int64_t r1 = r0 + tile < I.n ? r0 + tile : I.n;
    bg_tile_hist(I, p, r0, r1, hist);
    __syncthreads();
    for (int q = threadIdx.x; q < np; q += kGT)
      if (hist[q]) atomicAdd(&p.pcount[I.pbase + q], hist[q]);
  }
}

__global__ __launch_bounds__(kGT) void k_bg_scatter(const GInst* __restrict__ inst, const Item* __restrict__ items,
                                                    int nitems, unsigned int* cursor, int tile, GPlanB p)
{
  __shared__ int sh;
  __shared__ uint32_t hist[kMaxParts];
  __shared__ unsigned long long cur[kMaxParts];
  for (;;) {
    const int it = pop_item(cursor, &sh);
    if (it >= nitems) break;
    const Item m  = items[it];
    const GInst I = inst[m.inst];
    const int np  = 1 << I.pbits;
    for (int q = threadIdx.x; q < np; q += kGT) hist[q] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)m.tile * tile;
    const int64_t r1 = r0 + tile < I.n ? r0 + tile : I.n;
    bg_tile_hist(I, p, r0, r1, hist);
    __syncthreads();
    for (int q = threadIdx.x; q < np; q += kGT)
      if (hist[q]) cur[q] = atomicAdd(&p.pcursor[I.pbase + q], (unsigned long long)hist[q]);
    __syncthreads();
    for (int64_t r = r0 + threadIdx.x; r < r1; r += (int64_t)kBgRows * kGT) {
      uint64_t kb[kBgRows], x[kBgRows][kMaxVal];
      uint32_t fl[kBgRows];
#pragma unroll
      for (int u = 0; u < kBgRows; ++u) {  // every load of the thread's rows in flight together
        const int64_t rr = r + (int64_t)u * kGT;
        const int64_t rc = rr < r1 ? rr : r1 - 1;
        kb[u]            = load_bits_nb(I.key, p.kwidth, rc);
        fl[u]            = valid_bit_nb(I.kvalid, I.key, rc);
#pragma unroll
        for (int c = 0; c < kMaxVal; ++c) {  // columns past p.ncol read the key column again
          const bool use = c < p.ncol;
          const void* d  = use ? I.val[c] : I.key;
          x[u][c]        = load_bits_nb(d, use ? p.cwidth[c] : p.kwidth, rc);
          fl[u] |= (use ? valid_bit_nb(I.vvalid[c], d, rc) : 0u) << (c + 1);
        }
      }
#pragma unroll
      for (int u = 0; u < kBgRows; ++u) {
        if (r + (int64_t)u * kGT >= r1) break;
        const uint32_t q = part_of(kb[u], (fl[u] & 1u) != 0, I.pbits);
        const uint64_t d = atomicAdd(&cur[q], 1ull);
        p.pkey[d]        = kb[u];
#pragma unroll
        for (int c = 0; c < kMaxVal; ++c)
          if (c < p.ncol) p.pval[c][d] = x[u][c];
        p.pflag[d] = (uint8_t)fl[u];
      }
    }
  }
}

__global__ void k_bg_cursor_init(const uint64_t* __restrict__ pstart, int64_t n, unsigned long long* __restrict__ cursor)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    cursor[i] = pstart[i];
}

// LDS table of the aggregation kernel: keys[lslots + 2] (the last two hold the key whose bits
// equal the empty marker and the null key), then u32 row counters, then the accumulators
// (u32 valid counts, u64 integer sums, 6 x u64 exact-sum words, u64 ordered min / max).
__device__ __forceinline__ int bg_lds_slot(unsigned long long* keys, int lslots, uint64_t kb, bool kv, int* nkeys,
                                           int limit)
{
  if (!kv) return lslots + 1;
  if (kb == kBE) return lslots;
  const int mask = lslots - 1;
  int h          = (int)(mix64(kb) & (uint64_t)mask);
  for (int pr = 0; pr < lslots; ++pr) {
    unsigned long long cur = keys[h];
    if (cur == kb) return h;
    if (cur == kBE) {
      unsigned long long old = atomicCAS(&keys[h], kBE, (unsigned long long)kb);
      if (old == kBE) {
        if (atomicAdd(nkeys, 1) >= limit) return -2;  // table past its fill limit
        return h;
      }
      if (old == kb) return h;
    }
    h = (h + 1) & mask;
  }
  return -2;
}

__device__ __forceinline__ uint64_t pick4(uint64_t v0, uint64_t v1, uint64_t v2, uint64_t v3, int c)
{
  uint64_t x = v0;
  x          = c == 1 ? v1 : x;
  x          = c == 2 ? v2 : x;
  x          = c == 3 ? v3 : x;
  return x;
}

__device__ __forceinline__ void bg_store(void* out, int width, int64_t g, uint64_t bits)
{
  switch (width) {
    case 1: static_cast<uint8_t*>(out)[g] = (uint8_t)bits; break;
    case 2: static_cast<uint16_t*>(out)[g] = (uint16_t)bits; break;
    case 4: static_cast<uint32_t*>(out)[g] = (uint32_t)bits; break;
    default: static_cast<uint64_t*>(out)[g] = bits; break;
  }
}

// One block per (instance, partition): aggregate the partition's rows in LDS (in 1, 2, 4, ...
// sub-passes by key hash if its keys overflow the table) and stage the finished groups.
__global__ __launch_bounds__(kGT) void k_bg_aggregate(const GInst* __restrict__ inst, const Item* __restrict__ items,
                                                      int nitems, unsigned int* cursor, GPlanB p)
{
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int sh;
  __shared__ int nkeys;
  __shared__ int overflow;
  __shared__ int ngroups;
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(lds);
  uint32_t* rowsc          = reinterpret_cast<uint32_t*>(lds + p.rows_off);
  const int S              = p.lslots + 2;
  const int limit          = (p.lslots / 4) * 3;
  for (;;) {
    const int it = pop_item(cursor, &sh);
    if (it >= nitems) break;
    const Item m       = items[it];
    const GInst I      = inst[m.inst];
    const int64_t q    = I.pbase + m.tile;  // partition
    const int64_t r0   = (int64_t)p.pstart[q];
    const int64_t r1   = (int64_t)p.pstart[q + 1];
    const int64_t sbase = q * (int64_t)p.stage_cap;
    if (threadIdx.x == 0) ngroups = 0;
    int passes = 1;
    for (int pass = 0; pass < passes;) {
      // clear the table
      for (int h = threadIdx.x; h < S; h += kGT) {
        keys[h]  = kBE;
        rowsc[h] = 0;
      }
      for (int i = 0; i < p.nacc; ++i) {
        const BAccDev a = p.a[i];
        if (a.kind == B_ROWS) continue;
        if (a.kind == B_VCOUNT) {
          uint32_t* w = reinterpret_cast<uint32_t*>(lds + a.off);
          for (int h = threadIdx.x; h < S; h += kGT) w[h] = 0;
        } else {
          const int words = a.kind == B_FSUM ? 6 : 1;
          const unsigned long long v0 = a.kind == B_MIN ? ~0ull : 0ull;
          unsigned long long* w = reinterpret_cast<unsigned long long*>(lds + a.off);
          for (int h = threadIdx.x; h < S * words; h += kGT) w[h] = v0;
        }
      }
      if (threadIdx.x == 0) {
        nkeys    = 0;
        overflow = 0;
      }
      __syncthreads();
      for (int64_t r = r0 + threadIdx.x; r < r1; r += kGT) {
        const uint64_t kb = p.pkey[r];
        const uint32_t fl = p.pflag[r];
        const bool kv     = fl & 1u;
        if (passes > 1 && (kv ? (int)((mix64(kb) >> 20) % (uint64_t)passes) : 0) != pass) continue;
        const int h = bg_lds_slot(keys, p.lslots, kb, kv, &nkeys, limit);
        if (h < 0) {
          overflow = 1;
          break;
        }
        atomicAdd(&rowsc[h], 1u);
        const uint64_t v0 = p.ncol > 0 ? p.pval[0][r] : 0;
        const uint64_t v1 = p.ncol > 1 ? p.pval[1][r] : 0;
        const uint64_t v2 = p.ncol > 2 ? p.pval[2][r] : 0;
        const uint64_t v3 = p.ncol > 3 ? p.pval[3][r] : 0;
        for (int i = 0; i < p.nacc; ++i) {
          const BAccDev a = p.a[i];
          if (a.kind == B_ROWS || !((fl >> (1 + a.col)) & 1u)) continue;
          const uint64_t vb = pick4(v0, v1, v2, v3, a.col);
          const int ct = p.ctype[a.col], cw = p.cwidth[a.col];
          switch (a.kind) {
            case B_VCOUNT: atomicAdd(reinterpret_cast<uint32_t*>(lds + a.off) + h, 1u); break;
            case B_ISUM:
              atomicAdd(reinterpret_cast<unsigned long long*>(lds + a.off) + h,
                        (unsigned long long)bits_as_int(vb, ct, cw));
              break;
            case B_FSUM: {
              if (!lds_fsum_add(reinterpret_cast<unsigned long long*>(lds + a.off) + 6 * h, bits_as_double(vb, ct, cw)))
                p.flags[0] = 1;
              break;
            }
            case B_MIN:
            case B_MAX: {
              if (is_float_type(ct) && isnan(bits_as_double(vb, ct, cw))) break;
              const uint64_t o = ordered_of(vb, ct, cw);
              unsigned long long* w = reinterpret_cast<unsigned long long*>(lds + a.off) + h;
              if (a.kind == B_MIN) atomicMin(w, (unsigned long long)o);
              else atomicMax(w, (unsigned long long)o);
              break;
            }
            default: break;
          }
        }
      }
      __syncthreads();
      if (overflow) {  // redo the whole partition with twice as many sub-passes
        if (passes >= kMaxSubPass) {
          if (threadIdx.x == 0) {
            p.flags[1] = 1;
            ngroups    = p.stage_cap;  // report the partition as full: the plan gives it more partitions
          }
          break;
        }
        passes *= 2;
        pass = 0;
        if (threadIdx.x == 0) ngroups = 0;
        __syncthreads();
        continue;
      }
      // stage the groups of this pass
      for (int h = threadIdx.x; h < S; h += kGT) {
        if (keys[h] == kBE && h < p.lslots) continue;
        if (h >= p.lslots && rowsc[h] == 0) continue;  // special slot unused
        const int gi = atomicAdd(&ngroups, 1);
        if (gi >= p.stage_cap) {
          p.flags[1] = 1;
          continue;
        }
        const int64_t g = sbase + gi;
        for (int k = 0; k < p.nout; ++k) {
          const BOutDev o = p.o[k];
          bool ok         = true;
          uint64_t bits   = 0;
          const uint64_t vc =
            o.vcount >= 0 ? reinterpret_cast<const uint32_t*>(lds + p.a[o.vcount].off)[h] : rowsc[h];
          switch (o.kind) {
            case -1:
              ok   = h != p.lslots + 1;
              bits = h == p.lslots ? kBE : (h == p.lslots + 1 ? 0 : keys[h]);
              break;
            case LDF_AGG_COUNT:
              bits = o.mode == LDF_COUNT_ALL ? rowsc[h] : (o.mode == LDF_COUNT_ONLY_NULL ? rowsc[h] - vc : vc);
              break;
            case LDF_AGG_COUNT_ALL: bits = rowsc[h]; break;
            case LDF_AGG_SUM:
            case LDF_AGG_MEAN: {
              ok                 = vc > 0;
              const BAccDev a    = p.a[o.main];
              if (a.kind == B_FSUM) {
                uint64_t s0, s1, s2;
                lds_fsum_value(reinterpret_cast<const unsigned long long*>(lds + a.off) + 6 * h, s0, s1, s2);
                double v = ok ? fixed192_to_double(s0, s1, s2) : 0.0;
                if (o.kind == LDF_AGG_MEAN) v = ok ? v / (double)vc : 0.0;
                bits = (uint64_t)__double_as_longlong(v);
              } else {
                bits = ok ? reinterpret_cast<const unsigned long long*>(lds + a.off)[h] : 0;
              }
              break;
            }
            case LDF_AGG_MIN:
            case LDF_AGG_MAX: {
              ok                   = vc > 0;
              const uint64_t ord   = reinterpret_cast<const unsigned long long*>(lds + p.a[o.main].off)[h];
              const uint64_t ident = o.kind == LDF_AGG_MIN ? ~0ull : 0ull;
              if (is_float_type(o.in_type)) {
                double d;
                if (ord == ident) d = __longlong_as_double(0x7ff8000000000000ll);  // only NaN seen
                else d = __longlong_as_double((long long)((ord >> 63) ? (ord & 0x7fffffffffffffffull) : ~ord));
                bits = o.in_type == LDF_TYPE_FLOAT ? (uint64_t)__float_as_uint((float)d)
                                                   : (uint64_t)__double_as_longlong(d);
              } else if (is_unsigned_type(o.in_type)) {
                bits = ord;
              } else {
                bits = ord ^ 0x8000000000000000ull;
              }
              if (!ok) bits = 0;
              break;
            }
            default: break;
          }
          bg_store(o.stage, o.width, g, bits);
          if (o.stage_valid) o.stage_valid[g] = ok ? 1 : 0;
        }
      }
      __syncthreads();
      ++pass;
    }
    if (threadIdx.x == 0) p.gcount[q] = (uint32_t)(ngroups < p.stage_cap ? ngroups : p.stage_cap);
  }
}

// ooff[i] = first output row of instance i (gpos at its first partition)
__global__ void k_bg_ooff(const GInst* __restrict__ inst, int ninst, const uint64_t* __restrict__ gpos, int64_t nparts,
                          int64_t* __restrict__ ooff)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= ninst; i += gridDim.x * blockDim.x)
    ooff[i] = (int64_t)(i < ninst ? gpos[inst[i].pbase] : gpos[nparts]);
}

// Output row i comes from the partition q whose [gpos[q], gpos[q+1]) holds it, staged at
// q * stage_cap + (i - gpos[q]).  One wave covers 64 consecutive output rows.
__global__ __launch_bounds__(256) void k_bg_gather(GPlanB p, const uint64_t* __restrict__ gpos, int64_t nparts,
                                                   int64_t cap, unsigned long long* __restrict__ nulls)
{
  const int lane   = threadIdx.x & 63;
  const int64_t n0 = (int64_t)gpos[nparts];
  const int64_t n  = n0 < cap ? n0 : cap;
  for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + lane;
    int64_t q       = 0;
    if (lane == 0) {
      int64_t lo = 0, hi = nparts - 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if ((int64_t)gpos[mid] <= base) lo = mid;
        else hi = mid - 1;
      }
      q = lo;
    }
    q = __shfl(q, 0);
    if (i < n)
      while ((int64_t)gpos[q + 1] <= i) ++q;
    const int64_t si = q * (int64_t)p.stage_cap + (i - (int64_t)gpos[q]);
    for (int k = 0; k < p.nout; ++k) {
      const BOutDev o = p.o[k];
      bool ok         = false;
      if (i < n) {
        switch (o.width) {
          case 1: static_cast<uint8_t*>(o.out)[i] = static_cast<const uint8_t*>(o.stage)[si]; break;
          case 2: static_cast<uint16_t*>(o.out)[i] = static_cast<const uint16_t*>(o.stage)[si]; break;
          case 4: static_cast<uint32_t*>(o.out)[i] = static_cast<const uint32_t*>(o.stage)[si]; break;
          default: static_cast<uint64_t*>(o.out)[i] = static_cast<const uint64_t*>(o.stage)[si]; break;
        }
        ok = o.stage_valid == nullptr || o.stage_valid[si] != 0;
      }
      if (o.out_valid != nullptr) {
        const uint64_t b = __ballot(ok);
        if (lane == 0) {
          o.out_valid[base >> 6] = b;
          const int64_t in_range = n - base < 64 ? n - base : 64;
          const uint64_t z       = (uint64_t)in_range - (uint64_t)__popcll(b);
          if (z) atomicAdd(&nulls[k], (unsigned long long)z);
        }
      }
    }
  }
}

// distinct-count estimate per instance from up to kEstSample rows (GEE: sqrt(n/s) f1 + d - f1)
constexpr int kEstSample = 2048;
constexpr int kEstSlots  = 4096;
__global__ __launch_bounds__(256) void k_bg_estimate(const GInst* __restrict__ inst, int kwidth, double* __restrict__ est)
{
  __shared__ unsigned long long keys[kEstSlots];
  __shared__ uint32_t cnt[kEstSlots];
  __shared__ int d, f1;
  const GInst I = inst[blockIdx.x];
  for (int h = threadIdx.x; h < kEstSlots; h += blockDim.x) {
    keys[h] = kBE;
    cnt[h]  = 0;
  }
  if (threadIdx.x == 0) d = f1 = 0;
  __syncthreads();
  const int64_t s      = I.n < kEstSample ? I.n : kEstSample;
  const int64_t stride = s > 0 ? I.n / s : 1;
  for (int64_t j = threadIdx.x; j < s; j += blockDim.x) {
    const int64_t r  = j * stride;
    const bool kv    = valid_at(I.kvalid, r);
    uint64_t kb      = kv ? load_bits(I.key, kwidth, r) : 0x5bd1e995ull;
    if (kb == kBE) kb = 0x1234567ull;  // stands in for the special key in the sample
    int h = (int)(mix64(kb) & (kEstSlots - 1));
    while (true) {
      unsigned long long old = atomicCAS(&keys[h], kBE, (unsigned long long)kb);
      if (old == kBE || old == kb) break;
      h = (h + 1) & (kEstSlots - 1);
    }
    atomicAdd(&cnt[h], 1u);
  }
  __syncthreads();
  for (int h = threadIdx.x; h < kEstSlots; h += blockDim.x) {
    if (cnt[h] > 0) atomicAdd(&d, 1);
    if (cnt[h] == 1) atomicAdd(&f1, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double e = (double)d;
    if (s < I.n && s > 0) e = sqrt((double)I.n / (double)s) * f1 + (double)(d - f1);
    est[blockIdx.x] = e;
  }
}

// =========================================================================================
// host: batch objects
// =========================================================================================
int resident_blocks(int threads, size_t lds_bytes)
{
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  int per_cu = 2048 / threads;  // 32 waves per CU
  if (lds_bytes > 0) {
    const int by_lds = (int)((160 * 1024) / lds_bytes);
    per_cu           = std::max(1, std::min(per_cu, by_lds));
  }
  return cus * per_cu;
}

struct BatchImpl {
  virtual ~BatchImpl() { release(); }
  virtual void enqueue(hipStream_t s) = 0;
  virtual void result(hipStream_t s, ldf_table_out* out, int64_t* offsets) = 0;

  ldf_batch_config cfg{};
  hipStream_t cs = nullptr;  // stream of the plan's allocations
  std::vector<std::pair<void*, size_t>> owned;
  hipGraph_t graph     = nullptr;
  hipGraphExec_t exec  = nullptr;
  hipStream_t capture  = nullptr;

  void* dalloc(size_t bytes)
  {
    Alloc a(nullptr, cs);
    void* p = a.allocate(bytes > 0 ? bytes : 1);
    owned.push_back({p, bytes > 0 ? bytes : 1});
    return p;
  }
  template <typename T>
  T* dnew(int64_t n)
  {
    return static_cast<T*>(dalloc((size_t)std::max<int64_t>(1, n) * sizeof(T)));
  }
  template <typename T>
  T* upload(const std::vector<T>& v)
  {
    T* d = dnew<T>((int64_t)v.size());
    if (!v.empty()) LDF_HIP_TRY(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, cs));
    return d;
  }
  // Buffers go back to the pool under the stream they were allocated on; runs on other streams
  // (the public API allows any) are waited for with a device synchronisation unless the owner
  // only ever ran the batch on `cs`.
  bool only_cs = false;
  // frees every plan buffer except `keep` (after a build whose other buffers are no longer used)
  void keep_only(const std::vector<const void*>& keep)
  {
    LDF_HIP_TRY(hipStreamSynchronize(cs));
    Alloc a(nullptr, cs);
    std::vector<std::pair<void*, size_t>> rest;
    for (auto& o : owned) {
      if (std::find(keep.begin(), keep.end(), (const void*)o.first) != keep.end()) rest.push_back(o);
      else a.deallocate(o.first, o.second);
    }
    owned.swap(rest);
  }
  void release()
  {
    if (only_cs) (void)hipStreamSynchronize(cs);
    else (void)hipDeviceSynchronize();
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture) (void)hipStreamDestroy(capture);
    exec    = nullptr;
    graph   = nullptr;
    capture = nullptr;
    Alloc a(nullptr, cs);
    for (auto& o : owned) a.deallocate(o.first, o.second);
    owned.clear();
    (void)hipGetLastError();
  }
  // Runs the batch on `s`: the first run with use_graph captures the launch sequence (on a
  // private stream, so nothing executes during capture) and every run launches the graph.
  void run(hipStream_t s)
  {
    if (!cfg.use_graph) {
      enqueue(s);
      return;
    }
    if (exec == nullptr) {
      if (capture == nullptr) LDF_HIP_TRY(hipStreamCreateWithFlags(&capture, hipStreamNonBlocking));
      LDF_HIP_TRY(hipStreamBeginCapture(capture, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(capture);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(capture, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      LDF_HIP_TRY(hipStreamEndCapture(capture, &graph));
      LDF_HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    LDF_HIP_TRY(hipGraphLaunch(exec, s));
  }
};

void check_config(const ldf_batch_config& c, bool join = false)
{
  if (join) {  // joins: lds_slots = 1 probes every table in HBM
    LDF_EXPECTS(c.load_factor > 0.0 && c.load_factor <= 0.9, "batch: load_factor must be in (0, 0.9]");
    LDF_EXPECTS(c.tile_rows == 1024 || c.tile_rows == 2048 || c.tile_rows == 4096,
                "batch: tile_rows must be 1024, 2048 or 4096");
    LDF_EXPECTS(c.lds_slots == 1 || c.lds_slots == 256 || c.lds_slots == 512 || c.lds_slots == 1024 ||
                  c.lds_slots == 2048,
                "batch join: lds_slots must be 1 (tables in HBM), 256, 512, 1024 or 2048");
    return;
  }
  LDF_EXPECTS(c.load_factor > 0.0 && c.load_factor <= 0.9, "batch: load_factor must be in (0, 0.9]");
  LDF_EXPECTS(c.tile_rows == 1024 || c.tile_rows == 2048 || c.tile_rows == 4096,
              "batch: tile_rows must be 1024, 2048 or 4096");
  LDF_EXPECTS(c.lds_slots == 256 || c.lds_slots == 512 || c.lds_slots == 1024 || c.lds_slots == 2048,
              "batch: lds_slots must be 256, 512, 1024 or 2048");
}

// instances in decreasing size: the order in which their work items are listed
std::vector<int> by_size(const std::vector<int64_t>& sz)
{
  std::vector<int> o(sz.size());
  for (size_t i = 0; i < o.size(); ++i) o[i] = (int)i;
  std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return sz[a] > sz[b]; });
  return o;
}

// ---- join batch ---------------------------------------------------------------------------
struct JoinBatch : BatchImpl {
  int ninst = 0, w = 0, jt = 0, tile = 0, btile = 0, nout = 0, grid_q = 0;
  int lds_max = 0;  // tables of at most this many slots are probed from LDS (0: all in HBM)
  bool ne       = true;
  bool prebuilt = false;  // probes prebuilt indexes: no build phases
  int64_t nslot = 0, nbt = 0, npt = 0, ntiles = 0, out_cap = 0;
  int nbitems = 0, npitems = 0;
  JInst* d_inst     = nullptr;
  Item* d_bitems    = nullptr;
  Item* d_pitems    = nullptr;
  BSlot* tab        = nullptr;
  int64_t* slot_row = nullptr;
  uint32_t* rows    = nullptr;
  uint32_t* cnt32   = nullptr;
  uint64_t* start   = nullptr;
  uint32_t* fillc   = nullptr;
  uint64_t* info    = nullptr;
  uint64_t* btot    = nullptr;
  uint64_t* boff    = nullptr;
  int64_t* ooff     = nullptr;
  uint64_t* ws      = nullptr;
  uint32_t* cursors = nullptr;
  int32_t* flags    = nullptr;
  int32_t* lmap     = nullptr;
  int32_t* rmap     = nullptr;
  uint8_t* hit      = nullptr;  // FULL: build slots some probe row matched
  const void** d_src         = nullptr;
  const uint8_t** d_sval     = nullptr;
  unsigned long long* nulls  = nullptr;
  BGather g{};
  std::vector<ldf_column> cols;

  void enqueue_count(hipStream_t s)
  {
    fill_u32(cursors, 0, 6, s);
    fill_u32(reinterpret_cast<uint32_t*>(flags), 0, 2, s);
    if (!prebuilt) build(s);
    if (hit != nullptr) fill_u8(hit, 0, nslot, s);
    if (npitems > 0) {
      hipLaunchKernelGGL(k_bj_count, dim3(std::min(grid_q, npitems)), dim3(kQT),
                         lds_max > 0 ? (size_t)(lds_max + 2) * sizeof(BSlot) : 0, s, (const JInst*)d_inst,
                         (const Item*)d_pitems, npitems, cursors + 1, tile, w, ne, jt, info, btot, hit, lds_max);
      LDF_HIP_TRY(hipGetLastError());
    }
    if (hit != nullptr && nbitems > 0) {
      hipLaunchKernelGGL(k_bj_unmatched, dim3(std::min(grid_q, nbitems)), dim3(kQT), 0, s, (const JInst*)d_inst,
                         (const Item*)d_bitems, nbitems, cursors + 4, btile, (const int64_t*)slot_row,
                         (const uint8_t*)hit, btot, (const uint64_t*)nullptr, (int64_t)0, (int32_t*)nullptr,
                         (int32_t*)nullptr, flags);
      LDF_HIP_TRY(hipGetLastError());
    }
    exclusive_scan_u64_ws(btot, boff, ntiles, ws, s);
    hipLaunchKernelGGL(k_bj_ooff, dim3(grid_for(ninst + 1, 256)), dim3(256), 0, s, (const JInst*)d_inst, ninst,
                       (const uint64_t*)boff, ntiles, ooff);
    LDF_HIP_TRY(hipGetLastError());
  }

  // hash tables of all instances' build sides (CSR multimaps)
  void build(hipStream_t s)
  {
    hipLaunchKernelGGL(k_bj_fill, dim3(grid_for(nslot, 256)), dim3(256), 0, s, tab, nslot);
    LDF_HIP_TRY(hipGetLastError());
    if (nbitems > 0) {
      // FULL places null build keys in the null slot even when nulls do not match (no probe then
      // finds them, and they come out as unmatched build rows)
      hipLaunchKernelGGL(k_bj_build, dim3(std::min(grid_q, nbitems)), dim3(kQT), 0, s, (const JInst*)d_inst,
                         (const Item*)d_bitems, nbitems, cursors + 0, btile, w, ne || jt == LDF_JOIN_FULL, tab,
                         slot_row);
      LDF_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_bj_counts, dim3(grid_for(nslot, 256)), dim3(256), 0, s, (const BSlot*)tab, nslot, cnt32);
    LDF_HIP_TRY(hipGetLastError());
    exclusive_scan_u32_to_u64_ws(cnt32, start, nslot, ws, s);
    hipLaunchKernelGGL(k_bj_starts, dim3(grid_for(nslot, 256)), dim3(256), 0, s, tab, nslot, (const uint64_t*)start);
    LDF_HIP_TRY(hipGetLastError());
    fill_u32(fillc, 0, nslot, s);
    if (nbt > 0) {
      hipLaunchKernelGGL(k_bj_group, dim3(grid_for(nbt, 256)), dim3(256), 0, s, (const int64_t*)slot_row, nbt,
                         (const BSlot*)tab, fillc, rows);
      LDF_HIP_TRY(hipGetLastError());
    }
  }

  void enqueue(hipStream_t s) override
  {
    enqueue_count(s);
    enqueue_tail(s);
  }

  // emit + gather, after a count pass
  void enqueue_tail(hipStream_t s)
  {
    fill_u64(reinterpret_cast<uint64_t*>(nulls), 0, kMaxBOut, s);
    const size_t lds = (size_t)(tile + 1) * 8 + (size_t)tile * 8;
    if (npitems > 0 && out_cap > 0) {
      hipLaunchKernelGGL(k_bj_emit, dim3(std::min(grid_q, npitems)), dim3(kQT), lds, s, (const JInst*)d_inst,
                         (const Item*)d_pitems, npitems, cursors + 2, tile, jt, (const uint64_t*)info,
                         (const uint64_t*)boff, out_cap, lmap, rmap, flags);
      LDF_HIP_TRY(hipGetLastError());
    }
    if (hit != nullptr && nbitems > 0 && out_cap > 0) {
      hipLaunchKernelGGL(k_bj_unmatched, dim3(std::min(grid_q, nbitems)), dim3(kQT), 0, s, (const JInst*)d_inst,
                         (const Item*)d_bitems, nbitems, cursors + 5, btile, (const int64_t*)slot_row,
                         (const uint8_t*)hit, btot, (const uint64_t*)boff, out_cap, lmap, rmap, flags);
      LDF_HIP_TRY(hipGetLastError());
    }
    if (out_cap > 0 && nout > 0) {
      hipLaunchKernelGGL(k_bj_gather, dim3(grid_for(out_cap, 256, 256 * 16)), dim3(256), 0, s, g,
                         (const void* const*)d_src, (const uint8_t* const*)d_sval, ninst, (const int64_t*)ooff,
                         out_cap, (const int32_t*)lmap, (const int32_t*)rmap, nulls);
      LDF_HIP_TRY(hipGetLastError());
    }
  }

  void result(hipStream_t s, ldf_table_out* out, int64_t* offsets) override
  {
    LDF_EXPECTS(out != nullptr && out->cols != nullptr && out->ncols >= nout, "batch result: output table too small");
    LDF_HIP_TRY(hipStreamSynchronize(s));
    int32_t hf[2];
    std::vector<int64_t> ho(ninst + 1);
    uint64_t hn[kMaxBOut];
    LDF_HIP_TRY(hipMemcpy(hf, flags, sizeof(hf), hipMemcpyDeviceToHost));
    LDF_HIP_TRY(hipMemcpy(ho.data(), ooff, ho.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    LDF_HIP_TRY(hipMemcpy(hn, nulls, sizeof(hn), hipMemcpyDeviceToHost));
    if (hf[0] != 0 || ho[ninst] > out_cap)
      throw Error(LDF_E_INVALID, "batch result: the inputs produce more rows than when the batch was created");
    if (offsets) std::copy(ho.begin(), ho.end(), offsets);
    out->ncols = nout;
    for (int c = 0; c < nout; ++c) {
      ldf_column o = cols[c];
      o.length     = ho[ninst];
      o.null_count = o.validity ? (int64_t)hn[c] : 0;
      out->cols[c] = o;
    }
  }
};

// A built join index: the table and grouped rows of one build side (instance 0 of a JoinBatch
// whose probe side is empty).
struct JoinIndex {
  BSlot* tab            = nullptr;
  const uint32_t* rows  = nullptr;
  uint64_t cap          = 0;
};

JoinBatch* create_join(int32_t ninst, const ldf_table_view* lhs, const ldf_table_view* rhs, const int32_t* lhs_keys,
                       const int32_t* rhs_keys, int32_t nkeys, int32_t jt, bool ne, const int32_t* lhs_out,
                       int32_t n_lhs_out, const int32_t* rhs_out, int32_t n_rhs_out, const ldf_batch_config& cfg,
                       hipStream_t s, const std::vector<JoinIndex>* prebuilt = nullptr)
{
  LDF_EXPECTS(ninst >= 1 && lhs != nullptr && rhs != nullptr, "batch join: no instances");
  LDF_EXPECTS(nkeys == 1 && lhs_keys != nullptr && rhs_keys != nullptr, "batch join: exactly one key column");
  LDF_EXPECTS(jt >= LDF_JOIN_INNER && jt <= LDF_JOIN_ANTI, "Unknown JoinType");
  if (jt == LDF_JOIN_FULL && prebuilt != nullptr)
    throw Error(LDF_E_UNSUPPORTED, "batch join: FULL joins against prebuilt indexes are not supported");
  if ((jt == LDF_JOIN_SEMI || jt == LDF_JOIN_ANTI) && n_rhs_out > 0)
    throw Error(LDF_E_INVALID, "rhs must have no output columns for SEMI and ANTI join");
  LDF_EXPECTS(n_lhs_out >= 0 && n_rhs_out >= 0 && n_lhs_out + n_rhs_out <= kMaxBOut,
              "batch join: at most 16 output columns");
  check_config(cfg, true);
  std::unique_ptr<JoinBatch> b(new JoinBatch());
  b->lds_max = cfg.lds_slots > 1 && prebuilt == nullptr ? cfg.lds_slots : 0;
  b->cfg   = cfg;
  b->cs    = s;
  b->ninst = ninst;
  b->jt    = jt;
  b->ne    = ne;
  b->tile  = cfg.tile_rows;
  b->btile = cfg.tile_rows * 4;
  b->nout  = n_lhs_out + n_rhs_out;

  // schema: every instance has the types of instance 0
  const ldf_table_view& L0 = lhs[0];
  const ldf_table_view& R0 = rhs[0];
  const int lk = lhs_keys[0], rk = rhs_keys[0];
  LDF_EXPECTS(lk >= 0 && lk < L0.ncols && rk >= 0 && rk < R0.ncols, "batch join: key index out of range");
  const int32_t ktype = L0.cols[lk].type_id;
  LDF_EXPECTS(R0.cols[rk].type_id == ktype, "join: lhs and rhs key column types differ");
  b->w = type_width(ktype);
  if (b->w <= 0) throw Error(LDF_E_UNSUPPORTED, "batch join: the key must be a fixed-width column");
  std::vector<int> oside, ocol;
  for (int i = 0; i < n_lhs_out; ++i) {
    LDF_EXPECTS(lhs_out[i] >= 0 && lhs_out[i] < L0.ncols, "join: lhs_out index out of range");
    oside.push_back(0);
    ocol.push_back(lhs_out[i]);
  }
  for (int i = 0; i < n_rhs_out; ++i) {
    LDF_EXPECTS(rhs_out[i] >= 0 && rhs_out[i] < R0.ncols, "join: rhs_out index out of range");
    oside.push_back(1);
    ocol.push_back(rhs_out[i]);
  }
  std::vector<bool> nullable(b->nout, false);
  for (int c = 0; c < b->nout; ++c) {
    const int32_t t = (oside[c] ? R0 : L0).cols[ocol[c]].type_id;
    if (type_width(t) <= 0) throw Error(LDF_E_UNSUPPORTED, "batch join: output columns must be fixed-width");
    nullable[c] = jt == LDF_JOIN_FULL || (oside[c] == 1 && jt == LDF_JOIN_LEFT);
  }
  std::vector<int64_t> np(ninst), nb(ninst);
  for (int i = 0; i < ninst; ++i) {
    const ldf_table_view& L = lhs[i];
    const ldf_table_view& R = rhs[i];
    LDF_EXPECTS(L.ncols == L0.ncols && R.ncols == R0.ncols, "batch join: instances differ in their columns");
    np[i] = L.ncols > 0 ? L.cols[0].length : 0;
    nb[i] = R.ncols > 0 ? R.cols[0].length : 0;
    for (int c = 0; c < L.ncols; ++c) {
      LDF_EXPECTS(L.cols[c].type_id == L0.cols[c].type_id, "batch join: instances differ in their column types");
      validate_view(L.cols[c], np[i], "batch join lhs");
    }
    for (int c = 0; c < R.ncols; ++c) {
      LDF_EXPECTS(R.cols[c].type_id == R0.cols[c].type_id, "batch join: instances differ in their column types");
      validate_view(R.cols[c], nb[i], "batch join rhs");
    }
    LDF_EXPECTS(np[i] < INT32_MAX && nb[i] < INT32_MAX, "join: each side is limited to 2^31-1 rows per call");
    for (int c = 0; c < b->nout; ++c)
      if ((oside[c] ? R : L).cols[ocol[c]].validity != nullptr) nullable[c] = true;
  }

  // layout
  std::vector<JInst> hi(ninst);
  int64_t tabo = 0, brow = 0, prow = 0, ptile = 0;
  for (int i = 0; i < ninst; ++i) {
    JInst& I   = hi[i];
    I.pkey     = lhs[i].cols[lk].data;
    I.pvalid   = lhs[i].cols[lk].validity;
    I.np       = np[i];
    I.bkey     = rhs[i].cols[rk].data;
    I.bvalid   = rhs[i].cols[rk].validity;
    I.nb       = nb[i];
    I.cap      = prebuilt ? (*prebuilt)[i].cap
                          : next_pow2((uint64_t)std::max<double>(64.0, std::ceil((double)nb[i] / cfg.load_factor)));
    I.mask     = I.cap - 1;
    I.tab      = tabo;
    I.brow     = prebuilt ? 0 : brow;
    I.prow     = prow;
    I.ptile    = ptile;
    if (!prebuilt) {
      tabo += (int64_t)I.cap + 2;
      brow += nb[i];
    }
    prow += np[i];
    ptile += (np[i] + b->tile - 1) / b->tile;
    I.utile = ptile;  // FULL: the instance's unmatched build rows come after its probe rows
    if (jt == LDF_JOIN_FULL) ptile += (nb[i] + b->btile - 1) / b->btile;
  }
  b->nslot  = tabo;
  b->nbt    = brow;
  b->npt    = prow;
  b->ntiles = ptile;
  LDF_EXPECTS(brow < (int64_t)UINT32_MAX, "batch join: at most 2^32-1 build rows per batch");
  std::vector<Item> bi, pi;
  for (int i : by_size(np)) {
    if (!prebuilt)
      for (int64_t t = 0; t < (nb[i] + b->btile - 1) / b->btile; ++t) bi.push_back({i, (int32_t)t});
    for (int64_t t = 0; t < (np[i] + b->tile - 1) / b->tile; ++t) pi.push_back({i, (int32_t)t});
  }
  LDF_EXPECTS(bi.size() < (size_t)INT32_MAX / 2 && pi.size() < (size_t)INT32_MAX / 2, "batch join: too many tiles");
  b->nbitems = (int)bi.size();
  b->npitems = (int)pi.size();
  b->grid_q  = resident_blocks(kQT, (size_t)(b->tile + 1) * 8 + (size_t)b->tile * 8);

  b->prebuilt = prebuilt != nullptr;
  b->tab      = b->dnew<BSlot>(b->nslot);
  b->slot_row = b->dnew<int64_t>(b->nbt);
  b->rows     = b->dnew<uint32_t>(b->nbt);
  b->cnt32    = b->dnew<uint32_t>(b->nslot);
  b->start    = b->dnew<uint64_t>(b->nslot + 1);
  b->fillc    = b->dnew<uint32_t>(b->nslot);
  b->info     = b->dnew<uint64_t>(b->npt);
  b->btot     = b->dnew<uint64_t>(b->ntiles);
  b->boff     = b->dnew<uint64_t>(b->ntiles + 1);
  b->ooff     = b->dnew<int64_t>(ninst + 1);
  b->ws       = b->dnew<uint64_t>(std::max(scan_workspace_elems(b->nslot), scan_workspace_elems(b->ntiles)));
  b->cursors  = b->dnew<uint32_t>(6);
  if (jt == LDF_JOIN_FULL) b->hit = b->dnew<uint8_t>(b->nslot);
  b->flags    = b->dnew<int32_t>(2);
  b->nulls    = b->dnew<unsigned long long>(kMaxBOut);
  for (int i = 0; i < ninst; ++i) {
    hi[i].tabp  = prebuilt ? (*prebuilt)[i].tab : b->tab + hi[i].tab;
    hi[i].rowsp = prebuilt ? (*prebuilt)[i].rows : b->rows;
  }
  b->d_inst   = b->upload(hi);
  b->d_bitems = b->upload(bi);
  b->d_pitems = b->upload(pi);
  const size_t lds = (size_t)(b->tile + 1) * 8 + (size_t)b->tile * 8;
  if (lds > 64 * 1024) LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_bj_emit, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));

  // sizes: one count pass
  b->enqueue_count(s);
  uint64_t total = 0;
  LDF_HIP_TRY(hipMemcpyAsync(&total, b->boff + b->ntiles, sizeof(total), hipMemcpyDeviceToHost, s));
  LDF_HIP_TRY(hipStreamSynchronize(s));
  b->out_cap = (int64_t)total;
  b->lmap    = b->dnew<int32_t>(b->out_cap);
  b->rmap    = (jt == LDF_JOIN_SEMI || jt == LDF_JOIN_ANTI) ? nullptr : b->dnew<int32_t>(b->out_cap);

  // outputs and the per-instance source table of the gather
  b->g.ncols = b->nout;
  b->cols.resize(b->nout);
  for (int c = 0; c < b->nout; ++c) {
    const int32_t t = (oside[c] ? R0 : L0).cols[ocol[c]].type_id;
    const int wd    = type_width(t);
    ldf_column& o   = b->cols[c];
    std::memset(&o, 0, sizeof(o));
    o.type_id    = t;
    o.data_bytes = b->out_cap * wd;
    o.data       = b->dalloc((size_t)std::max<int64_t>(1, o.data_bytes));
    if (nullable[c]) {
      o.validity_bytes = ((b->out_cap + 63) / 64) * 8;
      o.validity       = static_cast<uint8_t*>(b->dalloc((size_t)std::max<int64_t>(8, o.validity_bytes)));
    }
    b->g.width[c] = wd;
    b->g.side[c]  = oside[c];
    b->g.dst[c]   = o.data;
    b->g.dval[c]  = reinterpret_cast<uint64_t*>(o.validity);
  }
  std::vector<const void*> src((size_t)ninst * std::max(1, b->nout));
  std::vector<const uint8_t*> sval(src.size());
  for (int i = 0; i < ninst; ++i)
    for (int c = 0; c < b->nout; ++c) {
      const ldf_column_view& v = (oside[c] ? rhs[i] : lhs[i]).cols[ocol[c]];
      src[(size_t)i * b->nout + c]  = v.data;
      sval[(size_t)i * b->nout + c] = v.validity;
    }
  b->d_src  = b->upload(src);
  b->d_sval = b->upload(sval);
  LDF_HIP_TRY(hipStreamSynchronize(s));
  return b.release();
}

// ---- groupby batch --------------------------------------------------------------------------
struct GroupbyBatch : BatchImpl {
  int ninst = 0, tile = 0, grid_g = 0, nout = 0;
  int64_t nrows = 0, nparts = 0, out_cap = 0;
  int nritems = 0, npitems = 0;
  GInst* d_inst   = nullptr;
  Item* d_ritems  = nullptr;
  Item* d_pitems  = nullptr;
  uint64_t* gpos  = nullptr;
  int64_t* ooff   = nullptr;
  uint64_t* ws    = nullptr;
  uint32_t* cursors = nullptr;
  unsigned long long* nulls = nullptr;
  GPlanB p{};
  std::vector<ldf_column> cols;
  std::vector<int64_t> pbase;  // host copy: first partition per instance
  std::vector<int32_t> pbits;

  void enqueue(hipStream_t s) override
  {
    fill_u32(cursors, 0, 4, s);
    fill_u32(reinterpret_cast<uint32_t*>(p.flags), 0, 2, s);
    fill_u32(p.pcount, 0, nparts, s);
    fill_u64(reinterpret_cast<uint64_t*>(nulls), 0, kMaxBAgg + 1, s);
    if (nritems > 0) {
      hipLaunchKernelGGL(k_bg_hist, dim3(std::min(grid_g, nritems)), dim3(kGT), 0, s, (const GInst*)d_inst,
                         (const Item*)d_ritems, nritems, cursors + 0, tile, p);
      LDF_HIP_TRY(hipGetLastError());
    }
    exclusive_scan_u32_to_u64_ws(p.pcount, p.pstart, nparts, ws, s);
    hipLaunchKernelGGL(k_bg_cursor_init, dim3(grid_for(nparts, 256)), dim3(256), 0, s, (const uint64_t*)p.pstart,
                       nparts, p.pcursor);
    LDF_HIP_TRY(hipGetLastError());
    if (nritems > 0) {
      hipLaunchKernelGGL(k_bg_scatter, dim3(std::min(grid_g, nritems)), dim3(kGT), 0, s, (const GInst*)d_inst,
                         (const Item*)d_ritems, nritems, cursors + 1, tile, p);
      LDF_HIP_TRY(hipGetLastError());
    }
    if (npitems > 0) {
      hipLaunchKernelGGL(k_bg_aggregate, dim3(std::min(resident_blocks(kGT, p.lds_bytes), npitems)), dim3(kGT),
                         p.lds_bytes, s, (const GInst*)d_inst, (const Item*)d_pitems, npitems, cursors + 2, p);
      LDF_HIP_TRY(hipGetLastError());
    }
    exclusive_scan_u32_to_u64_ws(p.gcount, gpos, nparts, ws, s);
    hipLaunchKernelGGL(k_bg_ooff, dim3(grid_for(ninst + 1, 256)), dim3(256), 0, s, (const GInst*)d_inst, ninst,
                       (const uint64_t*)gpos, nparts, ooff);
    LDF_HIP_TRY(hipGetLastError());
    if (out_cap > 0) {
      hipLaunchKernelGGL(k_bg_gather, dim3(grid_for(out_cap, 256, 256 * 16)), dim3(256), 0, s, p,
                         (const uint64_t*)gpos, nparts, out_cap, nulls);
      LDF_HIP_TRY(hipGetLastError());
    }
  }

  // the final gather only, after create's sizing run (whose partition results are still in place)
  void enqueue_tail(hipStream_t s)
  {
    fill_u64(reinterpret_cast<uint64_t*>(nulls), 0, kMaxBAgg + 1, s);
    if (out_cap > 0) {
      hipLaunchKernelGGL(k_bg_gather, dim3(grid_for(out_cap, 256, 256 * 16)), dim3(256), 0, s, p,
                         (const uint64_t*)gpos, nparts, out_cap, nulls);
      LDF_HIP_TRY(hipGetLastError());
    }
  }

  void result(hipStream_t s, ldf_table_out* out, int64_t* offsets) override
  {
    LDF_EXPECTS(out != nullptr && out->cols != nullptr && out->ncols >= nout, "batch result: output table too small");
    LDF_HIP_TRY(hipStreamSynchronize(s));
    int32_t hf[2];
    std::vector<int64_t> ho(ninst + 1);
    uint64_t hn[kMaxBAgg + 1];
    LDF_HIP_TRY(hipMemcpy(hf, p.flags, sizeof(hf), hipMemcpyDeviceToHost));
    LDF_HIP_TRY(hipMemcpy(ho.data(), ooff, ho.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    LDF_HIP_TRY(hipMemcpy(hn, nulls, sizeof(hn), hipMemcpyDeviceToHost));
    if (hf[0] != 0)
      throw Error(LDF_E_UNSUPPORTED,
                  "batch groupby: a SUM/MEAN input is non-finite or |x| >= 2^62 (outside the exact range); use ldf_groupby");
    if (hf[1] != 0 || ho[ninst] > out_cap)
      throw Error(LDF_E_INVALID, "batch result: the inputs produce more groups than when the batch was created");
    if (offsets) std::copy(ho.begin(), ho.end(), offsets);
    out->ncols = nout;
    for (int c = 0; c < nout; ++c) {
      ldf_column o = cols[c];
      o.length     = ho[ninst];
      o.null_count = o.validity ? (int64_t)hn[c] : 0;
      out->cols[c] = o;
    }
  }
};

struct GbSchema {
  int32_t ktype = 0;
  std::vector<int> vcols;  // table column of value column c
  std::vector<int32_t> vtype;
  std::vector<bool> vnull;  // some instance has nulls in value column c
  bool knull = false;
};

GroupbyBatch* create_groupby(int32_t ninst, const ldf_table_view* tables, const int32_t* keys, int32_t nkeys,
                             const ldf_agg* aggs, int32_t naggs, const ldf_batch_config& cfg, hipStream_t s,
                             const std::vector<int32_t>* pbits_in = nullptr)
{
  LDF_EXPECTS(ninst >= 1 && tables != nullptr, "batch groupby: no instances");
  LDF_EXPECTS(nkeys == 1 && keys != nullptr, "batch groupby: exactly one key column");
  LDF_EXPECTS(naggs >= 0 && (naggs == 0 || aggs != nullptr), "batch groupby: NULL aggregations");
  if (naggs > kMaxBAgg) throw Error(LDF_E_UNSUPPORTED, "batch groupby: at most 16 aggregations");
  check_config(cfg);
  const ldf_table_view& T0 = tables[0];
  LDF_EXPECTS(keys[0] >= 0 && keys[0] < T0.ncols, "groupby: key index out of range");
  GbSchema sc;
  sc.ktype = T0.cols[keys[0]].type_id;
  const int kw = type_width(sc.ktype);
  if (kw <= 0) throw Error(LDF_E_UNSUPPORTED, "batch groupby: the key must be a fixed-width column");
  auto vslot = [&](int col) {
    for (size_t c = 0; c < sc.vcols.size(); ++c)
      if (sc.vcols[c] == col) return (int)c;
    if (sc.vcols.size() >= (size_t)kMaxVal) throw Error(LDF_E_UNSUPPORTED, "batch groupby: at most 4 value columns");
    sc.vcols.push_back(col);
    sc.vtype.push_back(T0.cols[col].type_id);
    sc.vnull.push_back(false);
    return (int)sc.vcols.size() - 1;
  };
  for (int a = 0; a < naggs; ++a) {
    const ldf_agg& g = aggs[a];
    switch (g.kind) {
      case LDF_AGG_COUNT_ALL: break;
      case LDF_AGG_SUM:
      case LDF_AGG_MEAN:
      case LDF_AGG_COUNT:
      case LDF_AGG_MIN:
      case LDF_AGG_MAX: {
        LDF_EXPECTS(g.in_col >= 0 && g.in_col < T0.ncols, "groupby: aggregation column index out of range");
        const int32_t t = T0.cols[g.in_col].type_id;
        if (type_width(t) <= 0 || t == LDF_TYPE_BOOL)
          throw Error(LDF_E_UNSUPPORTED, "batch groupby: value columns must be fixed-width numeric");
        (void)vslot(g.in_col);
        break;
      }
      default: throw Error(LDF_E_UNSUPPORTED, "batch groupby: only SUM, MEAN, COUNT, COUNT_ALL, MIN and MAX");
    }
  }
  std::vector<int64_t> n(ninst);
  for (int i = 0; i < ninst; ++i) {
    const ldf_table_view& T = tables[i];
    LDF_EXPECTS(T.ncols == T0.ncols, "batch groupby: instances differ in their columns");
    n[i] = T.ncols > 0 ? T.cols[0].length : 0;
    // the exact sums' LDS words take fewer than 2^31 addends per table
    LDF_EXPECTS(n[i] < INT32_MAX, "batch groupby: each instance is limited to 2^31-1 rows");
    for (int c = 0; c < T.ncols; ++c) {
      LDF_EXPECTS(T.cols[c].type_id == T0.cols[c].type_id, "batch groupby: instances differ in their column types");
      validate_view(T.cols[c], n[i], "batch groupby");
    }
    if (T.cols[keys[0]].validity) sc.knull = true;
    for (size_t c = 0; c < sc.vcols.size(); ++c)
      if (T.cols[sc.vcols[c]].validity) sc.vnull[c] = true;
  }

  std::unique_ptr<GroupbyBatch> b(new GroupbyBatch());
  b->cfg   = cfg;
  b->cs    = s;
  b->ninst = ninst;
  b->tile  = cfg.tile_rows;
  GPlanB& p = b->p;
  p.ktype   = sc.ktype;
  p.kwidth  = kw;
  p.ncol    = (int)sc.vcols.size();
  for (int c = 0; c < p.ncol; ++c) {
    p.ctype[c]  = sc.vtype[c];
    p.cwidth[c] = type_width(sc.vtype[c]);
  }
  // accumulators (deduplicated) and their LDS arrays
  p.lslots      = cfg.lds_slots;
  const int S   = p.lslots + 2;
  int off       = S * 8;  // keys
  p.rows_off    = off;
  off += ((S * 4 + 7) / 8) * 8;
  p.nacc        = 0;
  auto acc = [&](int kind, int col) {
    for (int i = 0; i < p.nacc; ++i)
      if (p.a[i].kind == kind && p.a[i].col == col) return i;
    LDF_EXPECTS(p.nacc < kMaxBAcc, "batch groupby: too many accumulators");
    BAccDev& a = p.a[p.nacc];
    a.kind     = kind;
    a.col      = col;
    a.off      = off;
    off += kind == B_VCOUNT ? ((S * 4 + 7) / 8) * 8 : (kind == B_FSUM ? S * 48 : S * 8);
    return p.nacc++;
  };
  p.nout = 0;
  {  // key column
    BOutDev& o = p.o[p.nout++];
    std::memset(&o, 0, sizeof(o));
    o.kind     = -1;
    o.out_type = sc.ktype;
    o.width    = kw;
    o.nullable = sc.knull ? 1 : 0;
    o.main = o.vcount = -1;
  }
  for (int a = 0; a < naggs; ++a) {
    const ldf_agg& g = aggs[a];
    BOutDev& o       = p.o[p.nout++];
    std::memset(&o, 0, sizeof(o));
    o.kind   = g.kind;
    o.mode   = g.count_mode;
    o.main   = -1;
    o.vcount = -1;
    if (g.kind == LDF_AGG_COUNT_ALL) {
      o.out_type = LDF_TYPE_INT64;
    } else {
      const int c  = vslot(g.in_col);
      const int32_t t = sc.vtype[c];
      o.in_type       = t;
      o.vcount        = sc.vnull[c] ? acc(B_VCOUNT, c) : -1;
      switch (g.kind) {
        case LDF_AGG_SUM:
          o.main     = acc(is_float(t) ? B_FSUM : B_ISUM, c);
          o.out_type = is_float(t) ? LDF_TYPE_DOUBLE : (is_unsigned_int(t) ? LDF_TYPE_UINT64 : LDF_TYPE_INT64);
          o.nullable = 1;
          break;
        case LDF_AGG_MEAN:
          o.main     = acc(B_FSUM, c);
          o.out_type = LDF_TYPE_DOUBLE;
          o.nullable = 1;
          break;
        case LDF_AGG_MIN:
        case LDF_AGG_MAX:
          o.main     = acc(g.kind == LDF_AGG_MIN ? B_MIN : B_MAX, c);
          o.out_type = t;
          o.nullable = 1;
          break;
        default:  // COUNT
          o.out_type = LDF_TYPE_INT64;
          break;
      }
    }
    o.width = type_width(o.out_type);
  }
  p.lds_bytes = off;
  LDF_EXPECTS(p.lds_bytes <= 150 * 1024, "batch groupby: the LDS table does not fit; use a smaller lds_slots");
  p.stage_cap = kStageFactor * p.lslots;
  b->nout     = p.nout;

  // instances: distinct-count estimates -> partitions
  std::vector<GInst> hi(ninst);
  int64_t rows = 0;
  for (int i = 0; i < ninst; ++i) {
    GInst& I = hi[i];
    std::memset(&I, 0, sizeof(I));
    const ldf_table_view& T = tables[i];
    I.key    = T.cols[keys[0]].data;
    I.kvalid = T.cols[keys[0]].validity;
    I.n      = n[i];
    for (int c = 0; c < p.ncol; ++c) {
      I.val[c]    = T.cols[sc.vcols[c]].data;
      I.vvalid[c] = T.cols[sc.vcols[c]].validity;
    }
    rows += n[i];
  }
  std::vector<int32_t> pb(ninst, 0);
  if (pbits_in != nullptr) {
    pb = *pbits_in;
  } else {
    GInst* dtmp = static_cast<GInst*>(b->dalloc(sizeof(GInst) * ninst));
    double* dest = b->dnew<double>(ninst);
    LDF_HIP_TRY(hipMemcpyAsync(dtmp, hi.data(), sizeof(GInst) * ninst, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_bg_estimate, dim3(ninst), dim3(256), 0, s, (const GInst*)dtmp, kw, dest);
    LDF_HIP_TRY(hipGetLastError());
    std::vector<double> est(ninst);
    LDF_HIP_TRY(hipMemcpyAsync(est.data(), dest, sizeof(double) * ninst, hipMemcpyDeviceToHost, s));
    LDF_HIP_TRY(hipStreamSynchronize(s));
    const double per_part = 0.3 * p.lslots;  // keys a partition is planned for
    for (int i = 0; i < ninst; ++i) {
      const double want = std::ceil(1.25 * est[i] / per_part);
      int bits          = 0;
      while ((1 << bits) < want && (1 << bits) < kMaxParts) ++bits;
      pb[i] = bits;
    }
  }
  int64_t parts = 0;
  b->pbase.resize(ninst);
  b->pbits = pb;
  for (int i = 0; i < ninst; ++i) {
    hi[i].pbase  = parts;
    hi[i].pbits  = pb[i];
    b->pbase[i]  = parts;
    parts += (int64_t)1 << pb[i];
  }
  b->nrows  = rows;
  b->nparts = parts;
  std::vector<Item> ri, qi;
  for (int i : by_size(n)) {
    for (int64_t t = 0; t < (n[i] + b->tile - 1) / b->tile; ++t) ri.push_back({i, (int32_t)t});
    for (int q = 0; q < (1 << pb[i]); ++q) qi.push_back({i, q});
  }
  LDF_EXPECTS(ri.size() < (size_t)INT32_MAX / 2 && qi.size() < (size_t)INT32_MAX / 2, "batch groupby: too many tiles");
  b->nritems = (int)ri.size();
  b->npitems = (int)qi.size();
  b->grid_g  = resident_blocks(kGT, kMaxParts * 12);

  b->d_inst   = b->upload(hi);
  b->d_ritems = b->upload(ri);
  b->d_pitems = b->upload(qi);
  p.pkey      = b->dnew<unsigned long long>(rows);
  for (int c = 0; c < p.ncol; ++c) p.pval[c] = b->dnew<unsigned long long>(rows);
  p.pflag   = b->dnew<uint8_t>(rows);
  p.pcount  = b->dnew<uint32_t>(parts);
  p.pstart  = b->dnew<uint64_t>(parts + 1);
  p.pcursor = b->dnew<unsigned long long>(parts);
  p.gcount  = b->dnew<uint32_t>(parts);
  p.flags   = b->dnew<int32_t>(2);
  b->gpos   = b->dnew<uint64_t>(parts + 1);
  b->ooff   = b->dnew<int64_t>(ninst + 1);
  b->ws     = b->dnew<uint64_t>(scan_workspace_elems(parts));
  b->cursors = b->dnew<uint32_t>(4);
  b->nulls   = b->dnew<unsigned long long>(kMaxBAgg + 1);
  const int64_t staged = parts * (int64_t)p.stage_cap;
  for (int k = 0; k < p.nout; ++k) {
    BOutDev& o = p.o[k];
    o.stage    = b->dalloc((size_t)staged * o.width);
    o.stage_valid = (o.nullable || o.kind == -1) ? b->dnew<uint8_t>(staged) : nullptr;
  }
  if (p.lds_bytes > 64 * 1024)
    LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_bg_aggregate, hipFuncAttributeMaxDynamicSharedMemorySize, p.lds_bytes));

  // sizes: one run without the final gather
  b->out_cap = 0;
  b->enqueue(s);
  int32_t hf[2];
  uint64_t total = 0;
  LDF_HIP_TRY(hipMemcpyAsync(hf, p.flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  LDF_HIP_TRY(hipMemcpyAsync(&total, b->gpos + parts, sizeof(total), hipMemcpyDeviceToHost, s));
  LDF_HIP_TRY(hipStreamSynchronize(s));
  if (hf[1] != 0) {
    // some partition had more groups than its staging room: give the instances holding a full
    // partition four times as many partitions and plan again
    std::vector<uint32_t> gc(parts);
    LDF_HIP_TRY(hipMemcpy(gc.data(), p.gcount, sizeof(uint32_t) * parts, hipMemcpyDeviceToHost));
    std::vector<int32_t> nb2 = pb;
    bool grew                = false;
    for (int i = 0; i < ninst; ++i) {
      bool full = false;
      for (int64_t q = b->pbase[i]; q < b->pbase[i] + ((int64_t)1 << pb[i]); ++q) full |= gc[q] >= (uint32_t)p.stage_cap;
      if (full && (1 << pb[i]) < kMaxParts) {
        nb2[i] = std::min(pb[i] + 2, 12);
        grew   = true;
      }
    }
    if (!grew) throw Error(LDF_E_UNSUPPORTED, "batch groupby: an instance has too many groups for the batched form");
    b->only_cs = true;  // it only ever ran on s
    b.reset();
    return create_groupby(ninst, tables, keys, nkeys, aggs, naggs, cfg, s, &nb2);
  }
  b->out_cap = (int64_t)total;
  b->cols.resize(p.nout);
  for (int k = 0; k < p.nout; ++k) {
    BOutDev& o   = p.o[k];
    ldf_column& c = b->cols[k];
    std::memset(&c, 0, sizeof(c));
    c.type_id    = o.out_type;
    c.data_bytes = b->out_cap * o.width;
    c.data       = b->dalloc((size_t)std::max<int64_t>(1, c.data_bytes));
    o.out        = c.data;
    if (o.nullable) {
      c.validity_bytes = ((b->out_cap + 63) / 64) * 8;
      c.validity       = static_cast<uint8_t*>(b->dalloc((size_t)std::max<int64_t>(8, c.validity_bytes)));
      o.out_valid      = reinterpret_cast<uint64_t*>(c.validity);
    } else {
      o.out_valid = nullptr;
    }
  }
  LDF_HIP_TRY(hipStreamSynchronize(s));
  return b.release();
}

// ---- autotuner ------------------------------------------------------------------------------
// Picks the unset knobs (load factor, tile rows, LDS table slots) by timing candidate plans on a
// sample of the batch's instances (every k-th instance by size, at most 64): each candidate is
// planned, run once to warm up and run twice more under HIP events; the fastest wins.  Choices
// are cached per (operation, log2 instances, log2 mean rows) in memory and, when the
// environment variable LDF_AUTOTUNE_CACHE names a file, in that file ("key lf tile slots" lines).
struct TuneCache {
  std::mutex mu;
  std::map<std::string, ldf_batch_config> m;
  bool loaded = false;
  void load()
  {
    if (loaded) return;
    loaded           = true;
    const char* path = std::getenv("LDF_AUTOTUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char key[128];
    double lf;
    int tile, slots;
    while (std::fscanf(f, "%127s %lf %d %d", key, &lf, &tile, &slots) == 4) {
      ldf_batch_config c{};
      c.load_factor = lf;
      c.tile_rows   = tile;
      c.lds_slots   = slots;
      m[key]        = c;
    }
    std::fclose(f);
  }
  void store(const std::string& key, const ldf_batch_config& c)
  {
    m[key]           = c;
    const char* path = std::getenv("LDF_AUTOTUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %.3f %d %d\n", key.c_str(), c.load_factor, c.tile_rows, c.lds_slots);
    std::fclose(f);
  }
};
TuneCache& tune_cache()
{
  static TuneCache* c = new TuneCache();
  return *c;
}

std::string tune_key(const char* op, int ninst, const std::vector<int64_t>& rows)
{
  double mean = 0;
  for (int64_t r : rows) mean += (double)r;
  mean /= std::max<size_t>(1, rows.size());
  char buf[96];
  std::snprintf(buf, sizeof(buf), "%s:i%d:r%d", op, (int)std::lround(std::log2(std::max(1, ninst))),
                (int)std::lround(std::log2(std::max(1.0, mean))));
  return buf;
}

std::vector<int> tune_sample(const std::vector<int64_t>& rows)
{
  std::vector<int> o = by_size(rows);
  const int want     = std::min<int>(64, (int)o.size());
  std::vector<int> s;
  for (int j = 0; j < want; ++j) s.push_back(o[(size_t)j * o.size() / want]);
  return s;
}

template <typename Make>
ldf_batch_config tune(const ldf_batch_config& asked, const std::vector<ldf_batch_config>& cands, hipStream_t s, Make make)
{
  ldf_batch_config best = cands.front();
  float best_ms         = 1e30f;
  hipEvent_t e0, e1;
  LDF_HIP_TRY(hipEventCreate(&e0));
  LDF_HIP_TRY(hipEventCreate(&e1));
  for (const ldf_batch_config& c : cands) {
    std::unique_ptr<BatchImpl> b(make(c));
    b->cfg.use_graph = 0;
    b->run(s);
    LDF_HIP_TRY(hipEventRecord(e0, s));
    b->run(s);
    b->run(s);
    LDF_HIP_TRY(hipEventRecord(e1, s));
    LDF_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    LDF_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best    = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  best.use_graph = asked.use_graph;
  return best;
}

ldf_batch_config resolve_join_config(const ldf_batch_config* in, int32_t ninst, const ldf_table_view* lhs,
                                     const ldf_table_view* rhs, const int32_t* lk, const int32_t* rk, int32_t jt,
                                     bool ne, const int32_t* lo, int32_t nlo, const int32_t* ro, int32_t nro,
                                     hipStream_t s)
{
  ldf_batch_config c{};
  if (in) c = *in;
  if (c.load_factor != 0.0 && c.tile_rows != 0 && c.lds_slots != 0) return c;
  std::vector<int64_t> rows(ninst);
  for (int i = 0; i < ninst; ++i) rows[i] = lhs[i].ncols > 0 ? lhs[i].cols[0].length : 0;
  const std::string key = tune_key("join", ninst, rows);
  TuneCache& T          = tune_cache();
  {
    std::lock_guard<std::mutex> lk2(T.mu);
    T.load();
    auto it = T.m.find(key);
    if (it != T.m.end()) {
      if (c.load_factor == 0.0) c.load_factor = it->second.load_factor;
      if (c.tile_rows == 0) c.tile_rows = it->second.tile_rows;
      if (c.lds_slots == 0) c.lds_slots = it->second.lds_slots == 1 || it->second.lds_slots == 2048 ? it->second.lds_slots : 1;
      return c;
    }
  }
  // load factor x tile rows with the tables in HBM, then the best of those with the tables of up
  // to 2048 slots probed from LDS
  std::vector<ldf_batch_config> cands;
  for (double lf : {0.5, 0.7})
    for (int t : {1024, 2048, 4096}) {
      ldf_batch_config k = c;
      if (c.load_factor == 0.0) k.load_factor = lf;
      else if (lf != 0.5) continue;
      if (c.tile_rows == 0) k.tile_rows = t;
      else if (t != 1024) continue;
      if (c.lds_slots == 0) k.lds_slots = 1;
      cands.push_back(k);
    }
  const std::vector<int> smp = tune_sample(rows);
  std::vector<ldf_table_view> sl, sr;
  for (int i : smp) {
    sl.push_back(lhs[i]);
    sr.push_back(rhs[i]);
  }
  auto make = [&](const ldf_batch_config& k) -> BatchImpl* {
    return create_join((int32_t)smp.size(), sl.data(), sr.data(), lk, rk, 1, jt, ne, lo, nlo, ro, nro, k, s);
  };
  ldf_batch_config best = tune(c, cands, s, make);
  if (c.lds_slots == 0) {
    ldf_batch_config lds = best;
    lds.lds_slots        = 2048;
    best                 = tune(c, {best, lds}, s, make);
  }
  std::lock_guard<std::mutex> lk2(T.mu);
  T.store(key, best);
  return best;
}

ldf_batch_config resolve_groupby_config(const ldf_batch_config* in, int32_t ninst, const ldf_table_view* tables,
                                        const int32_t* keys, int32_t nkeys, const ldf_agg* aggs, int32_t naggs,
                                        hipStream_t s)
{
  ldf_batch_config c{};
  if (in) c = *in;
  if (c.load_factor == 0.0) c.load_factor = 0.5;  // unused by the groupby batch
  if (c.tile_rows != 0 && c.lds_slots != 0) return c;
  std::vector<int64_t> rows(ninst);
  for (int i = 0; i < ninst; ++i) rows[i] = tables[i].ncols > 0 ? tables[i].cols[0].length : 0;
  const std::string key = tune_key("groupby", ninst, rows);
  TuneCache& T          = tune_cache();
  {
    std::lock_guard<std::mutex> lk2(T.mu);
    T.load();
    auto it = T.m.find(key);
    if (it != T.m.end()) {
      if (c.tile_rows == 0) c.tile_rows = it->second.tile_rows;
      if (c.lds_slots == 0) c.lds_slots = it->second.lds_slots;
      return c;
    }
  }
  std::vector<ldf_batch_config> cands;
  for (int t : {2048, 4096})
    for (int sl : {512, 1024, 2048}) {
      ldf_batch_config k = c;
      if (c.tile_rows == 0) k.tile_rows = t;
      else if (t != 2048) continue;
      if (c.lds_slots == 0) k.lds_slots = sl;
      else if (sl != 512) continue;
      cands.push_back(k);
    }
  const std::vector<int> smp = tune_sample(rows);
  std::vector<ldf_table_view> st;
  for (int i : smp) st.push_back(tables[i]);
  std::vector<ldf_batch_config> ok;
  int first_code = 0;
  std::string first_msg;
  for (auto& k : cands) {  // drop the LDS layouts that do not fit
    try {
      std::unique_ptr<BatchImpl> b(create_groupby(1, st.data(), keys, nkeys, aggs, naggs, k, s));
      ok.push_back(k);
    } catch (const Error& e) {
      if (first_code == 0 || e.code == LDF_E_UNSUPPORTED) {
        first_code = e.code;
        first_msg  = e.what();
      }
    }
  }
  // every layout failed: a limit of the work-queue form (UNSUPPORTED: the caller batches instance
  // by instance), or the request itself
  if (ok.empty()) {
    if (first_code == LDF_E_UNSUPPORTED) throw Error(LDF_E_UNSUPPORTED, first_msg);
    if (first_code != 0) throw Error(first_code, first_msg);
    throw Error(LDF_E_UNSUPPORTED, "batch groupby: no LDS table layout fits these aggregations");
  }
  ldf_batch_config best = tune(c, ok, s, [&](const ldf_batch_config& k) -> BatchImpl* {
    return create_groupby((int32_t)smp.size(), st.data(), keys, nkeys, aggs, naggs, k, s);
  });
  std::lock_guard<std::mutex> lk2(T.mu);
  T.store(key, best);
  return best;
}

// =========================================================================================
// served form
// =========================================================================================
// dst bit j = src bit (off + j), for j < n (whole 64-bit destination words)
__global__ void k_copy_bits(const uint64_t* __restrict__ src, int64_t src_words, int64_t off, int64_t n,
                            uint64_t* __restrict__ dst)
{
  const int64_t nw = (n + 63) / 64;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nw; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b  = off + 64 * j;
    const int64_t w  = b >> 6;
    const int sh     = (int)(b & 63);
    uint64_t lo      = w < src_words ? src[w] : 0;
    uint64_t hi      = (sh != 0 && w + 1 < src_words) ? src[w + 1] : 0;
    uint64_t v       = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
    const int64_t rem = n - 64 * j;
    if (rem < 64) v &= (1ull << rem) - 1;
    dst[j] = v;
  }
}

// Bits [off, off + n) of a validity bitmap of src_bytes bytes (no padding assumed: only those
// bytes are read) into dst starting at bit 0, one 64-bit word per thread.
__global__ void k_slice_bits(const uint8_t* __restrict__ src, int64_t src_bytes, int64_t off, int64_t n,
                             uint64_t* __restrict__ dst)
{
  const int64_t nw = (n + 63) / 64;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nw; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b  = off + 64 * j;
    const int64_t by = b >> 3;
    const int sh     = (int)(b & 7);
    uint64_t v       = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (by + k < src_bytes) v |= (uint64_t)src[by + k] << (8 * k);
    if (sh != 0) {
      v >>= sh;
      if (by + 8 < src_bytes) v |= (uint64_t)src[by + 8] << (64 - sh);
    }
    const int64_t rem = n - 64 * j;
    if (rem < 64) v &= (1ull << rem) - 1;
    dst[j] = v;
  }
}

int64_t now_ns()
{
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
    .count();
}

struct ServedTable {
  std::vector<ldf_column_view> cols;
  int32_t key_col = 0;
  int64_t rows    = 0;
  bool fixed      = true;  // every column fixed-width: the table has a key index and batches
  std::unique_ptr<JoinBatch> index;  // instance 0 builds the table's key index
  JoinIndex idx;
  size_t bytes = 0;
};

struct Query {
  int64_t id = 0;
  ldf_query q{};
  std::vector<ldf_column_view> probe_cols;
  std::vector<ldf_agg> aggs;
  bool plain       = false;
  size_t est_bytes = 0;
  size_t res_bytes = 0;
  int status       = 1;  // 1 waiting or running, 0 done, < 0 failed
  std::string err;
  std::vector<ldf_column> cols;
  ldf_query_stats st{};
};

size_t column_bytes(const ldf_column& c)
{
  size_t b = (size_t)std::max<int64_t>(0, c.data_bytes) + (size_t)std::max<int64_t>(0, c.validity_bytes);
  if (c.type_id == LDF_TYPE_LARGE_STRING) b += (size_t)(c.length + 1) * 8;
  return b;
}

struct Server {
  ldf_server_config cfg{};
  int dev = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> waiting;
  std::map<int64_t, std::unique_ptr<Query>> queries;
  std::vector<std::unique_ptr<ServedTable>> tables;
  double held    = 0;  // bytes: indexes, batches in flight (estimates), unreleased results
  int running    = 0;
  bool stop      = false;
  int64_t next_id = 1;
  std::vector<std::thread> workers;
  std::vector<hipStream_t> streams;

  ~Server()
  {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    for (auto& t : workers) t.join();
    (void)hipSetDevice(dev);
    (void)hipDeviceSynchronize();
    Alloc a(nullptr, nullptr);
    for (auto& kv : queries)
      for (auto& c : kv.second->cols) free_column(&c, a);
    queries.clear();
    tables.clear();
    for (auto s : streams) (void)hipStreamDestroy(s);
    (void)hipGetLastError();
  }

  // estimated device bytes a query needs while it runs and for its result
  size_t estimate(const Query& q) const
  {
    if (q.q.kind == LDF_QUERY_JOIN) {
      const ServedTable& t = *tables[q.q.table];
      int64_t row_bytes    = 0;
      for (auto& c : q.probe_cols) row_bytes += std::max(1, type_width(c.type_id));
      for (auto& c : t.cols) row_bytes += std::max(1, type_width(c.type_id));
      const int64_t np = q.probe_cols.empty() ? 0 : q.probe_cols[0].length;
      return (size_t)np * (size_t)(16 + 2 * (row_bytes + 8)) + (1u << 20);
    }
    const int64_t n = q.q.row_end - q.q.row_begin;
    return (size_t)n * (q.plain ? 96 : 72) + (1u << 20);
  }

  bool same_batch(const Query& a, const Query& b) const
  {
    if (a.q.kind != b.q.kind || a.plain || b.plain) return false;
    if (a.q.kind == LDF_QUERY_JOIN) return true;
    if (a.aggs.size() != b.aggs.size()) return false;
    for (size_t i = 0; i < a.aggs.size(); ++i)
      if (a.aggs[i].in_col != b.aggs[i].in_col || a.aggs[i].kind != b.aggs[i].kind ||
          a.aggs[i].count_mode != b.aggs[i].count_mode)
        return false;
    return true;
  }

  // Takes the next batch: the oldest waiting query (run even over budget when nothing else runs,
  // so a query larger than the budget cannot wait forever) and the later ones that batch with it
  // and still fit.  Called with mu held.
  bool take(std::vector<Query*>& out)
  {
    if (waiting.empty()) return false;
    Query* head = queries[waiting.front()].get();
    if (held + head->est_bytes > cfg.hbm_budget_bytes && running > 0) return false;
    double used = held + head->est_bytes;
    out.push_back(head);
    std::deque<int64_t> rest;
    waiting.pop_front();
    while (!waiting.empty()) {
      const int64_t id = waiting.front();
      waiting.pop_front();
      Query* q = queries[id].get();
      if ((int)out.size() < cfg.max_batch && same_batch(*head, *q) && used + q->est_bytes <= cfg.hbm_budget_bytes) {
        used += q->est_bytes;
        out.push_back(q);
      } else {
        rest.push_back(id);
      }
    }
    waiting.swap(rest);
    held = used;
    return true;
  }

  void worker(int slot)
  {
    (void)hipSetDevice(dev);
    hipStream_t s = streams[slot];
    for (;;) {
      std::vector<Query*> batch;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || take(batch); });
        if (batch.empty()) return;  // stopping
        ++running;
        const int64_t t = now_ns();
        for (Query* q : batch) {
          q->st.start_ns      = t;
          q->st.batch_queries = (int32_t)batch.size();
          q->st.plain         = q->plain ? 1 : 0;
        }
      }
      int rc = LDF_OK;
      std::string msg;
      try {
        execute(batch, s);
      } catch (const Error& e) {
        rc  = e.code;
        msg = e.what();
      } catch (const std::exception& e) {
        rc  = LDF_E_INTERNAL;
        msg = e.what();
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        const int64_t t = now_ns();
        for (Query* q : batch) {
          held -= (double)q->est_bytes;
          if (rc == LDF_OK) {
            held += (double)q->res_bytes;
            q->status = 0;
          } else {
            q->status = rc;
            q->err    = msg;
          }
          q->st.end_ns = t;
        }
        --running;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  // copies rows [lo, lo + n) of batch column c into a result column of the query
  void copy_rows(const ldf_column& c, int64_t lo, int64_t n, ldf_column* o, Alloc& a, hipStream_t s)
  {
    alloc_column(c.type_id, n, c.validity != nullptr, 0, a, o);
    const int w = type_width(c.type_id);
    if (n > 0)
      LDF_HIP_TRY(hipMemcpyAsync(o->data, static_cast<const uint8_t*>(c.data) + lo * w, (size_t)n * w,
                                 hipMemcpyDeviceToDevice, s));
    if (c.validity != nullptr && n > 0) {
      hipLaunchKernelGGL(k_copy_bits, dim3(grid_for((n + 63) / 64, 256)), dim3(256), 0, s,
                         reinterpret_cast<const uint64_t*>(c.validity), (c.validity_bytes + 7) / 8, lo, n,
                         reinterpret_cast<uint64_t*>(o->validity));
      LDF_HIP_TRY(hipGetLastError());
    }
  }

  void execute(std::vector<Query*>& batch, hipStream_t s)
  {
    Alloc a(nullptr, s);
    std::vector<Buf<uint64_t>> keep;  // validity slices of unaligned row ranges
    Query& h = *batch.front();
    if (h.plain && h.q.kind == LDF_QUERY_JOIN) {  // the plain join of the probe with the table
      const ServedTable& t = *tables[h.q.table];
      ldf_table_view lv{(int32_t)h.probe_cols.size(), h.probe_cols.data()};
      ldf_table_view rv{(int32_t)t.cols.size(), t.cols.data()};
      std::vector<int32_t> lo(h.probe_cols.size()), ro(t.cols.size());
      for (size_t i = 0; i < lo.size(); ++i) lo[i] = (int32_t)i;
      for (size_t i = 0; i < ro.size(); ++i) ro[i] = (int32_t)i;
      h.cols.resize(lo.size() + ro.size());
      ldf_table_out out{(int32_t)h.cols.size(), h.cols.data()};
      const int rc = ldf_join(&lv, &rv, &h.q.probe_key, &t.key_col, 1, LDF_JOIN_INNER, 1, lo.data(), (int32_t)lo.size(),
                              ro.data(), (int32_t)ro.size(), nullptr, s, &out);
      if (rc != LDF_OK) throw Error(rc, ldf_last_error());
      LDF_HIP_TRY(hipStreamSynchronize(s));
      h.res_bytes = 0;
      for (auto& c : h.cols) h.res_bytes += column_bytes(c);
      return;
    }
    if (h.plain) {  // groupby: the plain entry point on this stream
      const ServedTable& t = *tables[h.q.table];
      std::vector<ldf_column_view> v = slice(t, h.q.row_begin, h.q.row_end, a, keep);
      ldf_table_view tv{(int32_t)v.size(), v.data()};
      h.cols.resize(1 + h.aggs.size());
      ldf_table_out out{(int32_t)h.cols.size(), h.cols.data()};
      const int rc = ldf_groupby(&tv, &t.key_col, 1, h.aggs.data(), (int32_t)h.aggs.size(), nullptr, s, &out);
      if (rc != LDF_OK) throw Error(rc, ldf_last_error());
      LDF_HIP_TRY(hipStreamSynchronize(s));
      h.res_bytes = 0;
      for (auto& c : h.cols) h.res_bytes += column_bytes(c);
      return;
    }
    const int n = (int)batch.size();
    ldf_batch_config cfg{0.5, 2048, 1024, 0};
    std::unique_ptr<BatchImpl> b;
    std::vector<std::vector<ldf_column_view>> views(n);
    std::vector<ldf_table_view> tv(n);
    if (h.q.kind == LDF_QUERY_JOIN) {
      std::vector<ldf_table_view> rv(n);
      std::vector<JoinIndex> idx(n);
      for (int k = 0; k < n; ++k) {
        const Query& q       = *batch[k];
        const ServedTable& t = *tables[q.q.table];
        tv[k]                = {(int32_t)q.probe_cols.size(), q.probe_cols.data()};
        rv[k]                = {(int32_t)t.cols.size(), t.cols.data()};
        idx[k]               = t.idx;
      }
      const int32_t lk = h.q.probe_key, rk = tables[h.q.table]->key_col;
      std::vector<int32_t> lo(h.probe_cols.size()), ro(tables[h.q.table]->cols.size());
      for (size_t i = 0; i < lo.size(); ++i) lo[i] = (int32_t)i;
      for (size_t i = 0; i < ro.size(); ++i) ro[i] = (int32_t)i;
      JoinBatch* jb = create_join(n, tv.data(), rv.data(), &lk, &rk, 1, LDF_JOIN_INNER, true, lo.data(),
                                  (int32_t)lo.size(), ro.data(), (int32_t)ro.size(), cfg, s, &idx);
      b.reset(jb);
      b->only_cs = true;
      jb->enqueue_tail(s);
    } else {
      for (int k = 0; k < n; ++k) {
        const Query& q = *batch[k];
        views[k]       = slice(*tables[q.q.table], q.q.row_begin, q.q.row_end, a, keep);
        tv[k]          = {(int32_t)views[k].size(), views[k].data()};
      }
      const int32_t key = tables[h.q.table]->key_col;
      GroupbyBatch* gb  = create_groupby(n, tv.data(), &key, 1, h.aggs.data(), (int32_t)h.aggs.size(), cfg, s);
      b.reset(gb);
      b->only_cs = true;
      gb->enqueue_tail(s);
    }
    std::vector<ldf_column> cols(kMaxBOut + kMaxBAgg + 1);
    ldf_table_out out{(int32_t)cols.size(), cols.data()};
    std::vector<int64_t> offs(n + 1);
    b->result(s, &out, offs.data());
    for (int k = 0; k < n; ++k) {
      Query& q = *batch[k];
      q.cols.resize(out.ncols);
      for (int c = 0; c < out.ncols; ++c) copy_rows(cols[c], offs[k], offs[k + 1] - offs[k], &q.cols[c], a, s);
    }
    // null counts of each result column (one small count per column)
    LDF_HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < n; ++k) {
      Query& q    = *batch[k];
      q.res_bytes = 0;
      for (auto& c : q.cols) {
        c.null_count = 0;
        if (c.validity != nullptr && c.length > 0) {
          std::vector<uint64_t> bits((size_t)(c.length + 63) / 64);
          LDF_HIP_TRY(hipMemcpy(bits.data(), c.validity, bits.size() * 8, hipMemcpyDeviceToHost));
          int64_t valid = 0;
          for (uint64_t w : bits) valid += __builtin_popcountll(w);
          c.null_count = c.length - valid;
        }
        q.res_bytes += column_bytes(c);
      }
    }
  }

  // table rows [lo, hi) as views: values by pointer; validity by pointer when lo is a multiple
  // of 64, else as a bit-shifted copy in `keep` (released in the stream's order)
  std::vector<ldf_column_view> slice(const ServedTable& t, int64_t lo, int64_t hi, Alloc& a,
                                     std::vector<Buf<uint64_t>>& keep) const
  {
    std::vector<ldf_column_view> v = t.cols;
    for (auto& c : v) {
      const int w = type_width(c.type_id);
      c.data      = static_cast<const uint8_t*>(c.data) + lo * w;
      if (w == 0) c.offsets += lo;  // strings: the rows' offsets, still into the whole character buffer
      if (c.validity && (lo & 63) == 0) {
        c.validity += lo / 8;
      } else if (c.validity) {
        const int64_t n = hi - lo;
        keep.emplace_back(a, (size_t)std::max<int64_t>(1, (n + 63) / 64));
        if (n > 0) {
          hipLaunchKernelGGL(k_slice_bits, dim3(grid_for((n + 63) / 64, 256)), dim3(256), 0, a.stream(), c.validity,
                             (t.rows + 7) / 8, lo, n, keep.back().get());
          LDF_HIP_TRY(hipGetLastError());
        }
        c.validity = reinterpret_cast<const uint8_t*>(keep.back().get());
      }
      c.length     = hi - lo;
      c.null_count = -1;
    }
    return v;
  }
};

// ---- instances the work-queue kernels do not take --------------------------------------------
// Composite or string keys, string output columns, and the groupby kinds COUNT_DISTINCT,
// PRODUCT, VARIANCE, STDDEV and APPROX_MEDIAN: such a batch runs each instance through the plain
// path (ldf_join / ldf_groupby on the run's stream: their allocations and one host
// synchronisation per instance) and concatenates the instances' output columns into buffers the
// batch owns.  No graph capture; the outputs are the plain calls' rows, instance by instance.

struct ConcatSrc {
  const uint8_t* data;    // values, or chars for strings
  const uint8_t* valid;   // NULL: every row valid
  const int64_t* offs;    // strings: length + 1 offsets
  int64_t cbase;          // strings: the instance's first char in the output
};

// one thread per output row: instance by binary search over the row bases (ninst + 1 entries)
__global__ void k_concat_rows(const ConcatSrc* __restrict__ src, const int64_t* __restrict__ bases, int ninst, int w,
                              int64_t total, uint8_t* __restrict__ dst, uint32_t* __restrict__ dvalid,
                              int64_t* __restrict__ doffs)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < total; r += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = ninst;  // bases[lo] <= r < bases[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (bases[mid] <= r) lo = mid;
      else hi = mid;
    }
    const ConcatSrc c = src[lo];
    const int64_t o   = r - bases[lo];
    if (w > 0) {
      for (int b = 0; b < w; ++b) dst[r * w + b] = c.data[o * w + b];
    } else {
      doffs[r] = c.offs[o] - c.offs[0] + c.cbase;
      if (r == total - 1) doffs[total] = c.offs[o + 1] - c.offs[0] + c.cbase;
    }
    if (dvalid != nullptr) {
      const uint32_t bit = c.valid == nullptr ? 1u : ((uint32_t)c.valid[o >> 3] >> (uint32_t)(o & 7)) & 1u;
      if (bit) atomicOr(&dvalid[r >> 5], 1u << (uint32_t)(r & 31));
    }
  }
}

struct BatchPlain : BatchImpl {
  bool is_join = false;
  int32_t ninst = 0, join_type = 0, nulls_equal = 0;
  std::vector<std::vector<ldf_column_view>> lcols, rcols;  // copies of the instances' views
  std::vector<int32_t> lkeys, rkeys, lout, rout;
  std::vector<ldf_agg> aggs;
  int ncols_out = 0;
  // the last run's concatenated columns (batch-owned, freed by the next run or destroy)
  std::vector<ldf_column> outs;
  std::vector<int64_t> bases;
  std::vector<std::pair<void*, size_t>> outbufs;
  hipStream_t out_stream = nullptr;

  ~BatchPlain() override { free_outputs(); }

  void free_outputs()
  {
    if (outbufs.empty()) return;
    Alloc a(nullptr, out_stream);
    for (auto& o : outbufs) a.deallocate(o.first, o.second);
    outbufs.clear();
  }
  void* out_alloc(hipStream_t s, size_t bytes)
  {
    Alloc a(nullptr, s);
    void* p = a.allocate(bytes > 0 ? bytes : 1);
    outbufs.push_back({p, bytes > 0 ? bytes : 1});
    return p;
  }

  void enqueue(hipStream_t s) override
  {
    free_outputs();
    out_stream = s;
    outs.assign(ncols_out, ldf_column{});
    bases.assign(ninst + 1, 0);
    std::vector<std::vector<ldf_column>> per(ninst, std::vector<ldf_column>(std::max(1, ncols_out)));
    auto free_all = [&](int upto) {
      for (int i = 0; i < upto; ++i) {
        ldf_table_out t{ncols_out, per[i].data()};
        ldf_free_table(&t, nullptr, s);
      }
    };
    for (int i = 0; i < ninst; ++i) {
      ldf_table_out t{ncols_out, per[i].data()};
      int rc;
      if (is_join) {
        const ldf_table_view lv{(int32_t)lcols[i].size(), lcols[i].data()};
        const ldf_table_view rv{(int32_t)rcols[i].size(), rcols[i].data()};
        rc = ldf_join(&lv, &rv, lkeys.data(), rkeys.data(), (int32_t)lkeys.size(), join_type, nulls_equal,
                      lout.data(), (int32_t)lout.size(), rout.data(), (int32_t)rout.size(), nullptr, s, &t);
      } else {
        const ldf_table_view tv{(int32_t)lcols[i].size(), lcols[i].data()};
        rc = ldf_groupby(&tv, lkeys.data(), (int32_t)lkeys.size(), aggs.data(), (int32_t)aggs.size(), nullptr, s, &t);
      }
      if (rc != LDF_OK) {
        const std::string msg = std::string("batch instance ") + std::to_string(i) + ": " + ldf_last_error();
        free_all(i);
        throw Error(rc, msg);
      }
      LDF_EXPECTS(t.ncols == ncols_out, "batch: an instance returned another column count");
      bases[i + 1] = bases[i] + (ncols_out > 0 ? per[i][0].length : 0);
    }
    const int64_t total = bases[ninst];
    int64_t* dbases     = static_cast<int64_t*>(out_alloc(s, (size_t)(ninst + 1) * 8));
    LDF_HIP_TRY(hipMemcpyAsync(dbases, bases.data(), (size_t)(ninst + 1) * 8, hipMemcpyHostToDevice, s));
    std::vector<ConcatSrc> hsrc(ninst);
    ConcatSrc* dsrc = static_cast<ConcatSrc*>(out_alloc(s, (size_t)ninst * sizeof(ConcatSrc)));
    for (int c = 0; c < ncols_out; ++c) {
      ldf_column& o    = outs[c];
      const int32_t ty = per[0][c].type_id;
      const bool str   = ty == LDF_TYPE_LARGE_STRING;
      const int w      = str ? 0 : type_width(ty);
      bool nullable    = false;
      int64_t nulls = 0, chars = 0;
      std::vector<int64_t> first(ninst, 0);  // strings: each instance's first offset
      for (int i = 0; i < ninst; ++i) {
        const ldf_column& pc = per[i][c];
        LDF_EXPECTS(pc.type_id == ty, "batch: instances returned different column types");
        nullable |= pc.validity != nullptr;
        nulls += pc.null_count > 0 ? pc.null_count : 0;
        hsrc[i] = ConcatSrc{static_cast<const uint8_t*>(pc.data), pc.validity, pc.offsets, chars};
        if (str && pc.length > 0) {
          int64_t ends[2] = {0, 0};  // the instance's first and last offsets
          LDF_HIP_TRY(hipMemcpyAsync(&ends[0], pc.offsets, 8, hipMemcpyDeviceToHost, s));
          LDF_HIP_TRY(hipMemcpyAsync(&ends[1], pc.offsets + pc.length, 8, hipMemcpyDeviceToHost, s));
          LDF_HIP_TRY(hipStreamSynchronize(s));
          first[i] = ends[0];
          chars += ends[1] - ends[0];
        }
      }
      o.type_id        = ty;
      o.length         = total;
      o.null_count     = nullable ? nulls : 0;
      o.data_bytes     = str ? chars : total * w;
      o.data           = out_alloc(s, (size_t)std::max<int64_t>(1, o.data_bytes));
      o.validity_bytes = nullable ? ((total + 63) / 64) * 8 : 0;
      o.validity       = nullable ? static_cast<uint8_t*>(out_alloc(s, (size_t)std::max<int64_t>(8, o.validity_bytes)))
                                  : nullptr;
      o.offsets        = str ? static_cast<int64_t*>(out_alloc(s, (size_t)(total + 1) * 8)) : nullptr;
      if (nullable) LDF_HIP_TRY(hipMemsetAsync(o.validity, 0, (size_t)std::max<int64_t>(8, o.validity_bytes), s));
      if (str) {
        if (total == 0) LDF_HIP_TRY(hipMemsetAsync(o.offsets, 0, 8, s));
        for (int i = 0; i < ninst; ++i) {  // characters: one copy per instance
          const ldf_column& pc = per[i][c];
          if (pc.length == 0) continue;
          const int64_t nb = (i + 1 < ninst ? hsrc[i + 1].cbase : chars) - hsrc[i].cbase;
          if (nb <= 0) continue;
          LDF_HIP_TRY(hipMemcpyAsync(static_cast<uint8_t*>(o.data) + hsrc[i].cbase,
                                     static_cast<const uint8_t*>(pc.data) + first[i], (size_t)nb,
                                     hipMemcpyDeviceToDevice, s));
        }
      }
      if (total > 0) {
        LDF_HIP_TRY(hipMemcpyAsync(dsrc, hsrc.data(), (size_t)ninst * sizeof(ConcatSrc), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_concat_rows, dim3(grid_for(total, 256)), dim3(256), 0, s, (const ConcatSrc*)dsrc,
                           (const int64_t*)dbases, ninst, w, total, static_cast<uint8_t*>(o.data),
                           reinterpret_cast<uint32_t*>(o.validity), o.offsets);
        LDF_HIP_TRY(hipGetLastError());
        // the source descriptors are rewritten for the next column only after this one's rows
        LDF_HIP_TRY(hipStreamSynchronize(s));
      }
    }
    free_all(ninst);
  }

  void result(hipStream_t s, ldf_table_out* out, int64_t* offsets) override
  {
    LDF_EXPECTS(out != nullptr && out->cols != nullptr, "batch: NULL output table");
    LDF_EXPECTS(out->ncols >= ncols_out, "batch: output table has too few columns");
    LDF_HIP_TRY(hipStreamSynchronize(s));
    LDF_EXPECTS((int)outs.size() == ncols_out, "batch: result before the first run");
    out->ncols = ncols_out;
    for (int c = 0; c < ncols_out; ++c) out->cols[c] = outs[c];
    if (offsets != nullptr)
      for (int i = 0; i <= ninst; ++i) offsets[i] = bases[i];
  }
};

std::vector<ldf_column_view> copy_view(const ldf_table_view& t)
{
  LDF_EXPECTS(t.ncols >= 0 && (t.ncols == 0 || t.cols != nullptr), "batch: NULL table columns");
  return std::vector<ldf_column_view>(t.cols, t.cols + t.ncols);
}

BatchImpl* create_plain_join(int32_t ninst, const ldf_table_view* lhs, const ldf_table_view* rhs,
                             const int32_t* lhs_keys, const int32_t* rhs_keys, int32_t nkeys, int32_t join_type,
                             bool nulls_equal, const int32_t* lhs_out, int32_t n_lhs_out, const int32_t* rhs_out,
                             int32_t n_rhs_out, hipStream_t s)
{
  std::unique_ptr<BatchPlain> b(new BatchPlain());
  b->is_join     = true;
  b->ninst       = ninst;
  b->join_type   = join_type;
  b->nulls_equal = nulls_equal ? 1 : 0;
  b->cs          = s;
  for (int i = 0; i < ninst; ++i) {
    b->lcols.push_back(copy_view(lhs[i]));
    b->rcols.push_back(copy_view(rhs[i]));
  }
  b->lkeys.assign(lhs_keys, lhs_keys + nkeys);
  b->rkeys.assign(rhs_keys, rhs_keys + nkeys);
  if (n_lhs_out > 0) b->lout.assign(lhs_out, lhs_out + n_lhs_out);
  if (n_rhs_out > 0) b->rout.assign(rhs_out, rhs_out + n_rhs_out);
  b->ncols_out = n_lhs_out + n_rhs_out;
  std::memset(&b->cfg, 0, sizeof(b->cfg));
  return b.release();
}

BatchImpl* create_plain_groupby(int32_t ninst, const ldf_table_view* tables, const int32_t* keys, int32_t nkeys,
                                const ldf_agg* aggs, int32_t naggs, hipStream_t s)
{
  std::unique_ptr<BatchPlain> b(new BatchPlain());
  b->ninst = ninst;
  b->cs    = s;
  for (int i = 0; i < ninst; ++i) b->lcols.push_back(copy_view(tables[i]));
  b->lkeys.assign(keys, keys + nkeys);
  if (naggs > 0) b->aggs.assign(aggs, aggs + naggs);
  b->ncols_out = nkeys + naggs;
  std::memset(&b->cfg, 0, sizeof(b->cfg));
  return b.release();
}

}  // namespace
}  // namespace ldf

using namespace ldf;

struct ldf_server {
  std::unique_ptr<Server> impl;
};

extern "C" int ldf_server_create(const ldf_server_config* cfg, ldf_server** server)
{
  return guarded([&] {
    LDF_EXPECTS(cfg != nullptr && server != nullptr, "server: NULL argument");
    LDF_EXPECTS(cfg->hbm_budget_bytes > 0, "server: hbm_budget_bytes must be positive");
    LDF_EXPECTS(cfg->slots >= 1 && cfg->slots <= 16, "server: slots must be in 1..16");
    LDF_EXPECTS(cfg->max_batch >= 1, "server: max_batch must be positive");
    *server = nullptr;
    std::unique_ptr<ldf_server> h(new ldf_server());
    h->impl.reset(new Server());
    Server& S = *h->impl;
    S.cfg     = *cfg;
    LDF_HIP_TRY(hipGetDevice(&S.dev));
    for (int i = 0; i < cfg->slots; ++i) {
      hipStream_t st;
      LDF_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      S.streams.push_back(st);
    }
    for (int i = 0; i < cfg->slots; ++i) S.workers.emplace_back([&S, i] { S.worker(i); });
    *server = h.release();
  });
}

extern "C" int ldf_server_add_table(ldf_server* server, const ldf_table_view* table, int32_t key_col,
                                    int32_t* table_id)
{
  return guarded([&] {
    LDF_EXPECTS(server != nullptr && table != nullptr && table_id != nullptr, "server: NULL argument");
    Server& S = *server->impl;
    LDF_EXPECTS(key_col >= 0 && key_col < table->ncols, "server: key column out of range");
    std::unique_ptr<ServedTable> t(new ServedTable());
    t->cols.assign(table->cols, table->cols + table->ncols);
    t->key_col = key_col;
    t->rows    = table->ncols > 0 ? table->cols[0].length : 0;
    for (auto& c : t->cols) {
      validate_view(c, t->rows, "server table");
      if (type_width(c.type_id) <= 0) t->fixed = false;  // its queries run through the plain calls
    }
    {
      std::lock_guard<std::mutex> lk(S.mu);
      if (!S.tables.empty()) {
        const ServedTable& t0 = *S.tables.front();
        LDF_EXPECTS(t0.cols.size() == t->cols.size() && t0.key_col == key_col,
                    "server: every table has the schema of the first");
        for (size_t c = 0; c < t->cols.size(); ++c)
          LDF_EXPECTS(t0.cols[c].type_id == t->cols[c].type_id, "server: every table has the schema of the first");
      }
    }
    if (!t->fixed) {  // no key index: joins with this table run through ldf_join
      std::lock_guard<std::mutex> lk(S.mu);
      *table_id = (int32_t)S.tables.size();
      S.tables.push_back(std::move(t));
      return;
    }
    // the key index: a join batch of one instance with an empty probe side
    hipStream_t s = S.streams[0];
    ldf_column_view empty = t->cols[key_col];
    empty.length          = 0;
    empty.validity        = nullptr;
    ldf_table_view lt{1, &empty};
    ldf_table_view rt{(int32_t)t->cols.size(), t->cols.data()};
    const int32_t zero = 0;
    ldf_batch_config cfg{0.5, 2048, 1024, 0};
    t->index.reset(create_join(1, &lt, &rt, &zero, &key_col, 1, LDF_JOIN_INNER, true, nullptr, 0, nullptr, 0, cfg, s));
    LDF_HIP_TRY(hipStreamSynchronize(s));
    t->idx.tab  = t->index->tab;
    t->idx.rows = t->index->rows;
    t->index->keep_only({t->index->tab, t->index->rows});
    uint64_t cap = next_pow2((uint64_t)std::max<double>(64.0, std::ceil((double)t->rows / cfg.load_factor)));
    t->idx.cap  = cap;
    for (auto& o : t->index->owned) t->bytes += o.second;
    std::lock_guard<std::mutex> lk(S.mu);
    LDF_EXPECTS(S.held + (double)t->bytes <= S.cfg.hbm_budget_bytes, "server: the table index exceeds the HBM budget");
    S.held += (double)t->bytes;
    *table_id = (int32_t)S.tables.size();
    S.tables.push_back(std::move(t));
  });
}

extern "C" int ldf_server_submit(ldf_server* server, const ldf_query* query, int64_t* query_id)
{
  return guarded([&] {
    LDF_EXPECTS(server != nullptr && query != nullptr && query_id != nullptr, "server: NULL argument");
    Server& S = *server->impl;
    std::unique_ptr<Query> q(new Query());
    q->q            = *query;
    q->st.submit_ns = now_ns();
    std::lock_guard<std::mutex> lk(S.mu);
    LDF_EXPECTS(query->table >= 0 && query->table < (int32_t)S.tables.size(), "server: unknown table");
    const ServedTable& t = *S.tables[query->table];
    if (query->kind == LDF_QUERY_JOIN) {
      LDF_EXPECTS(query->probe.ncols >= 1 && query->probe.cols != nullptr, "server: a join query needs probe columns");
      LDF_EXPECTS(query->probe_key >= 0 && query->probe_key < query->probe.ncols, "server: probe key out of range");
      q->probe_cols.assign(query->probe.cols, query->probe.cols + query->probe.ncols);
      const int64_t n = q->probe_cols[0].length;
      q->plain        = !t.fixed;  // string columns: the plain join, alone
      for (auto& c : q->probe_cols) {
        validate_view(c, n, "server probe");
        if (type_width(c.type_id) <= 0) q->plain = true;
      }
      LDF_EXPECTS(q->probe_cols[query->probe_key].type_id == t.cols[t.key_col].type_id,
                  "join: lhs and rhs key column types differ");
      for (auto& kv : S.queries) {  // one probe schema per server, so that joins batch together
        const Query& o = *kv.second;
        if (o.q.kind != LDF_QUERY_JOIN) continue;
        LDF_EXPECTS(o.probe_cols.size() == q->probe_cols.size() && o.q.probe_key == query->probe_key,
                    "server: every join query has the probe schema of the first");
        for (size_t c = 0; c < q->probe_cols.size(); ++c)
          LDF_EXPECTS(o.probe_cols[c].type_id == q->probe_cols[c].type_id,
                      "server: every join query has the probe schema of the first");
        break;
      }
    } else if (query->kind == LDF_QUERY_GROUPBY) {
      LDF_EXPECTS(query->row_begin >= 0 && query->row_begin <= query->row_end && query->row_end <= t.rows,
                  "server: groupby rows out of range");
      LDF_EXPECTS(query->naggs >= 0 && query->naggs <= kMaxBAgg && (query->naggs == 0 || query->aggs != nullptr),
                  "server: 0..16 aggregations");
      q->aggs.assign(query->aggs, query->aggs + query->naggs);
      // large row ranges, tables with string columns and the kinds the batched groupby does not
      // take run through the plain groupby, alone
      q->plain = (query->row_end - query->row_begin) > (int64_t)1 << 21 || !t.fixed;
      for (auto& g : q->aggs) {
        if (g.kind != LDF_AGG_SUM && g.kind != LDF_AGG_MEAN && g.kind != LDF_AGG_COUNT && g.kind != LDF_AGG_COUNT_ALL &&
            g.kind != LDF_AGG_MIN && g.kind != LDF_AGG_MAX)
          q->plain = true;
        LDF_EXPECTS(g.kind == LDF_AGG_COUNT_ALL || (g.in_col >= 0 && g.in_col < (int32_t)t.cols.size()),
                    "groupby: aggregation column index out of range");
      }
    } else {
      throw Error(LDF_E_INVALID, "server: unknown query kind");
    }
    q->q.probe.cols = nullptr;
    q->q.aggs       = nullptr;
    q->est_bytes    = S.estimate(*q);
    q->id           = S.next_id++;
    *query_id       = q->id;
    S.waiting.push_back(q->id);
    S.queries[q->id] = std::move(q);
    S.cv_work.notify_one();
  });
}

extern "C" int ldf_server_wait(ldf_server* server, int64_t query_id, ldf_table_out* out, ldf_query_stats* stats)
{
  return guarded([&] {
    LDF_EXPECTS(server != nullptr, "server: NULL argument");
    Server& S = *server->impl;
    std::unique_lock<std::mutex> lk(S.mu);
    auto it = S.queries.find(query_id);
    LDF_EXPECTS(it != S.queries.end(), "server: unknown query id");
    Query* q = it->second.get();
    S.cv_done.wait(lk, [&] { return q->status != 1; });
    if (stats) *stats = q->st;
    if (q->status < 0) throw Error(q->status, q->err);
    if (out) {
      LDF_EXPECTS(out->cols != nullptr && out->ncols >= (int32_t)q->cols.size(), "server: output table too small");
      out->ncols = (int32_t)q->cols.size();
      for (size_t c = 0; c < q->cols.size(); ++c) out->cols[c] = q->cols[c];
    }
  });
}

extern "C" int ldf_server_release(ldf_server* server, int64_t query_id)
{
  return guarded([&] {
    LDF_EXPECTS(server != nullptr, "server: NULL argument");
    Server& S = *server->impl;
    std::unique_ptr<Query> q;
    {
      std::unique_lock<std::mutex> lk(S.mu);
      auto it = S.queries.find(query_id);
      LDF_EXPECTS(it != S.queries.end(), "server: unknown query id");
      LDF_EXPECTS(it->second->status != 1, "server: the query has not finished");
      q = std::move(it->second);
      S.queries.erase(it);
      S.held -= (double)q->res_bytes;
    }
    Alloc a(nullptr, nullptr);
    for (auto& c : q->cols) free_column(&c, a);
    S.cv_work.notify_all();
  });
}

extern "C" void ldf_server_destroy(ldf_server* server) { delete server; }

struct ldf_batch {
  std::unique_ptr<BatchImpl> impl;
};

extern "C" int ldf_batch_join_create(int32_t ninst, const ldf_table_view* lhs, const ldf_table_view* rhs,
                                     const int32_t* lhs_keys, const int32_t* rhs_keys, int32_t nkeys, int32_t join_type,
                                     int32_t nulls_equal, const int32_t* lhs_out, int32_t n_lhs_out,
                                     const int32_t* rhs_out, int32_t n_rhs_out, const ldf_batch_config* cfg,
                                     void* stream, ldf_batch** batch)
{
  return guarded([&] {
    LDF_EXPECTS(batch != nullptr, "batch: NULL output handle");
    *batch             = nullptr;
    hipStream_t s      = (hipStream_t)stream;
    const bool ne      = nulls_equal != 0;
    LDF_EXPECTS(ninst >= 1 && lhs != nullptr && rhs != nullptr && lhs_keys && rhs_keys && nkeys >= 1,
                "batch join: needs >= 1 instance and a key column");
    std::unique_ptr<ldf_batch> b(new ldf_batch());
    if (nkeys == 1) {
      try {  // the work-queue form; a schema it does not take runs instance by instance
        ldf_batch_config c = resolve_join_config(cfg, ninst, lhs, rhs, lhs_keys, rhs_keys, join_type, ne, lhs_out,
                                                 n_lhs_out, rhs_out, n_rhs_out, s);
        b->impl.reset(create_join(ninst, lhs, rhs, lhs_keys, rhs_keys, nkeys, join_type, ne, lhs_out, n_lhs_out,
                                  rhs_out, n_rhs_out, c, s));
      } catch (const Error& e) {
        if (e.code != LDF_E_UNSUPPORTED) throw;
      }
    }
    if (!b->impl)
      b->impl.reset(create_plain_join(ninst, lhs, rhs, lhs_keys, rhs_keys, nkeys, join_type, ne, lhs_out, n_lhs_out,
                                      rhs_out, n_rhs_out, s));
    *batch = b.release();
  });
}

extern "C" int ldf_batch_groupby_create(int32_t ninst, const ldf_table_view* tables, const int32_t* keys,
                                        int32_t nkeys, const ldf_agg* aggs, int32_t naggs,
                                        const ldf_batch_config* cfg, void* stream, ldf_batch** batch)
{
  return guarded([&] {
    LDF_EXPECTS(batch != nullptr, "batch: NULL output handle");
    *batch        = nullptr;
    hipStream_t s = (hipStream_t)stream;
    LDF_EXPECTS(ninst >= 1 && tables != nullptr && keys != nullptr && nkeys >= 1,
                "batch groupby: needs >= 1 instance and a key column");
    std::unique_ptr<ldf_batch> b(new ldf_batch());
    if (nkeys == 1) {
      try {  // the work-queue form; a schema it does not take runs instance by instance
        ldf_batch_config c = resolve_groupby_config(cfg, ninst, tables, keys, nkeys, aggs, naggs, s);
        b->impl.reset(create_groupby(ninst, tables, keys, nkeys, aggs, naggs, c, s));
      } catch (const Error& e) {
        if (e.code != LDF_E_UNSUPPORTED) throw;
      }
    }
    if (!b->impl) b->impl.reset(create_plain_groupby(ninst, tables, keys, nkeys, aggs, naggs, s));
    *batch = b.release();
  });
}

extern "C" int ldf_batch_run(ldf_batch* batch, void* stream)
{
  return guarded([&] {
    LDF_EXPECTS(batch != nullptr && batch->impl, "batch: NULL handle");
    ProfScope prof("batch_run", (hipStream_t)stream);
    batch->impl->run((hipStream_t)stream);
  });
}

extern "C" int ldf_batch_result(ldf_batch* batch, void* stream, ldf_table_out* out, int64_t* row_offsets)
{
  return guarded([&] {
    LDF_EXPECTS(batch != nullptr && batch->impl, "batch: NULL handle");
    batch->impl->result((hipStream_t)stream, out, row_offsets);
  });
}

extern "C" int ldf_batch_config_used(const ldf_batch* batch, ldf_batch_config* cfg)
{
  return guarded([&] {
    LDF_EXPECTS(batch != nullptr && batch->impl && cfg != nullptr, "batch: NULL handle");
    *cfg = batch->impl->cfg;
  });
}

extern "C" void ldf_batch_destroy(ldf_batch* batch) { delete batch; }
