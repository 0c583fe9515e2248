// Hash join (INNER / LEFT / FULL / SEMI / ANTI) of lhs (probe side) with rhs (build side).
//
// Reference: JoinTask::gpu_variant (cpp/src/join.cu:187-244) -> cudf_join builds
// cudf::hash_join on the rhs keys and probes with the lhs keys (cpp/src/join.cu:32-74); the
// gather maps then gather lhs_out / rhs_out with NULLIFY on the side that may be unmatched
// (cpp/src/join.cu:79-155).  Null keys: nulls_equal -> null matches null, otherwise a row with
// a null key never matches (the CPU variant emulates this with coalesce + mask key columns,
// cpp/src/join.cpp:117-209).  Floating-point keys compare bitwise, as Arrow's hash join does.
//
// Layout of the build side in HBM (a multimap in CSR form, so a key repeated 10^5 times in
// the build table costs one slot, not a 10^5-long probe chain):
//   table   : open addressing, linear probing, 16-byte slots {key, cnt, start};
//             key = the 64-bit key bits (single fixed-width key column) or
//             (fingerprint32 << 32 | representative build row) for composite / string keys,
//             installed with one 64-bit atomicCAS.  Two extra slots hold the key whose bits
//             equal the empty marker and the null key.
//   rows    : build row ids grouped by slot; slot s owns rows[start, start + cnt).
// Probe: k_join_count finds each lhs row's slot once and records {start, cnt} (8 B/row) so
// the emit pass streams instead of re-probing; per-block output counts are scanned; the emit
// kernel expands every block's rows over all its threads (binary search over the block's row
// offsets in LDS), so one hot key with 10^5 matches does not serialise on one thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ldf_common.hpp"
#include "ldf_device.hpp"

namespace ldf {

struct alignas(16) JSlot {
  unsigned long long key;
  uint32_t cnt;
  uint32_t start;
};
constexpr unsigned long long kEmpty = ~0ull;

constexpr int kJBlock = 256;
constexpr int kJItems = 8;
constexpr int kJRows  = kJBlock * kJItems;  // rows per block in the probe passes

__global__ void k_jslot_fill(JSlot* __restrict__ tab, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    tab[i].key   = kEmpty;
    tab[i].cnt   = 0;
    tab[i].start = 0;
  }
}

// ---- slot lookup ------------------------------------------------------------------------
// Linear probing wraps inside a tile of tmask + 1 slots (the whole table when tmask = mask), so
// a tile can be built on its own in LDS.  Probe sequences stop after tmask + 1 slots: a full tile
// returns kTileFull from the build (the caller rebuilds with one tile) and "absent" from a probe.
constexpr int64_t kTileFull = -2;
__device__ __forceinline__ uint64_t tile_next(uint64_t i, uint64_t tmask) { return (i & ~tmask) | ((i + 1) & tmask); }

template <bool GENERIC>
__device__ __forceinline__ int64_t build_find_or_insert(const DKeys& bk, int64_t r, JSlot* tab, uint64_t mask,
                                                        uint64_t tmask, uint64_t cap, bool nulls_equal)
{
  if (!GENERIC) {
    const DCol& c = bk.c[0];
    if (!valid_at(c.valid, r)) return nulls_equal ? (int64_t)cap + 1 : -1;
    unsigned long long k = load_bits(c.data, c.width, r);
    if (k == kEmpty) return (int64_t)cap;
    uint64_t i = mix64(k) & mask;
    for (uint64_t probes = 0; probes <= tmask; ++probes) {
      unsigned long long cur = atomic_load_relaxed(&tab[i].key);
      if (cur == k) return (int64_t)i;
      if (cur == kEmpty) {
        unsigned long long old = atomicCAS(&tab[i].key, kEmpty, k);
        if (old == kEmpty || old == k) return (int64_t)i;
      }
      i = tile_next(i, tmask);
    }
    return kTileFull;
  } else {
    if (!nulls_equal && any_key_null(bk, r)) return -1;
    uint64_t h                = row_hash64(bk, r);
    uint32_t fp               = (uint32_t)(h >> 32);
    unsigned long long packed = ((unsigned long long)fp << 32) | (uint32_t)r;
    uint64_t i                = h & mask;
    for (uint64_t probes = 0; probes <= tmask; ++probes) {
      unsigned long long cur = atomic_load_relaxed(&tab[i].key);
      if (cur == kEmpty) {
        unsigned long long old = atomicCAS(&tab[i].key, kEmpty, packed);
        if (old == kEmpty) return (int64_t)i;
        cur = old;
      }
      if ((uint32_t)(cur >> 32) == fp && rows_equal(bk, r, bk, (int64_t)(uint32_t)cur)) return (int64_t)i;
      i = tile_next(i, tmask);
    }
    return kTileFull;
  }
}

// Returns the slot of the lhs row's key, or -1 when no build row has that key.
template <bool GENERIC>
__device__ __forceinline__ int64_t probe_find(const DKeys& pk, const DKeys& bk, int64_t r, const JSlot* tab,
                                              uint64_t mask, uint64_t tmask, uint64_t cap, bool nulls_equal)
{
  if (!GENERIC) {
    const DCol& c = pk.c[0];
    if (!valid_at(c.valid, r)) return nulls_equal ? (int64_t)cap + 1 : -1;
    unsigned long long k = load_bits(c.data, c.width, r);
    if (k == kEmpty) return (int64_t)cap;
    uint64_t i = mix64(k) & mask;
    for (uint64_t probes = 0; probes <= tmask; ++probes) {
      unsigned long long cur = tab[i].key;
      if (cur == k) return (int64_t)i;
      if (cur == kEmpty) return -1;
      i = tile_next(i, tmask);
    }
    return -1;
  } else {
    if (!nulls_equal && any_key_null(pk, r)) return -1;
    uint64_t h  = row_hash64(pk, r);
    uint32_t fp = (uint32_t)(h >> 32);
    uint64_t i  = h & mask;
    for (uint64_t probes = 0; probes <= tmask; ++probes) {
      unsigned long long cur = tab[i].key;
      if (cur == kEmpty) return -1;
      if ((uint32_t)(cur >> 32) == fp && rows_equal(pk, r, bk, (int64_t)(uint32_t)cur)) return (int64_t)i;
      i = tile_next(i, tmask);
    }
    return -1;
  }
}

// ---- build ------------------------------------------------------------------------------
template <bool GENERIC>
__global__ __launch_bounds__(256) void k_join_build(DKeys bk, int64_t nb, JSlot* __restrict__ tab, uint64_t mask,
                                                    uint64_t tmask, uint64_t cap, bool nulls_equal,
                                                    int32_t* __restrict__ slot_of_row, int32_t* __restrict__ err)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nb; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = build_find_or_insert<GENERIC>(bk, r, tab, mask, tmask, cap, nulls_equal);
    if (s == kTileFull) {
      *err = 1;
      s    = -1;
    }
    if (s >= 0) atomicAdd(&tab[s].cnt, 1u);
    slot_of_row[r] = (int32_t)s;
  }
}

__global__ void k_jslot_counts(const JSlot* __restrict__ tab, int64_t n, uint32_t* __restrict__ cnt)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    cnt[i] = tab[i].cnt;
}

__global__ void k_jslot_starts(JSlot* __restrict__ tab, int64_t n, const uint64_t* __restrict__ start)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    tab[i].start = (uint32_t)start[i];
}

__global__ void k_join_group_rows(const int32_t* __restrict__ slot_of_row, int64_t nb, const JSlot* __restrict__ tab,
                                  uint32_t* __restrict__ fill, uint32_t* __restrict__ rows)
{
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nb; r += (int64_t)gridDim.x * blockDim.x) {
    int32_t s = slot_of_row[r];
    if (s < 0) continue;
    uint32_t pos = tab[s].start + atomicAdd(&fill[s], 1u);
    rows[pos]    = (uint32_t)r;
  }
}

// ---- tiled build (single fixed-width key) -----------------------------------------------
// Build rows are grouped by the tile of their key's home slot (plus two bins: the key equal to
// the empty marker, and the null key when nulls are equal): per-block histograms, a scan, and a
// scatter of 16-byte (key, row) records.  Then one block per tile inserts its rows' keys into an
// LDS copy of the tile with LDS atomics, counts the rows per slot, writes the finished slots
// {key, cnt, start} and places its rows grouped by slot -- in the same row range the grouping
// gave the tile, which is exactly where the CSR layout wants them.
constexpr int kBTile    = 4096;   // slots per tile
constexpr int kBMaxBins = 32770;  // tiles + 2 special bins (a table of up to 2^27 slots)
constexpr int kBHist    = 1024;   // threads of the histogram / scatter blocks

__device__ __forceinline__ int jb_bin(const DCol& c, int64_t r, uint64_t mask, int tbits, int ntiles, bool nulls_equal,
                                      unsigned long long& k)
{
  if (!valid_at(c.valid, r)) return nulls_equal ? ntiles + 1 : -1;
  k = load_bits(c.data, c.width, r);
  if (k == kEmpty) return ntiles;
  return (int)((mix64(k) & mask) >> tbits);
}

__global__ __launch_bounds__(kBHist) void k_jb_hist(DCol c, int64_t nb, int64_t rpb, uint64_t mask, int tbits,
                                                    int ntiles, bool nulls_equal, uint32_t* __restrict__ counts)
{
  extern __shared__ uint32_t hist[];
  const int nbins = ntiles + 2;
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const int64_t begin = (int64_t)blockIdx.x * rpb;
  const int64_t end   = begin + rpb < nb ? begin + rpb : nb;
  for (int64_t r = begin + threadIdx.x; r < end; r += blockDim.x) {
    unsigned long long k;
    const int bin = jb_bin(c, r, mask, tbits, ntiles, nulls_equal, k);
    if (bin >= 0) atomicAdd(&hist[bin], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) counts[(int64_t)i * gridDim.x + blockIdx.x] = hist[i];
}

__global__ __launch_bounds__(kBHist) void k_jb_scatter(DCol c, int64_t nb, int64_t rpb, uint64_t mask, int tbits,
                                                       int ntiles, bool nulls_equal, const uint64_t* __restrict__ offs,
                                                       uint4* __restrict__ recs, int32_t* __restrict__ slot_of_row)
{
  extern __shared__ uint32_t cur[];
  const int nbins = ntiles + 2;
  for (int i = threadIdx.x; i < nbins; i += blockDim.x) cur[i] = (uint32_t)offs[(int64_t)i * gridDim.x + blockIdx.x];
  __syncthreads();
  const int64_t begin = (int64_t)blockIdx.x * rpb;
  const int64_t end   = begin + rpb < nb ? begin + rpb : nb;
  for (int64_t r = begin + threadIdx.x; r < end; r += blockDim.x) {
    unsigned long long k = 0;
    const int bin        = jb_bin(c, r, mask, tbits, ntiles, nulls_equal, k);
    if (bin < 0) {
      slot_of_row[r] = -1;
      continue;
    }
    const uint32_t pos = atomicAdd(&cur[bin], 1u);
    recs[pos]          = make_uint4((uint32_t)k, (uint32_t)(k >> 32), (uint32_t)r, 0u);
  }
}

// Probe filter of a tiled table: kFiltWords 32-bit words per tile, two bits per key in one word
// (a blocked Bloom filter), chosen by hash bits 32 and up, which the slot index (bits below 30)
// does not use.  A probe key whose two bits are not both set is not in the table, and the probe
// reads one filter word (a table of 2^27 slots has a 16 MB filter, mostly cache-resident)
// instead of a slot line of the 2 GB table.  The tile's block builds its words in LDS.
constexpr int kFiltWords = 128;
__device__ __forceinline__ uint32_t filt_word(uint64_t hk, int tbits, uint64_t mask)
{
  return (uint32_t)((((hk & mask) >> tbits) * kFiltWords) + ((hk >> 32) & (kFiltWords - 1)));
}
__device__ __forceinline__ uint32_t filt_bits(uint64_t hk) { return (1u << ((hk >> 40) & 31)) | (1u << ((hk >> 45) & 31)); }

// One block per bin: bins < ntiles are table tiles of `tile` slots; bin ntiles is slot cap (the
// empty-marker key), bin ntiles + 1 is slot cap + 1 (the null key).
__global__ __launch_bounds__(256) void k_jb_tiles(const uint4* __restrict__ recs, const uint64_t* __restrict__ offs,
                                                  int nblk_hist, int ntiles, int tile, uint64_t cap,
                                                  JSlot* __restrict__ tab, uint32_t* __restrict__ rows,
                                                  int32_t* __restrict__ slot_of_row, uint32_t* __restrict__ filt,
                                                  int32_t* __restrict__ err)
{
  __shared__ uint32_t tfilt[kFiltWords];
  __shared__ unsigned long long tkeys[kBTile];
  __shared__ uint32_t tcnt[kBTile];
  __shared__ uint32_t tpos[kBTile];
  __shared__ uint64_t smem[17];
  __shared__ int bad;
  const int t       = blockIdx.x;
  const int64_t r0  = (int64_t)offs[(int64_t)t * nblk_hist];
  const int64_t r1  = (int64_t)offs[(int64_t)(t + 1) * nblk_hist];
  if (t >= ntiles) {  // a special slot: all the bin's rows share it
    const int64_t slot = (int64_t)cap + (t - ntiles);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
      const uint32_t row = recs[i].z;
      rows[i]            = row;
      slot_of_row[row]   = (int32_t)slot;
    }
    if (threadIdx.x == 0) {
      tab[slot].key   = kEmpty;
      tab[slot].cnt   = (uint32_t)(r1 - r0);
      tab[slot].start = (uint32_t)r0;
    }
    return;
  }
  const uint32_t tmask = (uint32_t)tile - 1;
  for (int i = threadIdx.x; i < tile; i += blockDim.x) {
    tkeys[i] = kEmpty;
    tcnt[i]  = 0;
  }
  for (int i = threadIdx.x; i < kFiltWords; i += blockDim.x) tfilt[i] = 0;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  // keys into the tile (and their filter bits)
  for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const uint4 rc             = recs[i];
    const unsigned long long k = (unsigned long long)rc.x | ((unsigned long long)rc.y << 32);
    const uint64_t hk          = mix64(k);
    atomicOr(&tfilt[(hk >> 32) & (kFiltWords - 1)], filt_bits(hk));
    uint32_t j = (uint32_t)hk & tmask;
    int slot   = -1;
    for (int probes = 0; probes < tile; ++probes) {
      unsigned long long cur = tkeys[j];
      if (cur == kEmpty) cur = atomicCAS(&tkeys[j], kEmpty, k);
      if (cur == kEmpty || cur == k) {
        slot = (int)j;
        break;
      }
      j = (j + 1) & tmask;
    }
    if (slot < 0) bad = 1;
    else atomicAdd(&tcnt[slot], 1u);
  }
  __syncthreads();
  if (bad) {
    if (threadIdx.x == 0) atomicOr(err, 1);
    return;
  }
  for (int i = threadIdx.x; i < kFiltWords; i += blockDim.x) filt[(int64_t)t * kFiltWords + i] = tfilt[i];
  // row starts per slot: exclusive scan of the counts, tile / blockDim consecutive slots per thread
  const int per = tile / (int)blockDim.x > 0 ? tile / (int)blockDim.x : 1;
  const int my0 = threadIdx.x * per;
  uint64_t sum  = 0;
  for (int q = 0; q < per; ++q)
    if (my0 + q < tile) sum += tcnt[my0 + q];
  uint64_t total;
  uint64_t run = block_exclusive_scan(sum, smem, &total);
  for (int q = 0; q < per; ++q)
    if (my0 + q < tile) {
      tpos[my0 + q] = (uint32_t)run;
      run += tcnt[my0 + q];
    }
  __syncthreads();
  JSlot* base = tab + (int64_t)t * tile;
  for (int i = threadIdx.x; i < tile; i += blockDim.x) {
    JSlot sl;
    sl.key   = tkeys[i];
    sl.cnt   = tcnt[i];
    sl.start = (uint32_t)(r0 + tpos[i]);
    base[i]  = sl;
  }
  __syncthreads();
  // rows grouped by slot
  for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const uint4 rc             = recs[i];
    const unsigned long long k = (unsigned long long)rc.x | ((unsigned long long)rc.y << 32);
    uint32_t j                 = (uint32_t)mix64(k) & tmask;
    while (tkeys[j] != k) j = (j + 1) & tmask;  // present: inserted above
    const uint32_t pos = atomicAdd(&tpos[j], 1u);
    rows[r0 + pos]     = rc.z;
    slot_of_row[rc.z]  = (int32_t)((int64_t)t * tile + j);
  }
}

// ---- probe ------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t out_rows_for(uint32_t cnt, int jt)
{
  switch (jt) {
    case LDF_JOIN_INNER: return cnt;
    case LDF_JOIN_LEFT:
    case LDF_JOIN_FULL: return cnt > 0 ? cnt : 1;
    case LDF_JOIN_SEMI: return cnt > 0 ? 1 : 0;
    default: return cnt > 0 ? 0 : 1;  // ANTI
  }
}

// info[r] = start | (cnt << 32) of the lhs row's build rows (cnt = 0: no match).  With
// `mark` (FULL) a matched slot gets kSlotMatched in its count word, stored only by probes that
// find it unset: they have just read that word, so the flag costs no extra random read.
constexpr uint32_t kSlotMatched = 0x80000000u;  // counts stay below 2^31 (build side < 2^29 rows)
template <bool GENERIC>
__global__ __launch_bounds__(kJBlock) void k_join_count(DKeys pk, DKeys bk, int64_t np, JSlot* __restrict__ tab,
                                                        uint64_t mask, uint64_t tmask, uint64_t cap, bool nulls_equal, int jt,
                                                        const uint32_t* __restrict__ filt, int tbits,
                                                        uint64_t* __restrict__ info, bool mark,
                                                        uint64_t* __restrict__ block_tot)
{
  __shared__ uint64_t smem[17];
  int64_t base = (int64_t)blockIdx.x * kJRows;
  uint64_t sum = 0;
  if (np <= 0) {  // no probe rows (the grid still has one block): nothing to load
    if (threadIdx.x == 0) block_tot[blockIdx.x] = 0;
    return;
  }
  if (!GENERIC) {
    // All kJItems keys are read, then their filter words (tiled tables), then the home slots of
    // the rows the filter passes (one 16-byte load: key, count and start together), each step's
    // loads all in flight before any is used.  Only a row whose home slot holds another key
    // probes further.
    const DCol& c = pk.c[0];
    int64_t slot[kJItems];
    unsigned long long key[kJItems];
    uint32_t fw[kJItems], fb[kJItems];
    // (the loads are unconditional -- rows past the end read row np - 1, rows without a filter
    // word or home slot read word / slot 0 -- and masked afterwards: a load under a branch is
    // waited for before the next one is issued)
    uint32_t vb[kJItems];
#pragma unroll
    for (int j = 0; j < kJItems; ++j) {
      const int64_t r  = base + j * kJBlock + threadIdx.x;
      const int64_t rr = r < np ? r : np - 1;
      vb[j]            = valid_bit_nb(c.valid, c.data, rr);
      key[j]           = load_bits_nb(c.data, c.width, rr);
    }
#pragma unroll
    for (int j = 0; j < kJItems; ++j) {
      const int64_t r   = base + j * kJBlock + threadIdx.x;
      const bool in     = r < np;
      const bool valid  = vb[j] != 0;
      key[j]            = valid ? key[j] : 0;
      const uint64_t hk = mix64(key[j]);
      const int64_t sv  = key[j] == kEmpty ? (int64_t)cap : (int64_t)(hk & mask);
      slot[j]           = !in ? -1 : (valid ? sv : (nulls_equal ? (int64_t)cap + 1 : -1));
      const bool useft  = filt != nullptr && in && valid && key[j] != kEmpty;
      fw[j]             = useft ? filt_word(hk, tbits, mask) : 0;
      fb[j]             = useft ? filt_bits(hk) : 0;
    }
    if (filt != nullptr) {
      uint32_t word[kJItems];
#pragma unroll
      for (int j = 0; j < kJItems; ++j) word[j] = filt[fw[j]];
#pragma unroll
      for (int j = 0; j < kJItems; ++j)
        if (fb[j] != 0 && (word[j] & fb[j]) != fb[j]) slot[j] = -1;  // not in the table
    }
    uint4 home[kJItems];
#pragma unroll
    for (int j = 0; j < kJItems; ++j) home[j] = *reinterpret_cast<const uint4*>(&tab[slot[j] >= 0 ? slot[j] : 0]);
#pragma unroll
    for (int j = 0; j < kJItems; ++j)
      if (slot[j] < 0) home[j] = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < kJItems; ++j) {
      const int64_t r = base + j * kJBlock + threadIdx.x;
      if (r >= np) continue;
      int64_t s    = -1;
      uint32_t raw = 0, st = 0;  // raw: the count word with the matched flag
      if (slot[j] >= 0) {
        const unsigned long long hk = (unsigned long long)home[j].x | ((unsigned long long)home[j].y << 32);
        if (slot[j] >= (int64_t)cap || hk == key[j]) {  // the two special slots hold their rows whatever the key
          s   = slot[j];
          raw = home[j].z;
          st  = home[j].w;
        } else if (hk != kEmpty) {
          uint64_t i = tile_next((uint64_t)slot[j], tmask);
          for (uint64_t probes = 1; probes <= tmask; ++probes) {
            const unsigned long long cur = tab[i].key;
            if (cur == key[j]) {
              s   = (int64_t)i;
              raw = tab[i].cnt;
              st  = tab[i].start;
              break;
            }
            if (cur == kEmpty) break;
            i = tile_next(i, tmask);
          }
        }
      }
      const uint32_t cnt = raw & ~kSlotMatched;
      if (mark && cnt > 0 && !(raw & kSlotMatched)) tab[s].cnt = raw | kSlotMatched;  // same value from every writer
      info[r] = (uint64_t)st | ((uint64_t)cnt << 32);
      sum += out_rows_for(cnt, jt);
    }
  } else {
#pragma unroll 2
    for (int j = 0; j < kJItems; ++j) {
      int64_t r = base + j * kJBlock + threadIdx.x;
      if (r < np) {
        int64_t s    = probe_find<GENERIC>(pk, bk, r, tab, mask, tmask, cap, nulls_equal);
        uint32_t cnt = 0, st = 0;
        if (s >= 0) {
          const uint32_t raw = tab[s].cnt;
          cnt                = raw & ~kSlotMatched;
          st                 = tab[s].start;
          if (mark && cnt > 0 && !(raw & kSlotMatched)) tab[s].cnt = raw | kSlotMatched;
        }
        info[r] = (uint64_t)st | ((uint64_t)cnt << 32);
        sum += out_rows_for(cnt, jt);
      }
    }
  }
  uint64_t t = block_reduce_sum(sum, smem);
  if (threadIdx.x == 0) block_tot[blockIdx.x] = t;
}

__global__ __launch_bounds__(kJBlock) void k_join_emit(const uint64_t* __restrict__ info, int64_t np, int jt,
                                                       const uint32_t* __restrict__ rows,
                                                       const uint64_t* __restrict__ block_off, int32_t* __restrict__ lmap,
                                                       int32_t* __restrict__ rmap)
{
  __shared__ uint64_t off[kJRows + 1];
  __shared__ uint32_t st[kJRows];
  __shared__ uint32_t mc[kJRows];
  __shared__ uint64_t smem[17];
  int64_t base = (int64_t)blockIdx.x * kJRows;
  int nrows    = (int)((np - base) < kJRows ? (np - base) : kJRows);

  // each thread owns kJItems consecutive rows for the block scan
  uint64_t v[kJItems];
  uint64_t s = 0;
  uint64_t xs[kJItems];
#pragma unroll
  for (int j = 0; j < kJItems; ++j) {  // unconditional loads (past the block's rows: its last row)
    const int li = threadIdx.x * kJItems + j;
    xs[j]        = info[base + (li < nrows ? li : nrows - 1)];
  }
#pragma unroll
  for (int j = 0; j < kJItems; ++j) {
    int li = threadIdx.x * kJItems + j;
    uint64_t o = 0;
    if (li < nrows) {
      const uint64_t x = xs[j];
      uint32_t c       = (uint32_t)(x >> 32);
      st[li]           = (uint32_t)x;
      mc[li]           = c;
      o                = out_rows_for(c, jt);
    }
    v[j] = o;
    s += o;
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan(s, smem, &total);
#pragma unroll
  for (int j = 0; j < kJItems; ++j) {
    int li = threadIdx.x * kJItems + j;
    if (li < nrows) off[li] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) off[nrows] = total;
  __syncthreads();

  const uint64_t gbase = block_off[blockIdx.x];
  for (uint64_t q = threadIdx.x; q < total; q += blockDim.x) {
    // last row li with off[li] <= q (rows with zero output have equal offsets; take the last)
    int lo = 0, hi = nrows - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= q) lo = mid;
      else hi = mid - 1;
    }
    uint64_t k      = q - off[lo];
    int32_t lrow    = (int32_t)(base + lo);
    uint64_t gq     = gbase + q;
    lmap[gq]        = lrow;
    if (rmap != nullptr) rmap[gq] = (mc[lo] == 0) ? -1 : (int32_t)rows[st[lo] + k];
  }
}

// FULL join: build rows whose key no lhs row matched (or whose key is null with !nulls_equal).
__global__ __launch_bounds__(kJBlock) void k_join_unmatched(const int32_t* __restrict__ slot_of_row, int64_t nb,
                                                            const JSlot* __restrict__ tab,
                                                            const uint64_t* __restrict__ block_off, uint64_t base_out,
                                                            uint64_t* __restrict__ block_tot, int32_t* __restrict__ lmap,
                                                            int32_t* __restrict__ rmap)
{
  __shared__ uint64_t smem[17];
  int64_t base = (int64_t)blockIdx.x * kJRows;
  bool f[kJItems];
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < kJItems; ++j) {
    int64_t r = base + (int64_t)threadIdx.x * kJItems + j;
    bool u    = false;
    if (r < nb) {
      int32_t sl = slot_of_row[r];
      u          = (sl < 0) || !(tab[sl].cnt & kSlotMatched);
    }
    f[j] = u;
    s += u ? 1 : 0;
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan(s, smem, &total);
  if (block_off == nullptr) {
    if (threadIdx.x == 0) block_tot[blockIdx.x] = total;
    return;
  }
  uint64_t pos = base_out + block_off[blockIdx.x] + run;
#pragma unroll
  for (int j = 0; j < kJItems; ++j) {
    if (f[j]) {
      lmap[pos] = -1;
      rmap[pos] = (int32_t)(base + (int64_t)threadIdx.x * kJItems + j);
      ++pos;
    }
  }
}

// ---- host -------------------------------------------------------------------------------
static DKeys make_keys(const ldf_table_view* t, const int32_t* idx, int32_t nkeys, const char* side)
{
  DKeys k;
  k.n = nkeys;
  for (int i = 0; i < nkeys; ++i) {
    if (idx[i] < 0 || idx[i] >= t->ncols) throw Error(LDF_E_INVALID, std::string(side) + " key index out of range");
    k.c[i] = to_dcol(t->cols[idx[i]]);
  }
  return k;
}

static void run_join(const ldf_table_view* lhs, const ldf_table_view* rhs, const int32_t* lhs_keys,
                     const int32_t* rhs_keys, int32_t nkeys, int32_t jt, bool nulls_equal, const int32_t* lhs_out,
                     int32_t n_lhs_out, const int32_t* rhs_out, int32_t n_rhs_out, Alloc& a, ldf_table_out* out)
{
  hipStream_t s = a.stream();
  LDF_EXPECTS(lhs != nullptr && rhs != nullptr && out != nullptr && out->cols != nullptr, "join: NULL argument");
  LDF_EXPECTS(nkeys >= 1 && nkeys <= kMaxKeyCols, "join: between 1 and 8 key columns");
  LDF_EXPECTS(jt >= LDF_JOIN_INNER && jt <= LDF_JOIN_ANTI, "Unknown JoinType");
  if ((jt == LDF_JOIN_SEMI || jt == LDF_JOIN_ANTI) && n_rhs_out > 0)
    throw Error(LDF_E_INVALID, "rhs must have no output columns for SEMI and ANTI join");
  LDF_EXPECTS(out->ncols >= n_lhs_out + n_rhs_out, "join: output table has too few column slots");
  int64_t np = lhs->ncols > 0 ? lhs->cols[0].length : 0;
  int64_t nb = rhs->ncols > 0 ? rhs->cols[0].length : 0;
  for (int i = 0; i < lhs->ncols; ++i) validate_view(lhs->cols[i], np, "join lhs");
  for (int i = 0; i < rhs->ncols; ++i) validate_view(rhs->cols[i], nb, "join rhs");
  LDF_EXPECTS(np < INT32_MAX && nb < INT32_MAX, "join: each side is limited to 2^31-1 rows per call");
  for (int i = 0; i < n_lhs_out; ++i) LDF_EXPECTS(lhs_out[i] >= 0 && lhs_out[i] < lhs->ncols, "join: lhs_out index out of range");
  for (int i = 0; i < n_rhs_out; ++i) LDF_EXPECTS(rhs_out[i] >= 0 && rhs_out[i] < rhs->ncols, "join: rhs_out index out of range");
  DKeys pk = make_keys(lhs, lhs_keys, nkeys, "lhs");
  DKeys bk = make_keys(rhs, rhs_keys, nkeys, "rhs");
  for (int i = 0; i < nkeys; ++i)
    if (pk.c[i].type != bk.c[i].type) throw Error(LDF_E_INVALID, "join: lhs and rhs key column types differ");
  const bool generic = !(nkeys == 1 && pk.c[0].width > 0);

  uint64_t cap  = next_pow2((uint64_t)(nb > 32 ? 2 * nb : 64));
  uint64_t mask = cap - 1;
  LDF_EXPECTS(cap + 2 < (uint64_t)INT32_MAX, "join: build side too large for one call (2^29 rows)");
  int64_t nslot = (int64_t)cap + 2;
  Buf<JSlot> tab(a, nslot);
  Buf<int32_t> slot_of_row(a, nb > 0 ? nb : 1);
  Buf<uint32_t> rows(a, nb > 0 ? nb : 1);
  // Single fixed-width key: tiles of kBTile slots built in LDS (probing wraps inside a tile).
  // Composite keys, tables over kBMaxBins - 2 tiles, or a tile that fills up: the global build,
  // probing over the whole table.
  const int tile    = (int)std::min<uint64_t>(cap, (uint64_t)kBTile);
  const int ntiles  = (int)std::min<uint64_t>(cap / (uint64_t)tile, (uint64_t)INT32_MAX);
  uint64_t tmask    = mask;
  bool built        = false;
  Buf<int32_t> err(a, 1);
  Buf<uint32_t> filt;  // probe filter of the tiled build (ftbits = log2 tile; -1: none)
  int ftbits = -1;
  bool tiled = false;
  if (!generic && nb > 0 && (int64_t)ntiles + 2 <= kBMaxBins) {
    if (jt == LDF_JOIN_FULL)  // rows of a tile that fills keep -1: the unmatched pass stays in bounds
      fill_u32(reinterpret_cast<uint32_t*>(slot_of_row.get()), 0xFFFFFFFFu, nb, s);
    ProfScope prof("join_build", s);
    static bool attr = false;
    if (!attr) {
      LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_jb_hist, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
      LDF_HIP_TRY(hipFuncSetAttribute((const void*)k_jb_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
      attr = true;
    }
    int tbits = 0;
    while ((1 << tbits) < tile) ++tbits;
    const int nbins   = ntiles + 2;
    const int nblkh   = 256;  // one 1024-thread block per CU (the histogram takes the LDS)
    const int64_t rpb = (nb + nblkh - 1) / nblkh;
    Buf<uint32_t> counts(a, (size_t)nbins * nblkh);
    Buf<uint64_t> offs(a, (size_t)nbins * nblkh + 1);
    Buf<uint4> recs(a, (size_t)nb);
    fill_u32(reinterpret_cast<uint32_t*>(err.get()), 0, 1, s);
    hipLaunchKernelGGL(k_jb_hist, dim3(nblkh), dim3(kBHist), nbins * 4, s, bk.c[0], nb, rpb, mask, tbits, ntiles,
                       nulls_equal, counts.get());
    LDF_HIP_TRY(hipGetLastError());
    exclusive_scan_u32_to_u64(counts.get(), offs.get(), (int64_t)nbins * nblkh, a);
    hipLaunchKernelGGL(k_jb_scatter, dim3(nblkh), dim3(kBHist), nbins * 4, s, bk.c[0], nb, rpb, mask, tbits, ntiles,
                       nulls_equal, (const uint64_t*)offs.get(), recs.get(), slot_of_row.get());
    LDF_HIP_TRY(hipGetLastError());
    filt = Buf<uint32_t>(a, (size_t)ntiles * kFiltWords);
    hipLaunchKernelGGL(k_jb_tiles, dim3(nbins), dim3(256), 0, s, (const uint4*)recs.get(), (const uint64_t*)offs.get(),
                       nblkh, ntiles, tile, cap, tab.get(), rows.get(), slot_of_row.get(), filt.get(), err.get());
    LDF_HIP_TRY(hipGetLastError());
    // A tile that filled up is found when the output size is read back (no synchronisation
    // here); the probe of such a table reads only in-bounds slots, and its result is discarded.
    built  = true;
    tiled  = true;
    tmask  = (uint64_t)tile - 1;
    ftbits = tbits;
  }
  auto global_build = [&]() {
    ProfScope prof("join_build", s);
    tmask = mask;
    fill_u32(reinterpret_cast<uint32_t*>(err.get()), 0, 1, s);
    hipLaunchKernelGGL(k_jslot_fill, dim3(grid_for(nslot, 256)), dim3(256), 0, s, tab.get(), nslot);
    LDF_HIP_TRY(hipGetLastError());
    if (generic)
      hipLaunchKernelGGL(k_join_build<true>, dim3(grid_for(nb, 256)), dim3(256), 0, s, bk, nb, tab.get(), mask, tmask,
                         cap, nulls_equal, slot_of_row.get(), err.get());
    else
      hipLaunchKernelGGL(k_join_build<false>, dim3(grid_for(nb, 256)), dim3(256), 0, s, bk, nb, tab.get(), mask, tmask,
                         cap, nulls_equal, slot_of_row.get(), err.get());
    LDF_HIP_TRY(hipGetLastError());
    {
      Buf<uint32_t> cnt(a, nslot);
      Buf<uint64_t> start(a, nslot + 1);
      hipLaunchKernelGGL(k_jslot_counts, dim3(grid_for(nslot, 256)), dim3(256), 0, s, tab.get(), nslot, cnt.get());
      LDF_HIP_TRY(hipGetLastError());
      exclusive_scan_u32_to_u64(cnt.get(), start.get(), nslot, a);
      hipLaunchKernelGGL(k_jslot_starts, dim3(grid_for(nslot, 256)), dim3(256), 0, s, tab.get(), nslot, start.get());
      LDF_HIP_TRY(hipGetLastError());
      fill_u32(cnt.get(), 0, nslot, s);
      hipLaunchKernelGGL(k_join_group_rows, dim3(grid_for(nb, 256)), dim3(256), 0, s, slot_of_row.get(), nb, tab.get(),
                         cnt.get(), rows.get());
      LDF_HIP_TRY(hipGetLastError());
    }
  };
  if (!built && nb > 0) {
    global_build();
  } else if (nb == 0) {
    hipLaunchKernelGGL(k_jslot_fill, dim3(grid_for(nslot, 256)), dim3(256), 0, s, tab.get(), nslot);
    LDF_HIP_TRY(hipGetLastError());
  }

  // probe: count
  int64_t nblk = (np + kJRows - 1) / kJRows;
  if (nblk < 1) nblk = 1;
  Buf<uint64_t> info(a, np > 0 ? np : 1);
  Buf<uint64_t> btot(a, nblk);
  Buf<uint64_t> boff(a, nblk + 1);
  const bool mark = jt == LDF_JOIN_FULL;  // matched slots flagged in their count words
  int64_t ublk = (nb + kJRows - 1) / kJRows;
  if (ublk < 1) ublk = 1;
  Buf<uint64_t> utot, uoff;
  if (jt == LDF_JOIN_FULL) {
    utot = Buf<uint64_t>(a, ublk);
    uoff = Buf<uint64_t>(a, ublk + 1);
  }
  auto probe = [&]() {
  ProfScope prof("join_count", s);
  if (generic)
    hipLaunchKernelGGL(k_join_count<true>, dim3(nblk), dim3(kJBlock), 0, s, pk, bk, np, tab.get(), mask, tmask, cap,
                       nulls_equal, (int)jt, (const uint32_t*)nullptr, 0, info.get(), mark, btot.get());
  else
    hipLaunchKernelGGL(k_join_count<false>, dim3(nblk), dim3(kJBlock), 0, s, pk, bk, np, tab.get(), mask, tmask, cap,
                       nulls_equal, (int)jt, ftbits >= 0 ? (const uint32_t*)filt.get() : nullptr, ftbits, info.get(),
                       mark, btot.get());
  LDF_HIP_TRY(hipGetLastError());
  exclusive_scan_u64(btot.get(), boff.get(), nblk, a);
  if (jt == LDF_JOIN_FULL) {
    hipLaunchKernelGGL(k_join_unmatched, dim3(ublk), dim3(kJBlock), 0, s, slot_of_row.get(), nb, (const JSlot*)tab.get(),
                       (const uint64_t*)nullptr, (uint64_t)0, utot.get(), (int32_t*)nullptr, (int32_t*)nullptr);
    LDF_HIP_TRY(hipGetLastError());
    exclusive_scan_u64(utot.get(), uoff.get(), ublk, a);
  }
  };
  // output sizes (and the tiled build's overflow flag) read back with one synchronisation
  uint64_t n1 = 0, n2 = 0;
  auto read_sizes = [&](bool with_err) -> bool {
    uint64_t* st = static_cast<uint64_t*>(pinned_stage(24));
    st[0] = st[1] = st[2] = 0;
    LDF_HIP_TRY(hipMemcpyAsync(st, boff.get() + nblk, 8, hipMemcpyDeviceToHost, s));
    if (jt == LDF_JOIN_FULL) LDF_HIP_TRY(hipMemcpyAsync(st + 1, uoff.get() + ublk, 8, hipMemcpyDeviceToHost, s));
    if (with_err) LDF_HIP_TRY(hipMemcpyAsync(st + 2, err.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LDF_HIP_TRY(hipStreamSynchronize(s));
    n1 = st[0];
    n2 = st[1];
    return st[2] == 0;
  };
  probe();
  if (!read_sizes(tiled)) {  // a tile filled up: the global build, and the probe again
    tmask  = mask;
    ftbits = -1;
    global_build();
    probe();
    read_sizes(false);
  }
  int64_t n_out = (int64_t)(n1 + n2);

  // emit gather maps
  const bool with_rmap = !(jt == LDF_JOIN_SEMI || jt == LDF_JOIN_ANTI);
  Buf<int32_t> lmap(a, n_out > 0 ? n_out : 1);
  Buf<int32_t> rmap;
  if (with_rmap) rmap = Buf<int32_t>(a, n_out > 0 ? n_out : 1);
  if (n1 > 0) {
    ProfScope prof("join_emit", s);
    hipLaunchKernelGGL(k_join_emit, dim3(nblk), dim3(kJBlock), 0, s, (const uint64_t*)info.get(), np, (int)jt,
                       (const uint32_t*)rows.get(), (const uint64_t*)boff.get(), lmap.get(), rmap.get());
    LDF_HIP_TRY(hipGetLastError());
  }
  if (n2 > 0) {
    hipLaunchKernelGGL(k_join_unmatched, dim3(ublk), dim3(kJBlock), 0, s, slot_of_row.get(), nb, (const JSlot*)tab.get(),
                       (const uint64_t*)uoff.get(), n1, utot.get(), lmap.get(), rmap.get());
    LDF_HIP_TRY(hipGetLastError());
  }
  // release the probe-side temporaries before the output columns are allocated
  info.reset();
  tab.reset();

  // gather output columns
  out->ncols = n_lhs_out + n_rhs_out;
  std::vector<const ldf_column_view*> lsrc(n_lhs_out), rsrc(n_rhs_out);
  for (int i = 0; i < n_lhs_out; ++i) lsrc[i] = &lhs->cols[lhs_out[i]];
  for (int i = 0; i < n_rhs_out; ++i) rsrc[i] = &rhs->cols[rhs_out[i]];
  // both sides' null counts land in one device array, read back with one synchronisation
  const int nout_cols = n_lhs_out + n_rhs_out;
  Buf<unsigned long long> ncnt(a, nout_cols > 0 ? nout_cols : 1);
  fill_u64(reinterpret_cast<uint64_t*>(ncnt.get()), 0, nout_cols > 0 ? nout_cols : 1, s);
  gather_columns(lsrc.data(), n_lhs_out, lmap.get(), n_out, jt == LDF_JOIN_FULL, a, out->cols, false, ncnt.get());
  int allocated = n_lhs_out;  // output columns this function owns (a failed gather frees its own)
  try {
    gather_columns(rsrc.data(), n_rhs_out, rmap.get(), n_out, jt == LDF_JOIN_LEFT || jt == LDF_JOIN_FULL, a,
                   out->cols + n_lhs_out, /*random_map=*/true, ncnt.get() + n_lhs_out);
    allocated = nout_cols;
    bool any_nullable = false;  // no output bitmap: every null count is 0, no readback
    for (int i = 0; i < nout_cols; ++i) any_nullable |= out->cols[i].validity != nullptr;
    if (any_nullable) {
      std::vector<unsigned long long> hc(nout_cols);
      copy_to_host(hc.data(), ncnt.get(), hc.size() * sizeof(unsigned long long), s);
      for (int i = 0; i < nout_cols; ++i) out->cols[i].null_count = out->cols[i].validity ? (int64_t)hc[i] : 0;
    } else {
      for (int i = 0; i < nout_cols; ++i) out->cols[i].null_count = 0;
    }
  } catch (...) {
    for (int i = 0; i < allocated; ++i) free_column(&out->cols[i], a);
    out->ncols = 0;
    throw;
  }
}

}  // namespace ldf

using namespace ldf;

extern "C" int ldf_join(const ldf_table_view* lhs, const ldf_table_view* rhs, const int32_t* lhs_keys,
                        const int32_t* rhs_keys, int32_t nkeys, int32_t join_type, int32_t nulls_equal,
                        const int32_t* lhs_out, int32_t n_lhs_out, const int32_t* rhs_out, int32_t n_rhs_out,
                        const ldf_allocator* alloc, void* stream, ldf_table_out* out)
{
  return guarded([&] {
    Alloc a(alloc, (hipStream_t)stream);
    run_join(lhs, rhs, lhs_keys, rhs_keys, nkeys, join_type, nulls_equal != 0, lhs_out, n_lhs_out, rhs_out, n_rhs_out, a,
             out);
  });
}
