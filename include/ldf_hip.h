/*
 * ldf_hip.h -- C ABI of the MI355X (gfx950) hash join / groupby / hash-partition path.
 *
 * This is the boundary the GPU task variants of legate-dataframe would bind: plain pointers
 * and sizes, no C++ types, no exceptions crossing it.  Every entry point enqueues its work on
 * the caller's HIP stream; the host synchronises only for small readbacks that the next
 * allocation or the output needs (row counts, string sizes, sample statistics, null counts).
 *
 * Reference interfaces replaced (paths relative to the legate-dataframe source tree):
 *   ldf_join            <- JoinTask<bool>::gpu_variant (cpp/src/join.cu:187-244), which calls
 *                          cudf_join + cudf_join_and_gather (cpp/src/join.cu:32-155); argument
 *                          order follows cpp/src/join.cpp:410-420.
 *   ldf_groupby         <- GroupByAggregationTask::gpu_variant (cpp/src/groupby_aggregation.cu:90-167);
 *                          argument packing follows cpp/src/groupby_aggregation.cpp:211-223.
 *   ldf_hash_partition  <- the local step of task::repartition_by_hash
 *                          (cpp/src/core/repartition_by_hash.cu:296-310: cudf::hash_partition with
 *                          HASH_MURMUR3 and DEFAULT_HASH_SEED, then split at the offsets).
 *   ldf_murmur3_hash    <- the row hash used by that partition step (libcudf MurmurHash3_x86_32
 *                          per column, combined with hash_combine).
 *   ldf_distinct        <- DistinctTask::gpu_variant (cpp/src/stream_compaction.cu:38-72):
 *                          cudf::distinct on the key columns, every column returned
 *                          (client: cpp/src/stream_compaction.cpp:120-146).
 *
 * Column layout in HBM (Arrow-compatible, see ldf_column_view):
 *   fixed width : `data` = length * width bytes, little endian.
 *   validity    : LSB-first bitmap, bit i of byte i/8 set = row i valid; NULL = no nulls and the
 *                 column is not nullable.
 *   large_utf8  : `offsets` = length+1 int64 (first is 0), `data` = chars.
 *
 * Status codes: 0 = success, < 0 = error (see LDF_E_*); ldf_last_error() returns a
 * thread-local message for the last failing call on this thread.
 */
#ifndef LDF_HIP_H
#define LDF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Type ids: the numeric values of arrow::Type::type for the supported subset
 * (the reference describes columns by Arrow type, cpp/include/legate_dataframe/core/column.hpp). */
enum {
  LDF_TYPE_BOOL         = 1, /* one byte per row, 0 or 1 (Legate's bool layout) */
  LDF_TYPE_UINT8        = 2,
  LDF_TYPE_INT8         = 3,
  LDF_TYPE_UINT16       = 4,
  LDF_TYPE_INT16        = 5,
  LDF_TYPE_UINT32       = 6,
  LDF_TYPE_INT32        = 7,
  LDF_TYPE_UINT64       = 8,
  LDF_TYPE_INT64        = 9,
  LDF_TYPE_FLOAT        = 11,
  LDF_TYPE_DOUBLE       = 12,
  LDF_TYPE_DATE32       = 16,
  LDF_TYPE_TIMESTAMP    = 18, /* int64 storage; the unit is the caller's metadata */
  LDF_TYPE_LARGE_STRING = 34
};

/* JoinType / BroadcastInput values are ABI (cpp/include/legate_dataframe/join.hpp:25-26). */
enum { LDF_JOIN_INNER = 0, LDF_JOIN_LEFT = 1, LDF_JOIN_FULL = 2, LDF_JOIN_SEMI = 3, LDF_JOIN_ANTI = 4 };

/* Aggregation kinds (cpp/src/groupby_aggregation.cu:46-88 lists the same set). */
enum {
  LDF_AGG_SUM            = 0,
  LDF_AGG_PRODUCT        = 1,
  LDF_AGG_MIN            = 2,
  LDF_AGG_MAX            = 3,
  LDF_AGG_COUNT          = 4,
  LDF_AGG_MEAN           = 5,
  LDF_AGG_VARIANCE       = 6,
  LDF_AGG_STDDEV         = 7,
  LDF_AGG_APPROX_MEDIAN  = 8,
  LDF_AGG_COUNT_DISTINCT = 9,
  LDF_AGG_COUNT_ALL      = 10
};

/* arrow::compute::CountOptions::CountMode values. */
enum { LDF_COUNT_ONLY_VALID = 0, LDF_COUNT_ONLY_NULL = 1, LDF_COUNT_ALL = 2 };

enum {
  LDF_OK             = 0,
  LDF_E_INVALID      = -1, /* invalid argument (std::invalid_argument in the reference) */
  LDF_E_OOM          = -2, /* the allocator returned NULL */
  LDF_E_HIP          = -3, /* a HIP runtime call failed */
  LDF_E_UNSUPPORTED  = -4, /* valid request this build does not implement */
  LDF_E_INTERNAL     = -5
};

typedef struct ldf_column_view {
  int32_t type_id;
  int32_t reserved;
  int64_t length;
  const void* data;          /* values, or chars for LDF_TYPE_LARGE_STRING */
  const uint8_t* validity;   /* LSB bitmap or NULL */
  const int64_t* offsets;    /* large_utf8 only */
  int64_t null_count;        /* -1 = unknown */
} ldf_column_view;

typedef struct ldf_table_view {
  int32_t ncols;
  const ldf_column_view* cols;
} ldf_table_view;

/* Output column: every buffer comes from the caller's allocator and is owned by the caller. */
typedef struct ldf_column {
  int32_t type_id;
  int32_t reserved;
  int64_t length;
  void* data;
  uint8_t* validity; /* NULL when the output column is not nullable */
  int64_t* offsets;  /* large_utf8 only */
  int64_t null_count;
  int64_t data_bytes;     /* bytes allocated for `data` (chars for strings) */
  int64_t validity_bytes; /* bytes allocated for `validity` */
} ldf_column;

typedef struct ldf_table_out {
  int32_t ncols;     /* set by the caller to the capacity of `cols`, by the callee to the count */
  ldf_column* cols;  /* caller-provided array */
} ldf_table_out;

/* Allocation callbacks (the role TaskMemoryResource plays in the reference,
 * cpp/src/core/allocator.cu:27-81).  `stream` is the call's HIP stream; allocations are
 * stream ordered and must be aligned to 256 bytes (as hipMalloc's are; a misaligned pointer is
 * freed again and the call fails with LDF_E_INVALID).  A NULL ldf_allocator* selects the library default: a caching device
 * allocator that keeps freed blocks (in size classes) per device and stream and reuses them in
 * stream order; ldf_pool_release returns its cached blocks to the device. */
typedef void* (*ldf_alloc_fn)(void* ctx, size_t bytes, void* stream);
typedef void (*ldf_free_fn)(void* ctx, void* ptr, size_t bytes, void* stream);
typedef struct ldf_allocator {
  ldf_alloc_fn alloc;
  ldf_free_fn free;
  void* ctx;
} ldf_allocator;

typedef struct ldf_agg {
  int32_t in_col;     /* input column index (ignored for COUNT_ALL) */
  int32_t kind;       /* LDF_AGG_* */
  int32_t count_mode; /* LDF_COUNT_* for COUNT / COUNT_DISTINCT */
  int32_t reserved;
} ldf_agg;

/* Join of lhs (probe side) with rhs (build side), as cudf::hash_join(rhs keys) probed with the
 * lhs keys in cpp/src/join.cu:32-57, gathered as cpp/src/join.cu:110-155.  The output holds
 * the lhs_out columns then the rhs_out columns (cpp/src/join.cpp:398-400).  An output column
 * is nullable when its input column is, and additionally every rhs column for LEFT and every
 * column for FULL (cpp/src/join.cpp:348-383; unmatched rows are null,
 * cpp/src/join.cu:79-101).  SEMI/ANTI take no rhs output columns.  `nulls_equal` != 0 makes
 * a null key equal a null key (cpp/include/legate_dataframe/join.hpp:85).
 * `out->cols` must have room for n_lhs_out + n_rhs_out columns. */
int ldf_join(const ldf_table_view* lhs,
             const ldf_table_view* rhs,
             const int32_t* lhs_keys,
             const int32_t* rhs_keys,
             int32_t nkeys,
             int32_t join_type,
             int32_t nulls_equal,
             const int32_t* lhs_out,
             int32_t n_lhs_out,
             const int32_t* rhs_out,
             int32_t n_rhs_out,
             const ldf_allocator* alloc,
             void* stream,
             ldf_table_out* out);

/* Groupby-aggregate (cudf::groupby with null_policy::INCLUDE on the keys,
 * cpp/src/groupby_aggregation.cu:143; Arrow hash_<kind> result types and null rules,
 * cpp/src/groupby_aggregation.cpp:41-132).  Output = the key columns then one column per
 * aggregation in order.  Group order is unspecified.  `out->cols` must have room for
 * nkeys + naggs columns. */
int ldf_groupby(const ldf_table_view* table,
                const int32_t* keys,
                int32_t nkeys,
                const ldf_agg* aggs,
                int32_t naggs,
                const ldf_allocator* alloc,
                void* stream,
                ldf_table_out* out);

/* Local hash partition (cudf::hash_partition(table, cols, nparts, HASH_MURMUR3, seed),
 * cpp/src/core/repartition_by_hash.cu:296-302): row r goes to partition
 * murmur3_row_hash(r) mod nparts.  `out` receives every column of `table`, rows grouped by
 * partition (rows keep their input order inside a partition); `offsets` (host memory,
 * nparts+1 entries) receives the partition start rows, offsets[nparts] = row count. */
int ldf_hash_partition(const ldf_table_view* table,
                       const int32_t* cols,
                       int32_t ncols,
                       int32_t nparts,
                       uint32_t seed,
                       const ldf_allocator* alloc,
                       void* stream,
                       ldf_table_out* out,
                       int64_t* offsets);

/* Allocation-free hash partition.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Same result as ldf_hash_partition for a table whose every column is
 * fixed width and non-nullable (1 to 8 columns): the caller passes the outputs -- out->cols[i].data
 * with out->cols[i].data_bytes >= rows * width of column i -- and the call fills in each
 * out->cols[i] (type, length, no validity) and the host `offsets` (nparts + 1 entries).  The
 * counts and scans live in a fixed 16 MiB device array that each GPU holds until the process
 * ends; the call never allocates (more partitions use fewer, larger row blocks so that they fit).
 * Calls that overlap on one GPU race on that array: serialise them (one stream, or one call at a
 * time per device).  Returns LDF_E_INVALID for nullable, string or too many columns or too small
 * output buffers, LDF_E_UNSUPPORTED above 2048 partitions.  Synchronises `stream` to return the
 * offsets. */
int ldf_hash_partition_noalloc(const ldf_table_view* table,
                               const int32_t* cols,
                               int32_t ncols,
                               int32_t nparts,
                               uint32_t seed,
                               void* stream,
                               ldf_table_out* out,
                               int64_t* offsets);

/* Distinct rows by the key columns `keys` (cudf::distinct with KEEP_ANY, null_equality::EQUAL,
 * cpp/src/stream_compaction.cu:38-72): one row per distinct key tuple -- its first row, so the
 * output rows keep their input order -- with every column of `table` in order.  A null key equals
 * a null key; every NaN key equals every NaN and -0.0 equals 0.0 (nan_equality::ALL_EQUAL; the
 * groupby and the join compare floating-point keys bitwise, as Arrow does).
 * `out->cols` must have room for table->ncols columns; one or more key columns (a column named
 * twice counts once). */
int ldf_distinct(const ldf_table_view* table,
                 const int32_t* keys,
                 int32_t nkeys,
                 const ldf_allocator* alloc,
                 void* stream,
                 ldf_table_out* out);

/* Row hash of the selected columns into a device uint32 array of table rows (the value the
 * partition step reduces mod nparts). */
int ldf_murmur3_hash(const ldf_table_view* table,
                     const int32_t* cols,
                     int32_t ncols,
                     uint32_t seed,
                     uint32_t* out_hashes,
                     void* stream);

/* Legate-layout adapters (cpp/src/core/null_mask.cu:38-114, cpp/src/core/ranges.cu:33-138):
 * one byte per row validity <-> LSB bitmap, Rect<1>{lo,hi} (inclusive, global) string ranges
 * <-> local int64 offsets.
 * ldf_bools_to_bitmask writes the (n+7)/8 bitmap bytes (bits past n in the last byte are 0) and,
 * when null_count_dev (device memory) is not NULL, stores the count of zero bools there.
 * ldf_ranges_to_offsets: offsets[i] = ranges[i].lo - ranges[0].lo, offsets[n] = ranges[n-1].hi + 1
 * - ranges[0].lo (an empty string has hi = lo - 1); n + 1 offsets are written (offsets[0] = 0
 * when n = 0).  ldf_offsets_to_ranges: ranges[i] = {base + offsets[i], base + offsets[i+1] - 1}
 * (base 0 gives the reference's local ranges, cudf_offsets_to_local_ranges). */
int ldf_bools_to_bitmask(const uint8_t* bools, int64_t n, uint8_t* bitmask, int64_t* null_count_dev,
                         void* stream);
int ldf_bitmask_to_bools(const uint8_t* bitmask, int64_t n, uint8_t* bools, void* stream);
int ldf_ranges_to_offsets(const int64_t* ranges /* 2 per row */, int64_t n, int64_t* offsets,
                          void* stream);
int ldf_offsets_to_ranges(const int64_t* offsets, int64_t n, int64_t global_base, int64_t* ranges,
                          void* stream);

/* Per-kernel timing with HIP events recorded on the call's stream around the main kernels of
 * each entry point (off by default; no counterpart in the reference, whose only tracing is
 * Legion Prof provenance labels, python/legate_dataframe/utils.py:35-46).  ldf_profile_get
 * returns the launches and summed milliseconds recorded for `kernel` since the last reset. */
void ldf_profile_enable(int on);
void ldf_profile_reset(void);
int ldf_profile_get(const char* kernel, int64_t* launches, double* total_ms);

/* Measures what HBM gives on the current GPU: one launch that reads `nbytes` at `src` (copy = 0;
 * `dst` then only needs 16 bytes) or copies them to `dst` (copy = 1), with 16-byte loads, four in
 * flight per thread, over `nblocks` blocks of 256 threads; `*ms` is its duration from HIP events
 * on `stream`.  A measurement helper with no counterpart in the reference: it gives the
 * practical ceilings the kernels' bandwidth is compared with beside the published peak. */
int ldf_hbm_probe(const void* src, void* dst, int64_t nbytes, int32_t copy, int32_t nblocks, double* ms,
                  void* stream);

/* Frees every buffer of an output table through `alloc` (NULL = library default). */
void ldf_free_table(ldf_table_out* out, const ldf_allocator* alloc, void* stream);

/* Synchronises each device that holds cached blocks of the default allocator and frees those
 * blocks (blocks in use stay allocated).  No counterpart in the reference, whose task memory
 * comes from Legate's pool (cpp/src/core/allocator.cu:27-81). */
void ldf_pool_release(void);

/* The default allocator's own entry points, for a host layer that owns buffers the library
 * returned with a NULL allocator (the single-owner hand-off of cpp/src/core/column.cu:64-70 and
 * cpp/src/core/allocator.cu:60-81: the buffer is released to the allocator that made it).
 * ldf_pool_alloc returns a block of at least `bytes` (256-byte aligned) usable on `stream`, or
 * NULL when the device is out of memory; ldf_pool_free returns a block of the default allocator
 * to its cache for reuse in `stream`'s order (other pointers are ignored).  ldf_pool_stats
 * reports the blocks in use and the cached blocks with their bytes (any pointer may be NULL). */
void* ldf_pool_alloc(size_t bytes, void* stream);
void ldf_pool_free(void* ptr, void* stream);
void ldf_pool_stats(int64_t* live_blocks, int64_t* live_bytes, int64_t* cached_blocks, int64_t* cached_bytes);

/* ---- batched form -------------------------------------------------------------------------
 * Many independent joins or groupbys of one schema (the per-partition point tasks a Legate
 * launch runs one by one: JoinTask / GroupByAggregationTask over each partition,
 * cpp/src/join.cpp:410-420 and cpp/src/groupby_aggregation.cpp:211-226) executed together.
 * ldf_batch_*_create inspects the instances (one host synchronisation for the sizes), allocates
 * every table, work list and output buffer of the batch, and may autotune; ldf_batch_run then
 * enqueues the whole computation -- one persistent work-queue kernel per phase draining
 * (instance, tile) items -- with no host synchronisation and no allocation, captured in a
 * hipGraph on its first call and replayed afterwards when cfg->use_graph is set.
 * ldf_batch_result waits for the stream and describes the outputs: one table whose rows are the
 * instances' results concatenated, instance i owning rows [row_offsets[i], row_offsets[i+1]);
 * the buffers belong to the batch (valid until the next run or destroy).  Results are those
 * of ldf_join / ldf_groupby on each instance (rows in another order).
 * The work-queue kernels take: one fixed-width key column; join types INNER, LEFT, FULL, SEMI,
 * ANTI (FULL: the unmatched build rows of an instance follow its probe rows); fixed-width output
 * columns; groupby kinds SUM, MEAN, COUNT, COUNT_ALL, MIN, MAX over at most 4 value columns
 * (an FP64 SUM / MEAN input outside the exact range |x| < 2^62 makes ldf_batch_result return
 * LDF_E_UNSUPPORTED).  Any other schema ldf_join / ldf_groupby take (several or string keys,
 * string outputs, COUNT_DISTINCT, PRODUCT, VARIANCE, STDDEV, APPROX_MEDIAN, or more groups than
 * the LDS plan holds) is batched instance by instance: ldf_batch_run calls the plain entry point
 * on each instance (allocating, one host synchronisation each; no graph) and concatenates the
 * outputs, each instance's rows in the plain call's order; ldf_batch_config_used then reports
 * all zeros.  A run over inputs whose sizes changed after create may make ldf_batch_result
 * return LDF_E_INVALID (create the batch again).  No counterpart in the reference. */
typedef struct ldf_batch_config {
  double load_factor; /* hash-table load factor in (0, 0.9]; 0 = autotuned */
  int32_t tile_rows;  /* rows per work item: 1024, 2048 or 4096; 0 = autotuned */
  int32_t lds_slots;  /* groupby: slots of the LDS hash table that aggregates one partition
                         (256, 512, 1024 or 2048; the partition fan-out follows from it, so it
                         sets how much of each instance's table lives in LDS); join: instance
                         tables of at most this many slots are probed from an LDS copy (1 = every
                         table in HBM; 256 .. 2048); 0 = autotuned */
  int32_t use_graph;  /* 1 = capture the run's launch sequence in a hipGraph and replay it */
} ldf_batch_config;

typedef struct ldf_batch ldf_batch;

int ldf_batch_join_create(int32_t ninst, const ldf_table_view* lhs, const ldf_table_view* rhs,
                          const int32_t* lhs_keys, const int32_t* rhs_keys, int32_t nkeys, int32_t join_type,
                          int32_t nulls_equal, const int32_t* lhs_out, int32_t n_lhs_out, const int32_t* rhs_out,
                          int32_t n_rhs_out, const ldf_batch_config* cfg, void* stream, ldf_batch** batch);
int ldf_batch_groupby_create(int32_t ninst, const ldf_table_view* tables, const int32_t* keys, int32_t nkeys,
                             const ldf_agg* aggs, int32_t naggs, const ldf_batch_config* cfg, void* stream,
                             ldf_batch** batch);
int ldf_batch_run(ldf_batch* batch, void* stream);
int ldf_batch_result(ldf_batch* batch, void* stream, ldf_table_out* out, int64_t* row_offsets);
int ldf_batch_config_used(const ldf_batch* batch, ldf_batch_config* cfg);
void ldf_batch_destroy(ldf_batch* batch);

/* ---- served form --------------------------------------------------------------------------
 * A query server over device-resident tables.  Tables are registered once (each gets a
 * persistent join index on its key column, built by the join kernels).  Queries are submitted
 * asynchronously:
 *   LDF_QUERY_JOIN     inner join of a probe table (lhs, key column probe_key) with a registered
 *                      table (rhs, its key column), nulls_equal = true; output = every probe
 *                      column then every table column, as ldf_join would return them;
 *   LDF_QUERY_GROUPBY  groupby of the table rows [row_begin, row_end) on the table's key column
 *                      with `aggs` (any kind ldf_groupby takes), as ldf_groupby would return it.
 * Continuous batching: `slots` worker threads, each with its own HIP stream, repeatedly take the
 * waiting queries -- the oldest one and, behind it, those of the same kind (and aggregations)
 * that still fit -- while the memory they are estimated to need fits in `hbm_budget_bytes`
 * (resident indexes and unreleased results count against it), and run them as one batch; the
 * batches of different slots overlap on the GPU.  A groupby over more than 2^21 rows or with
 * a kind the batched groupby does not take, and every query on a table with a string column
 * (such a table gets no key index) or with a string probe column, runs alone through
 * ldf_groupby / ldf_join on the slot's stream.  ldf_server_wait blocks until a query finished and
 * describes its result (buffers owned by the server until ldf_server_release).  Results are
 * those of the plain entry points on the same query (row order aside), alone or in any mix.
 * No counterpart in the reference. */
enum { LDF_QUERY_JOIN = 0, LDF_QUERY_GROUPBY = 1 };

typedef struct ldf_server_config {
  double hbm_budget_bytes; /* memory the server may hold: indexes, batches in flight, results */
  int32_t slots;           /* concurrent batches (worker threads / HIP streams), 1..16 */
  int32_t max_batch;       /* queries per batch */
} ldf_server_config;

typedef struct ldf_query {
  int32_t kind;  /* LDF_QUERY_* */
  int32_t table; /* registered table id */
  ldf_table_view probe; /* JOIN: probe columns (device, valid until the query finished) */
  int32_t probe_key;
  int32_t naggs;        /* GROUPBY */
  const ldf_agg* aggs;  /* GROUPBY (copied at submit) */
  int64_t row_begin;    /* GROUPBY: any row (a range not starting on a 64-row word reads a
                           bit-shifted copy of the validity) */
  int64_t row_end;
} ldf_query;

typedef struct ldf_query_stats {
  int64_t submit_ns; /* steady-clock times */
  int64_t start_ns;
  int64_t end_ns;
  int32_t batch_queries; /* queries in the batch it ran in */
  int32_t plain;         /* 1 = ran through the plain entry point */
} ldf_query_stats;

typedef struct ldf_server ldf_server;

int ldf_server_create(const ldf_server_config* cfg, ldf_server** server);
int ldf_server_add_table(ldf_server* server, const ldf_table_view* table, int32_t key_col, int32_t* table_id);
int ldf_server_submit(ldf_server* server, const ldf_query* query, int64_t* query_id);
int ldf_server_wait(ldf_server* server, int64_t query_id, ldf_table_out* out, ldf_query_stats* stats);
int ldf_server_release(ldf_server* server, int64_t query_id);
void ldf_server_destroy(ldf_server* server);

const char* ldf_last_error(void);
const char* ldf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LDF_HIP_H */
