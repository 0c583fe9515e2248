/*
 * ORACLE -- test infrastructure only.  A single-threaded CPU restatement of the reference's
 * CPU task variants for the join / groupby / hash-partition path, used by tests/ as the
 * parity checker and by bench.py as the CPU baseline ("port").  Nothing in the product links
 * or calls this file.
 *
 * What each function restates (paths relative to the legate-dataframe source tree):
 *   orc_join        arrow_join_and_gather (cpp/src/join.cpp:168-228): Acero hashjoin of lhs
 *                   (probe) with rhs (build), INNER/LEFT_OUTER/FULL_OUTER/LEFT_SEMI/LEFT_ANTI
 *                   (legate_to_arrow_join_type, cpp/src/join.cpp:60-74).  nulls_equal=true is
 *                   the coalesce + "_mask" key emulation of cpp/src/join.cpp:117-209, i.e. a
 *                   null key equals a null key; nulls_equal=false never matches a null key.
 *                   Keys compare bitwise (Arrow's hash join does so for floating point).
 *                   Output: pairs (lhs row, rhs row), -1 for the null-filled side
 *                   (cudf NULLIFY, cpp/src/join.cu:79-101).
 *   orc_groupby     GroupByAggregationTask::cpu_variant (cpp/src/groupby_aggregation.cpp:41-94):
 *                   Acero "aggregate" with hash_<kind>; null keys form one group; SUM/MEAN/
 *                   MIN/MAX/PRODUCT/VARIANCE/STDDEV of a group without valid values is null;
 *                   MIN/MAX skip NaN unless all valid values are NaN; integer SUM/PRODUCT wrap;
 *                   COUNT and COUNT_DISTINCT follow CountOptions modes; APPROX_MEDIAN is the
 *                   exact median the GPU variant computes (cpp/src/groupby_aggregation.cu:75-77;
 *                   Arrow's own CPU kind is a t-digest estimate).  Floating sums use
 *                   Neumaier compensated summation (the FP64 oracle for SUM/MEAN).
 *                   Groups are returned in order of first appearance.
 *   orc_murmur3_32  MurmurHash3_x86_32 (Austin Appleby, public domain).
 *   orc_partition   cudf::hash_partition(HASH_MURMUR3, seed) as called by
 *                   cpp/src/core/repartition_by_hash.cu:296-302: row hash = fold of
 *                   hash_combine(h, murmur3(value, seed)) from h = seed over the hashed
 *                   columns (a null element hashes to 0xFFFFFFFF; a float/double element is
 *                   hashed after NaN -> canonical quiet NaN and -0.0 -> 0.0); partition =
 *                   hash mod P; rows keep input order within a partition.  The murmur3 and
 *                   float normalisation are pinned by published vectors; the hash_combine
 *                   fold, its seed start and the null sentinel are parity unpinned (libcudf,
 *                   where they live, is absent and the reference's tests hold no vectors).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_col {
  int32_t type;            /* LDF_TYPE_* id (arrow::Type::type values) */
  int32_t width;           /* bytes per value, 0 for strings */
  const void* data;        /* values or chars */
  const uint8_t* valid;    /* one byte per row (nonzero = valid) or NULL */
  const int64_t* offsets;  /* strings only */
} orc_col;

enum { T_BOOL = 1, T_UINT8 = 2, T_INT8 = 3, T_UINT16 = 4, T_INT16 = 5, T_UINT32 = 6, T_INT32 = 7, T_UINT64 = 8,
       T_INT64 = 9, T_FLOAT = 11, T_DOUBLE = 12 };
enum { A_SUM = 0, A_PRODUCT, A_MIN, A_MAX, A_COUNT, A_MEAN, A_VARIANCE, A_STDDEV, A_MEDIAN, A_COUNT_DISTINCT,
       A_COUNT_ALL };
enum { M_ONLY_VALID = 0, M_ONLY_NULL = 1, M_ALL = 2 };

void orc_free(void* p) { free(p); }

/* ---------------------------------------------------------------------------------------- */
/* element access                                                                            */
/* ---------------------------------------------------------------------------------------- */
static inline int is_valid(const orc_col* c, int64_t r) { return c->valid == NULL || c->valid[r] != 0; }

static inline uint64_t bits_at(const orc_col* c, int64_t r)
{
  switch (c->width) {
    case 1: return ((const uint8_t*)c->data)[r];
    case 2: return ((const uint16_t*)c->data)[r];
    case 4: return ((const uint32_t*)c->data)[r];
    default: return ((const uint64_t*)c->data)[r];
  }
}

static inline int64_t signed_at(const orc_col* c, int64_t r)
{
  switch (c->width) {
    case 1: return ((const int8_t*)c->data)[r];
    case 2: return ((const int16_t*)c->data)[r];
    case 4: return ((const int32_t*)c->data)[r];
    default: return ((const int64_t*)c->data)[r];
  }
}

static inline int is_float_t(int t) { return t == T_FLOAT || t == T_DOUBLE; }
static inline int is_unsigned_t(int t)
{
  return t == T_BOOL || t == T_UINT8 || t == T_UINT16 || t == T_UINT32 || t == T_UINT64;
}

static inline double double_at(const orc_col* c, int64_t r)
{
  if (c->type == T_DOUBLE) return ((const double*)c->data)[r];
  if (c->type == T_FLOAT) return (double)((const float*)c->data)[r];
  if (is_unsigned_t(c->type)) return (double)bits_at(c, r);
  return (double)signed_at(c, r);
}

static inline uint64_t mix64(uint64_t k)
{
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

static uint64_t bytes_hash(const uint8_t* p, int64_t n)
{
  uint64_t h = 1469598103934665603ull;
  for (int64_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
  return mix64(h ^ (uint64_t)n);
}

static uint64_t key_hash(const orc_col* k, int nk, int64_t r)
{
  uint64_t h = 0x243f6a8885a308d3ull;
  for (int i = 0; i < nk; ++i) {
    uint64_t e;
    if (!is_valid(&k[i], r)) e = 0x13198a2e03707344ull;
    else if (k[i].width == 0) {
      int64_t b = k[i].offsets[r], en = k[i].offsets[r + 1];
      e = bytes_hash((const uint8_t*)k[i].data + b, en - b);
    } else e = mix64(bits_at(&k[i], r) + 0x9e3779b97f4a7c15ull);
    h = mix64(h ^ e) + (uint64_t)i;
  }
  return h;
}

static int elem_equal(const orc_col* a, int64_t ra, const orc_col* b, int64_t rb)
{
  if (a->width == 0) {
    int64_t la = a->offsets[ra + 1] - a->offsets[ra], lb = b->offsets[rb + 1] - b->offsets[rb];
    if (la != lb) return 0;
    return memcmp((const uint8_t*)a->data + a->offsets[ra], (const uint8_t*)b->data + b->offsets[rb], (size_t)la) == 0;
  }
  return bits_at(a, ra) == bits_at(b, rb);
}

/* key tuples equal; a null equals a null */
static int keys_equal(const orc_col* ka, int64_t ra, const orc_col* kb, int64_t rb, int nk)
{
  for (int i = 0; i < nk; ++i) {
    int va = is_valid(&ka[i], ra), vb = is_valid(&kb[i], rb);
    if (va != vb) return 0;
    if (va && !elem_equal(&ka[i], ra, &kb[i], rb)) return 0;
  }
  return 1;
}

static int any_null(const orc_col* k, int nk, int64_t r)
{
  for (int i = 0; i < nk; ++i)
    if (!is_valid(&k[i], r)) return 1;
  return 0;
}

static uint64_t pow2_at_least(uint64_t x)
{
  uint64_t p = 16;
  while (p < x) p <<= 1;
  return p;
}

/* ---------------------------------------------------------------------------------------- */
/* join                                                                                      */
/* ---------------------------------------------------------------------------------------- */
typedef struct {
  int32_t* l;
  int32_t* r;
  int64_t n, cap;
} pairs_t;

static int push_pair(pairs_t* p, int32_t l, int32_t r)
{
  if (p->n == p->cap) {
    int64_t nc  = p->cap ? p->cap * 2 : 1024;
    int32_t* nl = (int32_t*)realloc(p->l, (size_t)nc * sizeof(int32_t));
    int32_t* nr = (int32_t*)realloc(p->r, (size_t)nc * sizeof(int32_t));
    if (!nl || !nr) return -1;
    p->l = nl;
    p->r = nr;
    p->cap = nc;
  }
  p->l[p->n] = l;
  p->r[p->n] = r;
  p->n++;
  return 0;
}

/* Returns the number of output rows (or -1 on allocation failure); *lmap / *rmap are
 * malloc'ed (free with orc_free). */
int64_t orc_join(const orc_col* lk, const orc_col* rk, int nk, int64_t nl, int64_t nr, int join_type, int nulls_equal,
                 int32_t** lmap, int32_t** rmap)
{
  /* build: chained buckets over distinct-key groups */
  uint64_t nb      = pow2_at_least((uint64_t)(nr * 2 + 1));
  int64_t* head    = (int64_t*)malloc(nb * sizeof(int64_t));
  int64_t* next    = (int64_t*)malloc((size_t)(nr > 0 ? nr : 1) * sizeof(int64_t));
  uint64_t* hashes = (uint64_t*)malloc((size_t)(nr > 0 ? nr : 1) * sizeof(uint64_t));
  uint8_t* matched = (uint8_t*)calloc((size_t)(nr > 0 ? nr : 1), 1);
  pairs_t out      = {0};
  if (!head || !next || !hashes || !matched) return -1;
  for (uint64_t i = 0; i < nb; ++i) head[i] = -1;
  /* insert in reverse so each chain lists rows in increasing order */
  for (int64_t r = nr - 1; r >= 0; --r) {
    next[r] = -2;
    if (!nulls_equal && any_null(rk, nk, r)) continue;
    uint64_t h = key_hash(rk, nk, r);
    hashes[r]  = h;
    uint64_t b = h & (nb - 1);
    next[r]    = head[b];
    head[b]    = r;
  }
  for (int64_t l = 0; l < nl; ++l) {
    int found = 0;
    if (nulls_equal || !any_null(lk, nk, l)) {
      uint64_t h = key_hash(lk, nk, l);
      for (int64_t r = head[h & (nb - 1)]; r >= 0; r = next[r]) {
        if (hashes[r] != h || !keys_equal(lk, l, rk, r, nk)) continue;
        found = 1;
        if (join_type == 3 || join_type == 4) break; /* SEMI / ANTI need existence only */
        matched[r] = 1;
        if (push_pair(&out, (int32_t)l, (int32_t)r)) return -1;
      }
    }
    if (!found && (join_type == 1 || join_type == 2 || join_type == 4)) {
      if (push_pair(&out, (int32_t)l, -1)) return -1;
    }
    if (found && join_type == 3) {
      if (push_pair(&out, (int32_t)l, -1)) return -1;
    }
  }
  if (join_type == 2) {
    for (int64_t r = 0; r < nr; ++r)
      if (!matched[r])
        if (push_pair(&out, -1, (int32_t)r)) return -1;
  }
  free(head);
  free(next);
  free(hashes);
  free(matched);
  *lmap = out.l;
  *rmap = out.r;
  return out.n;
}

/* ---------------------------------------------------------------------------------------- */
/* groupby                                                                                   */
/* ---------------------------------------------------------------------------------------- */
typedef struct orc_agg {
  int32_t kind;
  int32_t mode;
  orc_col in; /* ignored for COUNT_ALL */
} orc_agg;

/* Per-aggregation result arrays, one entry per group (allocated by the oracle). */
typedef struct orc_agg_out {
  uint64_t* bits; /* integer results as raw 64-bit (two's complement), MIN/MAX raw value bits */
  double* f64;    /* floating results */
  uint8_t* valid;
} orc_agg_out;

typedef struct {
  double s, c;
} neumaier_t;

static inline void neu_add(neumaier_t* a, double x)
{
  double t = a->s + x;
  if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
  else a->c += (x - t) + a->s;
  a->s = t;
}
static inline double neu_val(const neumaier_t* a) { return a->s + a->c; }

/* group id of every row, groups numbered in order of first appearance */
static int64_t assign_groups(const orc_col* keys, int nk, int64_t n, int64_t* gid, int64_t** rep_out)
{
  uint64_t cap   = pow2_at_least((uint64_t)(n * 2 + 1));
  int64_t* slots = (int64_t*)malloc(cap * sizeof(int64_t)); /* group id */
  int64_t* rep   = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  uint64_t* gh   = (uint64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
  if (!slots || !rep || !gh) return -1;
  for (uint64_t i = 0; i < cap; ++i) slots[i] = -1;
  int64_t ng = 0;
  for (int64_t r = 0; r < n; ++r) {
    uint64_t h = key_hash(keys, nk, r);
    uint64_t i = h & (cap - 1);
    while (1) {
      int64_t g = slots[i];
      if (g < 0) {
        slots[i] = ng;
        rep[ng]  = r;
        gh[ng]   = h;
        gid[r]   = ng++;
        break;
      }
      if (gh[g] == h && keys_equal(keys, r, keys, rep[g], nk)) {
        gid[r] = g;
        break;
      }
      i = (i + 1) & (cap - 1);
    }
  }
  free(slots);
  free(gh);
  *rep_out = rep;
  return ng;
}

static inline uint64_t ordered_key(const orc_col* c, int64_t r)
{
  if (is_float_t(c->type)) {
    double d   = double_at(c, r);
    uint64_t b;
    memcpy(&b, &d, 8);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
  }
  if (is_unsigned_t(c->type)) return bits_at(c, r);
  return (uint64_t)signed_at(c, r) ^ 0x8000000000000000ull;
}

/* median: a valid value and its sort key (ordered_key, NaN after every number) */
typedef struct {
  uint64_t key;
  double value;
} med_item;
static int med_cmp(const void* a, const void* b)
{
  uint64_t x = ((const med_item*)a)->key, y = ((const med_item*)b)->key;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* distinct (group, value) pairs */
typedef struct {
  int64_t* g;
  int64_t* row;
  uint64_t* h;
  uint64_t cap, n;
} pairset_t;

static int pairset_insert(pairset_t* s, const orc_col* c, int64_t g, int64_t r)
{
  uint64_t vh;
  if (c->width == 0) vh = bytes_hash((const uint8_t*)c->data + c->offsets[r], c->offsets[r + 1] - c->offsets[r]);
  else vh = mix64(bits_at(c, r) ^ 0x51afd7ed558ccdffull);
  uint64_t h = mix64(vh ^ mix64((uint64_t)g + 0x9e3779b97f4a7c15ull));
  uint64_t i = h & (s->cap - 1);
  while (1) {
    if (s->g[i] < 0) {
      s->g[i]   = g;
      s->row[i] = r;
      s->h[i]   = h;
      s->n++;
      return 1;
    }
    if (s->h[i] == h && s->g[i] == g && elem_equal(c, s->row[i], c, r)) return 0;
    i = (i + 1) & (s->cap - 1);
  }
}

/* Returns the number of groups (or -1); *rep receives the first row of each group. */
int64_t orc_groupby(const orc_col* keys, int nk, int64_t n, const orc_agg* aggs, int na, int64_t** rep,
                    orc_agg_out* outs)
{
  int64_t* gid = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  if (!gid) return -1;
  int64_t ng = assign_groups(keys, nk, n, gid, rep);
  if (ng < 0) return -1;
  size_t G = (size_t)(ng > 0 ? ng : 1);
  for (int a = 0; a < na; ++a) {
    const orc_agg* ag = &aggs[a];
    const orc_col* c  = &ag->in;
    orc_agg_out* o    = &outs[a];
    o->bits           = (uint64_t*)calloc(G, sizeof(uint64_t));
    o->f64            = (double*)calloc(G, sizeof(double));
    o->valid          = (uint8_t*)calloc(G, 1);
    int64_t* cnt      = (int64_t*)calloc(G, sizeof(int64_t));
    if (!o->bits || !o->f64 || !o->valid || !cnt) return -1;
    switch (ag->kind) {
      case A_COUNT_ALL:
        for (int64_t r = 0; r < n; ++r) cnt[gid[r]]++;
        for (size_t g = 0; g < G; ++g) o->bits[g] = (uint64_t)cnt[g], o->valid[g] = 1;
        break;
      case A_COUNT:
        for (int64_t r = 0; r < n; ++r) {
          int v = is_valid(c, r);
          if (ag->mode == M_ALL || (ag->mode == M_ONLY_NULL ? !v : v)) cnt[gid[r]]++;
        }
        for (size_t g = 0; g < G; ++g) o->bits[g] = (uint64_t)cnt[g], o->valid[g] = 1;
        break;
      case A_SUM:
      case A_MEAN:
      case A_VARIANCE:
      case A_STDDEV: {
        neumaier_t* acc = (neumaier_t*)calloc(G, sizeof(neumaier_t));
        if (!acc) return -1;
        int fp = is_float_t(c->type) || ag->kind != A_SUM;
        for (int64_t r = 0; r < n; ++r) {
          if (!is_valid(c, r)) continue;
          int64_t g = gid[r];
          cnt[g]++;
          if (fp) neu_add(&acc[g], double_at(c, r));
          else o->bits[g] += is_unsigned_t(c->type) ? bits_at(c, r) : (uint64_t)signed_at(c, r);
        }
        if (ag->kind == A_VARIANCE || ag->kind == A_STDDEV) {
          double* mean = (double*)malloc(G * sizeof(double));
          neumaier_t* m2 = (neumaier_t*)calloc(G, sizeof(neumaier_t));
          if (!mean || !m2) return -1;
          for (size_t g = 0; g < G; ++g) mean[g] = cnt[g] ? neu_val(&acc[g]) / (double)cnt[g] : 0.0;
          for (int64_t r = 0; r < n; ++r) {
            if (!is_valid(c, r)) continue;
            double d = double_at(c, r) - mean[gid[r]];
            neu_add(&m2[gid[r]], d * d);
          }
          for (size_t g = 0; g < G; ++g) {
            double v = cnt[g] ? neu_val(&m2[g]) / (double)cnt[g] : 0.0;
            o->f64[g] = ag->kind == A_STDDEV ? sqrt(v) : v;
          }
          free(mean);
          free(m2);
        } else if (fp) {
          for (size_t g = 0; g < G; ++g)
            o->f64[g] = ag->kind == A_MEAN ? (cnt[g] ? neu_val(&acc[g]) / (double)cnt[g] : 0.0) : neu_val(&acc[g]);
        }
        for (size_t g = 0; g < G; ++g) o->valid[g] = cnt[g] > 0;
        free(acc);
        break;
      }
      case A_PRODUCT: {
        int fp = is_float_t(c->type);
        for (size_t g = 0; g < G; ++g) o->bits[g] = 1, o->f64[g] = 1.0;
        for (int64_t r = 0; r < n; ++r) {
          if (!is_valid(c, r)) continue;
          int64_t g = gid[r];
          cnt[g]++;
          if (fp) o->f64[g] *= double_at(c, r);
          else o->bits[g] *= is_unsigned_t(c->type) ? bits_at(c, r) : (uint64_t)signed_at(c, r);
        }
        for (size_t g = 0; g < G; ++g) o->valid[g] = cnt[g] > 0;
        break;
      }
      case A_MIN:
      case A_MAX: {
        int is_min   = ag->kind == A_MIN;
        uint64_t* m  = (uint64_t*)malloc(G * sizeof(uint64_t));
        uint8_t* has = (uint8_t*)calloc(G, 1); /* a non-NaN value was seen */
        int64_t* any = (int64_t*)malloc(G * sizeof(int64_t));
        if (!m || !has || !any) return -1;
        for (size_t g = 0; g < G; ++g) any[g] = -1;
        for (int64_t r = 0; r < n; ++r) {
          if (!is_valid(c, r)) continue;
          int64_t g = gid[r];
          cnt[g]++;
          if (is_float_t(c->type) && isnan(double_at(c, r))) {
            if (any[g] < 0) any[g] = r;
            continue;
          }
          uint64_t k = ordered_key(c, r);
          if (!has[g] || (is_min ? k < m[g] : k > m[g])) {
            m[g]   = k;
            any[g] = r;
            has[g] = 1;
          }
        }
        for (size_t g = 0; g < G; ++g) {
          o->valid[g] = cnt[g] > 0;
          if (cnt[g] > 0) o->bits[g] = bits_at(c, any[g]); /* the raw bits of the chosen value */
        }
        free(m);
        free(has);
        free(any);
        break;
      }
      case A_COUNT_DISTINCT: {
        pairset_t s;
        s.cap = pow2_at_least((uint64_t)(n * 2 + 1));
        s.n   = 0;
        s.g   = (int64_t*)malloc(s.cap * sizeof(int64_t));
        s.row = (int64_t*)malloc(s.cap * sizeof(int64_t));
        s.h   = (uint64_t*)malloc(s.cap * sizeof(uint64_t));
        uint8_t* hasnull = (uint8_t*)calloc(G, 1);
        if (!s.g || !s.row || !s.h || !hasnull) return -1;
        for (uint64_t i = 0; i < s.cap; ++i) s.g[i] = -1;
        for (int64_t r = 0; r < n; ++r) {
          int64_t g = gid[r];
          if (!is_valid(c, r)) {
            hasnull[g] = 1;
            continue;
          }
          if (pairset_insert(&s, c, g, r)) cnt[g]++;
        }
        for (size_t g = 0; g < G; ++g) {
          int64_t v = cnt[g];
          if (ag->mode == M_ALL) v += hasnull[g];
          if (ag->mode == M_ONLY_NULL) v = hasnull[g];
          o->bits[g]  = (uint64_t)v;
          o->valid[g] = 1;
        }
        free(s.g);
        free(s.row);
        free(s.h);
        free(hasnull);
        break;
      }
      case A_MEDIAN: {
        /* the reference's GPU variant maps approximate_median to cudf's exact median
         * (cpp/src/groupby_aggregation.cu:75-77): the valid values of a group sorted (any NaN
         * after every number), the middle one, or 0.5 a + 0.5 b of the two middle ones for an
         * even count (quantile 0.5, linear interpolation); null for a group without valid
         * values.  Values are ordered by their typed value; the result is their double. */
        int64_t* start = (int64_t*)calloc(G + 1, sizeof(int64_t));
        med_item* it   = (med_item*)malloc((size_t)(n > 0 ? n : 1) * sizeof(med_item));
        if (!start || !it) return -1;
        for (int64_t r = 0; r < n; ++r)
          if (is_valid(c, r)) start[gid[r] + 1]++;
        for (size_t g = 0; g < G; ++g) start[g + 1] += start[g];
        int64_t* fill = (int64_t*)malloc(G * sizeof(int64_t));
        if (!fill) return -1;
        memcpy(fill, start, G * sizeof(int64_t));
        for (int64_t r = 0; r < n; ++r) {
          if (!is_valid(c, r)) continue;
          med_item m;
          m.value = double_at(c, r);
          m.key   = (is_float_t(c->type) && isnan(m.value)) ? ~0ull : ordered_key(c, r);
          it[fill[gid[r]]++] = m;
        }
        for (size_t g = 0; g < G; ++g) {
          int64_t k = start[g + 1] - start[g];
          o->valid[g] = k > 0;
          if (k == 0) continue;
          qsort(it + start[g], (size_t)k, sizeof(med_item), med_cmp);
          double a = it[start[g] + (k - 1) / 2].value, b = it[start[g] + k / 2].value;
          o->f64[g] = (k & 1) ? a : 0.5 * a + 0.5 * b;
        }
        free(start);
        free(it);
        free(fill);
        break;
      }
      default: return -1;
    }
    free(cnt);
  }
  free(gid);
  return ng;
}

void orc_free_agg_out(orc_agg_out* o, int na)
{
  for (int a = 0; a < na; ++a) {
    free(o[a].bits);
    free(o[a].f64);
    free(o[a].valid);
  }
}

/* ---------------------------------------------------------------------------------------- */
/* murmur3 partition                                                                         */
/* ---------------------------------------------------------------------------------------- */
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

uint32_t orc_murmur3_32(const uint8_t* p, int64_t len, uint32_t seed)
{
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  uint32_t h        = seed;
  int64_t nblocks   = len / 4;
  for (int64_t i = 0; i < nblocks; ++i) {
    uint32_t k = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
                 ((uint32_t)p[4 * i + 3] << 24);
    k *= c1;
    k = rotl32(k, 15);
    k *= c2;
    h ^= k;
    h = rotl32(h, 13);
    h = h * 5 + 0xe6546b64u;
  }
  const uint8_t* tail = p + nblocks * 4;
  uint32_t k1         = 0;
  switch (len & 3) {
    case 3: k1 ^= (uint32_t)tail[2] << 16; /* fall through */
    case 2: k1 ^= (uint32_t)tail[1] << 8;  /* fall through */
    case 1:
      k1 ^= tail[0];
      k1 *= c1;
      k1 = rotl32(k1, 15);
      k1 *= c2;
      h ^= k1;
  }
  h ^= (uint32_t)len;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

static inline uint32_t hash_combine32(uint32_t lhs, uint32_t rhs)
{
  return lhs ^ (rhs + 0x9e3779b9u + (lhs << 6) + (lhs >> 2));
}

uint32_t orc_row_hash(const orc_col* cols, int nc, int64_t r, uint32_t seed)
{
  uint32_t h = seed;
  for (int i = 0; i < nc; ++i) {
    const orc_col* c = &cols[i];
    uint32_t e;
    if (!is_valid(c, r)) e = 0xFFFFFFFFu;
    else if (c->width == 0)
      e = orc_murmur3_32((const uint8_t*)c->data + c->offsets[r], c->offsets[r + 1] - c->offsets[r], seed);
    else {
      uint64_t b = bits_at(c, r);
      /* libcudf's MurmurHash3_x86_32<float/double> hashes normalize_nans_and_zeros(value):
       * any NaN -> the canonical quiet NaN, -0.0 -> 0.0 */
      if (c->type == T_FLOAT) {
        float f;
        uint32_t w = (uint32_t)b;
        memcpy(&f, &w, 4);
        if (f != f) b = 0x7fc00000u;
        else if (f == 0.0f) b = 0;
      } else if (c->type == T_DOUBLE) {
        double d;
        memcpy(&d, &b, 8);
        if (d != d) b = 0x7ff8000000000000ull;
        else if (d == 0.0) b = 0;
      }
      uint8_t buf[8];
      memcpy(buf, &b, 8); /* little endian: the value's own bytes */
      e = orc_murmur3_32(buf, c->width, seed);
    }
    h = hash_combine32(h, e);
  }
  return h;
}

void orc_row_hashes(const orc_col* cols, int nc, int64_t n, uint32_t seed, uint32_t* out)
{
  for (int64_t r = 0; r < n; ++r) out[r] = orc_row_hash(cols, nc, r, seed);
}

/* perm[i] = input row placed at output position i; offsets has nparts+1 entries */
int orc_partition(const orc_col* cols, int nc, int64_t n, int32_t nparts, uint32_t seed, int64_t* perm,
                  int64_t* offsets)
{
  uint32_t* p = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
  if (!p) return -1;
  for (int32_t i = 0; i <= nparts; ++i) offsets[i] = 0;
  for (int64_t r = 0; r < n; ++r) {
    p[r] = orc_row_hash(cols, nc, r, seed) % (uint32_t)nparts;
    offsets[p[r] + 1]++;
  }
  for (int32_t i = 0; i < nparts; ++i) offsets[i + 1] += offsets[i];
  int64_t* cur = (int64_t*)malloc((size_t)nparts * sizeof(int64_t));
  if (!cur) return -1;
  memcpy(cur, offsets, (size_t)nparts * sizeof(int64_t));
  for (int64_t r = 0; r < n; ++r) perm[cur[p[r]]++] = r;
  free(cur);
  free(p);
  return 0;
}
