// Groupby kind approximate_median, computed as the reference's GPU variant computes it: the
// variant maps it to cudf::make_median_aggregation (cpp/src/groupby_aggregation.cu:75-77), the
// exact median -- nulls excluded, the middle value of a group's sorted valid values, and for an
// even count 0.5 a + 0.5 b of the two middle values (quantile 0.5 with linear interpolation),
// float64, null for a group without valid values.  Floating-point values sort by value with every
// NaN after every number; -0.0 sorts before 0.0.
//
// Not on the hot path, so it is built from the library's own pieces: the groupby gives the
// groups (keys, and every other aggregation of the call); a join of the input rows with the
// groups' keys (nulls equal) gives each row its group index; two stable radix sorts (rocPRIM)
// order the (group, value) pairs by value and then by group; each group's median is read at
// its middle.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <string>
#include <vector>

#include "ldf_agg.hpp"
#include "ldf_common.hpp"
#include "ldf_device.hpp"

namespace ldf {

void run_groupby(const ldf_table_view* tv, const int32_t* key_idx, int32_t nkeys, const ldf_agg* aggs, int32_t naggs,
                 Alloc& a, ldf_table_out* out);

constexpr uint64_t kMedNaNKey = ~0ull;  // every NaN after every number

__global__ void k_med_iota(int32_t* out, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (int32_t)i;
}

// sort keys: the order-preserving bits of each valid value (group index as payload); a null
// value gets the group index G so it sorts after every group
__global__ void k_med_keys(DCol v, const int32_t* __restrict__ gidx, int64_t n, uint32_t G,
                           unsigned long long* __restrict__ key, uint32_t* __restrict__ grp)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const bool ok    = valid_at(v.valid, i);
    const uint64_t b = load_bits(v.data, v.width, i);
    uint64_t k       = 0;
    if (ok) k = (is_float_type(v.type) && isnan(bits_as_double(b, v.type, v.width))) ? kMedNaNKey : ordered_of(b, v.type, v.width);
    key[i] = k;
    grp[i] = ok ? (uint32_t)gidx[i] : G;
  }
}

// first and one-past-last sorted position of every group that has valid values
__global__ void k_med_bounds(const uint32_t* __restrict__ grp, int64_t n, uint32_t G, int64_t* __restrict__ start,
                             int64_t* __restrict__ end)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t g = grp[i];
    if (g >= G) continue;
    if (i == 0 || grp[i - 1] != g) start[g] = i;
    if (i == n - 1 || grp[i + 1] != g) end[g] = i + 1;
  }
}

__device__ __forceinline__ double med_value(uint64_t k, int type)
{
  if (k == kMedNaNKey) return __longlong_as_double(0x7ff8000000000000ll);
  if (is_float_type(type)) {
    const uint64_t db = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)db);
  }
  if (is_unsigned_type(type)) return (double)k;
  return (double)(int64_t)(k ^ 0x8000000000000000ull);
}

// one thread per group, one wave per 64 groups: the median and the validity word
__global__ void k_med_result(const unsigned long long* __restrict__ key, const int64_t* __restrict__ start,
                             const int64_t* __restrict__ end, int64_t G, int type, double* __restrict__ out,
                             uint64_t* __restrict__ valid, unsigned long long* __restrict__ nulls)
{
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t g0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) & ~(int64_t)(kWave - 1); g0 < G;
       g0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = g0 + lane;
    bool ok         = false;
    if (g < G) {
      const int64_t s = start[g];
      ok              = s >= 0;
      double r        = 0.0;
      if (ok) {
        const int64_t c = end[g] - s;
        const double x  = med_value(key[s + (c - 1) / 2], type);
        const double y  = med_value(key[s + c / 2], type);
        r               = (c & 1) ? x : 0.5 * x + 0.5 * y;
      }
      out[g] = r;
    }
    const uint64_t m = __ballot(ok);
    if (lane == 0) {
      valid[g0 / 64] = m;
      const int64_t in_word = G - g0 < 64 ? G - g0 : 64;
      const unsigned long long nn = (unsigned long long)(in_word - __popcll(m));
      if (nn) atomicAdd(nulls, nn);
    }
  }
}

static void median_column(const ldf_table_view* tv, const int32_t* key_idx, int32_t nkeys, int32_t in_col,
                          const ldf_table_view& groups, int64_t G, Alloc& a, ldf_column* res)
{
  hipStream_t s          = a.stream();
  const ldf_column_view& vc = tv->cols[in_col];
  const int64_t n        = vc.length;
  alloc_column(LDF_TYPE_DOUBLE, G, true, 0, a, res);
  if (G == 0) return;
  // each input row's group: INNER join of the rows with the group keys, nulls equal (every row
  // matches exactly its group)
  std::vector<int32_t> gkeys(nkeys);
  for (int i = 0; i < nkeys; ++i) gkeys[i] = i;
  const int32_t lhs_out = in_col, rhs_out = nkeys;
  ldf_column jc[2];
  ldf_table_out jout{2, jc};
  const int rc = ldf_join(tv, &groups, key_idx, gkeys.data(), nkeys, LDF_JOIN_INNER, 1, &lhs_out, 1, &rhs_out, 1, a.raw(),
                          (void*)s, &jout);
  if (rc != LDF_OK) throw Error(rc, std::string("approximate_median: ") + ldf_last_error());
  struct Free {
    ldf_table_out* t;
    Alloc* a;
    ~Free()
    {
      for (int i = 0; i < t->ncols; ++i) free_column(&t->cols[i], *a);
    }
  } free_join{&jout, &a};
  LDF_EXPECTS(jc[0].length == n, "approximate_median: every row must fall in exactly one group");
  DCol v = to_dcol(vc);
  v.data  = jc[0].data;
  v.valid = jc[0].validity;

  Buf<unsigned long long> k0(a, n), k1(a, n);
  Buf<uint32_t> g0(a, n), g1(a, n);
  hipLaunchKernelGGL(k_med_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, v, (const int32_t*)jc[1].data, n, (uint32_t)G,
                     k0.get(), g0.get());
  LDF_HIP_TRY(hipGetLastError());
  unsigned gbits = 1;
  while (gbits < 32 && ((uint64_t)G >> gbits) != 0) ++gbits;  // G itself (the null marker) fits
  size_t t0 = 0, t1 = 0;
  LDF_HIP_TRY(rocprim::radix_sort_pairs(nullptr, t0, k0.get(), k1.get(), g0.get(), g1.get(), (size_t)n, 0, 64, s));
  LDF_HIP_TRY(rocprim::radix_sort_pairs(nullptr, t1, g1.get(), g0.get(), k1.get(), k0.get(), (size_t)n, 0, gbits, s));
  Buf<unsigned char> tmp(a, std::max<size_t>(std::max(t0, t1), 1));
  size_t tb = std::max(t0, t1);
  // by value, then (stable) by group: sorted by (group, value)
  LDF_HIP_TRY(rocprim::radix_sort_pairs(tmp.get(), tb, k0.get(), k1.get(), g0.get(), g1.get(), (size_t)n, 0, 64, s));
  tb = std::max(t0, t1);
  LDF_HIP_TRY(rocprim::radix_sort_pairs(tmp.get(), tb, g1.get(), g0.get(), k1.get(), k0.get(), (size_t)n, 0, gbits, s));
  Buf<int64_t> start(a, G), end(a, G);
  fill_u64(reinterpret_cast<uint64_t*>(start.get()), ~0ull, G, s);
  hipLaunchKernelGGL(k_med_bounds, dim3(grid_for(n, 256)), dim3(256), 0, s, g0.get(), n, (uint32_t)G, start.get(),
                     end.get());
  LDF_HIP_TRY(hipGetLastError());
  Buf<unsigned long long> nulls(a, 1);
  fill_u64(reinterpret_cast<uint64_t*>(nulls.get()), 0, 1, s);
  hipLaunchKernelGGL(k_med_result, dim3(grid_for((G + 63) / 64 * 64, 256)), dim3(256), 0, s, k0.get(), start.get(),
                     end.get(), G, vc.type_id, static_cast<double*>(res->data), reinterpret_cast<uint64_t*>(res->validity),
                     nulls.get());
  LDF_HIP_TRY(hipGetLastError());
  unsigned long long nn = 0;
  copy_to_host(&nn, nulls.get(), sizeof(nn), s);
  res->null_count = (int64_t)nn;
}

// ldf_groupby with one or more approximate_median aggregations: the groupby of every other
// aggregation (plus a row count, so the call always has one) gives the groups; each median
// column is computed over them and placed at its position.
void run_groupby_with_medians(const ldf_table_view* tv, const int32_t* key_idx, int32_t nkeys, const ldf_agg* aggs,
                              int32_t naggs, Alloc& a, ldf_table_out* out)
{
  LDF_EXPECTS(tv != nullptr && out != nullptr && out->cols != nullptr && key_idx != nullptr, "groupby: NULL argument");
  LDF_EXPECTS(nkeys >= 1, "must pass at least one key column");
  LDF_EXPECTS(out->ncols >= nkeys + naggs, "groupby: output table has too few column slots");
  std::vector<ldf_agg> rest;
  for (int i = 0; i < naggs; ++i) {
    if (aggs[i].kind != LDF_AGG_APPROX_MEDIAN) {
      rest.push_back(aggs[i]);
      continue;
    }
    LDF_EXPECTS(aggs[i].in_col >= 0 && aggs[i].in_col < tv->ncols, "groupby: aggregation input column out of range");
    const int32_t t = tv->cols[aggs[i].in_col].type_id;
    if (type_width(t) == 0 || t == LDF_TYPE_BOOL)
      throw Error(LDF_E_UNSUPPORTED, "groupby: approximate_median of a string or bool column");
  }
  rest.push_back(ldf_agg{0, LDF_AGG_COUNT_ALL, LDF_COUNT_ALL, 0});
  std::vector<ldf_column> tmp(nkeys + rest.size());
  ldf_table_out tout{(int32_t)tmp.size(), tmp.data()};
  run_groupby(tv, key_idx, nkeys, rest.data(), (int32_t)rest.size(), a, &tout);
  const int64_t G = tmp[0].length;
  // the groups' keys and their index 0..G-1: the build side of the row -> group join
  Buf<int32_t> gid(a, G > 0 ? G : 1);
  if (G > 0) {
    hipLaunchKernelGGL(k_med_iota, dim3(grid_for(G, 256)), dim3(256), 0, a.stream(), gid.get(), G);
    LDF_HIP_TRY(hipGetLastError());
  }
  std::vector<ldf_column_view> gv(nkeys + 1);
  for (int i = 0; i < nkeys; ++i) {
    const ldf_column& c = tmp[i];
    gv[i]               = ldf_column_view{c.type_id, 0, c.length, c.data, c.validity, c.offsets, c.null_count};
  }
  gv[nkeys]             = ldf_column_view{LDF_TYPE_INT32, 0, G, gid.get(), nullptr, nullptr, 0};
  const ldf_table_view groups{nkeys + 1, gv.data()};
  std::vector<ldf_column> med(naggs);
  try {
    for (int i = 0; i < naggs; ++i)
      if (aggs[i].kind == LDF_AGG_APPROX_MEDIAN) median_column(tv, key_idx, nkeys, aggs[i].in_col, groups, G, a, &med[i]);
  } catch (...) {
    for (auto& c : med)
      if (c.data) free_column(&c, a);
    for (auto& c : tmp) free_column(&c, a);
    throw;
  }
  for (int i = 0; i < nkeys; ++i) out->cols[i] = tmp[i];
  int r = nkeys;
  for (int i = 0; i < naggs; ++i) out->cols[nkeys + i] = aggs[i].kind == LDF_AGG_APPROX_MEDIAN ? med[i] : tmp[r++];
  free_column(&tmp[r], a);  // the row count added above
  out->ncols = nkeys + naggs;
}

}  // namespace ldf
