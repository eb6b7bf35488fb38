// Host driver of the hash group-by: validates the request, flattens the
// aggregations into per-column accumulators, sizes the hash table, launches
// the kernels and assembles the result columns.
//
// Reference behaviour restated here:
//   validation        cpp/src/groupby/groupby.cu:175-221
//   hash eligibility  cpp/src/groupby/hash/groupby.cu:73-84, 635-651
//   flattening        cpp/src/groupby/hash/groupby.cu:113-181, 351-383
//                     (MEAN and VARIANCE/STD reduce to SUM + COUNT_VALID)
//   target types      cpp/include/cudf/detail/aggregation/aggregation.hpp:1211-1331
//   result nullability cpp/src/groupby/hash/groupby.cu:431-433 (SUM/MIN/MAX are
//                     nullable iff the values have nulls, VARIANCE/STD always,
//                     COUNT never) and the DIV of MEAN (hash/groupby.cu:298-304)
//   result order      cpp/src/groupby/common/utils.hpp:32-60 (request order)
#include "groupby.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <map>
#include <optional>
#include <unordered_map>

#include "host_plan.hpp"
#include "kernels.hpp"

namespace mig {

double* thread_stats()
{
  static thread_local double stats[STAT_COUNT] = {0};
  return stats;
}

namespace {
std::atomic<int> g_phase_timing{0};  // 0 off, 1 every phase, 2 the main aggregation kernel only
std::atomic<int> g_strategy{MIG_STRATEGY_AUTO};

// Partitioned path: LDS hash slots and byte budget of the aggregating
// workgroup (two 512-thread workgroups fit in a CU's 160 KB of LDS).
constexpr int32_t RPA_LDS_SLOTS      = 4096;
constexpr int32_t RPA_LDS_BUDGET     = 80 * 1024;  // two aggregating workgroups per CU (160 KB)
constexpr int RPA_MAX_LOG_BINS       = 16;  // 256 x 256 through the two scatter passes
constexpr int RPA_MAX_LOG_FANOUT     = 8;
constexpr int32_t RPA_STAGE_BUDGET   = 56 * 1024;  // LDS staging of a scatter tile (3 workgroups per CU)
// Global path, low cardinality: LDS partial tables of up to 2048 slots (3/4
// fillable), used when the estimate is at most LDS_PREAGG_MAX_GROUPS groups.
constexpr int32_t LDS_PREAGG_SLOTS      = 2048;
constexpr int32_t LDS_PREAGG_MAX_GROUPS = 1024;
constexpr int32_t LDS_PREAGG_BUDGET     = 96 * 1024;

// hipEvents marking phase boundaries on the call's stream.
struct phase_timer {
  static constexpr int N = 16;
  int level;
  bool on;
  hipStream_t stream;
  hipEvent_t ev[N];
  bool recorded[N] = {false};
  explicit phase_timer(hipStream_t s, bool allowed = true)
    : level(allowed ? g_phase_timing.load() : 0), on(level > 0), stream(s)
  {
    if (!on) return;
    static thread_local std::unordered_map<int, std::vector<hipEvent_t>> pool;
    int dev = 0;
    MIG_HIP_TRY(hipGetDevice(&dev));
    auto& v = pool[dev];
    if (v.empty()) {
      v.resize(N);
      for (auto& e : v) MIG_HIP_TRY(hipEventCreate(&e));
    }
    for (int i = 0; i < N; i++) ev[i] = v[i];
  }
  void mark(int i)
  {
    // level 2 brackets only the accumulate kernel of either strategy (marks 2, 3 and 12, 13)
    if (!on || (level == 2 && i != 2 && i != 3 && i != 12 && i != 13)) return;
    MIG_HIP_TRY(hipEventRecord(ev[i], stream));
    recorded[i] = true;
  }
  double between(int a, int b)
  {
    if (!on || !recorded[a] || !recorded[b]) return 0.0;
    float ms = 0.f;
    MIG_HIP_TRY(hipEventElapsedTime(&ms, ev[a], ev[b]));
    return ms;
  }
};
}  // namespace

void set_phase_timing(int level) { g_phase_timing.store(level); }
void set_strategy(int s) { g_strategy.store(s); }

// ---------------------------------------------------------------------------
void* device_allocator::allocate(size_t bytes)
{
  if (borrowed) fail(MIG_ERROR_LOGIC, "the allocation-free group-by cannot allocate an output buffer");
  if (bytes == 0) bytes = 1;
  void* p = nullptr;
  if (has_mr) {
    p = mr.allocate(mr.ctx, bytes, stream);
    if (!p) fail(MIG_ERROR_ALLOC, "memory resource returned null for " + std::to_string(bytes) + " bytes");
  } else {
    hipError_t e = hipMallocAsync(&p, bytes, stream);
    if (e != hipSuccess) fail(MIG_ERROR_ALLOC, std::string("hipMallocAsync failed: ") + hipGetErrorString(e));
  }
  return p;
}

void device_allocator::deallocate(void* p, size_t bytes)
{
  if (!p || borrowed) return;
  if (bytes == 0) bytes = 1;
  if (has_mr) mr.deallocate(mr.ctx, p, bytes, stream);
  else (void)hipFreeAsync(p, stream);
}

namespace {

struct device_props {
  int num_cus = 256;
};

device_props const& props()
{
  static std::mutex mu;
  static std::unordered_map<int, device_props> cache;
  int dev = 0;
  MIG_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  device_props p;
  MIG_HIP_TRY(hipDeviceGetAttribute(&p.num_cus, hipDeviceAttributeMultiprocessorCount, dev));
  // Keep freed scratch in the stream-ordered pool instead of returning it to
  // the driver at every synchronisation: a group-by's table and partition
  // buffers are gigabytes, and re-mapping them per call costs far more than
  // the kernels.
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t keep = ~0ull;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  return cache.emplace(dev, p).first->second;
}

// Device scratch blocks kept across calls, per device.  A call's hash table,
// accumulators and partition buffers total gigabytes; mapping fresh device
// memory for them on every call costs ~40 ms per GB, far more than the
// kernels, so freed blocks are cached and reused by later calls (best fit
// within 2x of the request).  Every call synchronises its stream before its
// blocks return to the cache, so a reused block is never still in use.
struct scratch_cache {
  std::mutex mu;
  std::unordered_map<int, std::multimap<size_t, void*>> free_blocks;

  static scratch_cache& get()
  {
    static scratch_cache c;
    return c;
  }
  void* take(int dev, size_t bytes)
  {
    {
      std::lock_guard<std::mutex> lock(mu);
      auto& m = free_blocks[dev];
      auto it = m.lower_bound(bytes);
      if (it != m.end() && it->first <= 2 * bytes + (1u << 20)) {
        void* p = it->second;
        m.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      // release cached blocks of this device and try once more
      release_all(dev);
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess)
      fail(MIG_ERROR_ALLOC, "scratch allocation of " + std::to_string(bytes) + " bytes failed");
    return p;
  }
  void give(int dev, void* p, size_t bytes)
  {
    std::lock_guard<std::mutex> lock(mu);
    free_blocks[dev].emplace(bytes, p);
  }
  void release_all(int dev)
  {
    std::lock_guard<std::mutex> lock(mu);
    for (auto& kv : free_blocks[dev]) (void)hipFree(kv.second);
    free_blocks[dev].clear();
  }
};

// Scratch memory of one call; blocks go back to the cache when the call ends
// (after its stream has been synchronised).  Given an arena (the
// allocation-free entry point's fixed device array), blocks are carved from
// it instead and a request past its end fails with MIG_ERROR_OVERFLOW.
struct scratch {
  hipStream_t stream;
  int dev = 0;
  std::vector<std::pair<void*, size_t>> blocks;
  uint8_t* arena     = nullptr;
  size_t arena_bytes = 0;
  size_t arena_used  = 0;
  explicit scratch(hipStream_t s, uint8_t* a = nullptr, size_t a_bytes = 0)
    : stream(s), arena(a), arena_bytes(a_bytes)
  {
    MIG_HIP_TRY(hipGetDevice(&dev));
  }
  template <typename T>
  T* get(size_t count)
  {
    size_t bytes = std::max<size_t>(256, ((count * sizeof(T)) + 255) & ~size_t(255));
    if (arena) {
      if (bytes > arena_bytes - arena_used)
        fail(MIG_ERROR_OVERFLOW, "the call needs more than the " + std::to_string(arena_bytes >> 20) +
                                   " MiB of fixed device scratch (" + std::to_string((arena_used + bytes) >> 20) +
                                   " MiB so far)");
      void* p = arena + arena_used;
      arena_used += bytes;
      return static_cast<T*>(p);
    }
    void* p = scratch_cache::get().take(dev, bytes);
    blocks.emplace_back(p, bytes);
    return static_cast<T*>(p);
  }
  // Arena position, and a return to it (the blocks above it are free once
  // the work that used them has finished); no-ops for cached blocks.
  size_t mark() const { return arena_used; }
  void release_to(size_t m)
  {
    if (arena) arena_used = m;
  }
  ~scratch()
  {
    (void)hipStreamSynchronize(stream);
    for (auto& b : blocks) scratch_cache::get().give(dev, b.first, b.second);
  }
};

// Small pinned host buffer per thread for parameter uploads and readbacks.
struct pinned_host {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t need)
  {
    if (need > bytes) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      MIG_HIP_TRY(hipHostMalloc(&p, need, hipHostMallocDefault));
      bytes = need;
    }
    return p;
  }
  ~pinned_host()
  {
    if (p) (void)hipHostFree(p);
  }
};

pinned_host& thread_pinned()
{
  static thread_local pinned_host h;
  return h;
}

constexpr int64_t SAMPLE_ROWS = 65536;
// Host bytes a call stages parameter blocks and readbacks in: two gb_params
// blocks, then the readback area.
constexpr size_t HOST_STAGING_BYTES = 2 * sizeof(gb_params) + 4 * SAMPLE_ROWS + 64;

// The allocation-free entry point's host staging: static per thread
// (pageable memory, so copies from it complete before hipMemcpyAsync returns).
uint8_t* host_staging()
{
  alignas(64) static thread_local uint8_t buf[HOST_STAGING_BYTES];
  return buf;
}

}  // namespace

// Planning helpers shared with the batched driver (declared in host_plan.hpp).

void* scratch_take(int dev, size_t bytes) { return scratch_cache::get().take(dev, bytes); }
int num_cus_of_device() { return props().num_cus; }
void scratch_give(int dev, void* p, size_t bytes) { scratch_cache::get().give(dev, p, bytes); }

bool same_view(mig_column_view const& a, mig_column_view const& b)
{
  return a.type_id == b.type_id && a.size == b.size && a.offset == b.offset && a.data == b.data &&
         a.null_mask == b.null_mask && a.offsets == b.offsets;
}

int32_t null_count_of(mig_column_view const& v, hipStream_t s)
{
  if (v.null_mask == nullptr) return 0;
  if (v.null_count >= 0) return v.null_count;
  dcol c{};
  c.mask        = v.null_mask;
  c.mask_offset = v.offset;
  uint32_t* d   = nullptr;
  MIG_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(uint32_t), s));
  MIG_HIP_TRY(hipMemsetAsync(d, 0, sizeof(uint32_t), s));
  launch_count_nulls(c, v.size, d, s);
  uint32_t valid = 0;
  MIG_HIP_TRY(hipMemcpyAsync(&valid, d, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MIG_HIP_TRY(hipStreamSynchronize(s));
  MIG_HIP_TRY(hipFreeAsync(d, s));
  return v.size - int32_t(valid);
}

dcol make_dcol(mig_column_view const& v, bool use_mask)
{
  dcol c{};
  c.storage     = storage_of(v.type_id);
  c.mask        = use_mask ? v.null_mask : nullptr;
  c.mask_offset = v.offset;
  if (c.storage == ST_STRING) {
    c.data      = static_cast<uint8_t const*>(v.data);
    c.offsets64 = v.offsets_type_id == MIG_TYPE_INT64;
    int w       = c.offsets64 ? 8 : 4;
    c.offsets   = static_cast<uint8_t const*>(v.offsets) + int64_t(w) * v.offset;
  } else {
    c.data = static_cast<uint8_t const*>(v.data) + int64_t(storage_width(c.storage)) * v.offset;
  }
  return c;
}

bool is_numeric_type(int32_t t)
{
  return (t >= MIG_TYPE_INT8 && t <= MIG_TYPE_FLOAT64) || t == MIG_TYPE_BOOL8;
}
bool is_chrono_type(int32_t t) { return t >= MIG_TYPE_TIMESTAMP_DAYS && t <= MIG_TYPE_DURATION_NANOSECONDS; }
// fixed-point types with a 32- or 64-bit representation (DECIMAL128 is not supported)
bool is_decimal_type(int32_t t) { return t == MIG_TYPE_DECIMAL32 || t == MIG_TYPE_DECIMAL64; }

char const* kind_name(int32_t k)
{
  switch (k) {
    case MIG_AGG_SUM: return "SUM";
    case MIG_AGG_PRODUCT: return "PRODUCT";
    case MIG_AGG_MIN: return "MIN";
    case MIG_AGG_MAX: return "MAX";
    case MIG_AGG_COUNT_VALID: return "COUNT_VALID";
    case MIG_AGG_COUNT_ALL: return "COUNT_ALL";
    case MIG_AGG_SUM_OF_SQUARES: return "SUM_OF_SQUARES";
    case MIG_AGG_MEAN: return "MEAN";
    case MIG_AGG_VARIANCE: return "VARIANCE";
    case MIG_AGG_STD: return "STD";
    case MIG_AGG_ARGMIN: return "ARGMIN";
    case MIG_AGG_ARGMAX: return "ARGMAX";
    case MIG_AGG_NUNIQUE: return "NUNIQUE";
    case MIG_AGG_NTH_ELEMENT: return "NTH_ELEMENT";
    case MIG_AGG_MEDIAN: return "MEDIAN";
    case MIG_AGG_QUANTILE: return "QUANTILE";
    default: return "UNKNOWN";
  }
}

// Result type of (source type, aggregation): aggregation.hpp:1211-1331.
int32_t target_type_of(int32_t src, int32_t kind)
{
  switch (kind) {
    case MIG_AGG_MIN:
    case MIG_AGG_MAX:
    case MIG_AGG_NTH_ELEMENT: return src;
    case MIG_AGG_COUNT_VALID:
    case MIG_AGG_COUNT_ALL:
    case MIG_AGG_ARGMIN:  // size_type row index (aggregation.hpp:1346-1357)
    case MIG_AGG_ARGMAX:
    case MIG_AGG_NUNIQUE: return MIG_TYPE_INT32;  // size_type count (group_nunique.cu:90-91)
    case MIG_AGG_MEAN:
      if (is_decimal_type(src)) return src;  // fixed point keeps its type (aggregation.hpp:1265-1272)
      return MIG_TYPE_FLOAT64;
    case MIG_AGG_VARIANCE:
    case MIG_AGG_STD:
    case MIG_AGG_MEDIAN:    // always double (aggregation.hpp:1334-1344)
    case MIG_AGG_QUANTILE: return MIG_TYPE_FLOAT64;
    case MIG_AGG_SUM:
    case MIG_AGG_SUM_OF_SQUARES:
    case MIG_AGG_PRODUCT:  // aggregation.hpp:1276-1306
      if (src == MIG_TYPE_FLOAT32 || src == MIG_TYPE_FLOAT64) return src;
      if (kind == MIG_AGG_SUM && is_decimal_type(src)) return src;  // aggregation.hpp:1286-1293
      return MIG_TYPE_INT64;
    default: return MIG_TYPE_EMPTY;
  }
}

// Validity of (source type, aggregation) on this hash path.  Mirrors the
// type rules of cudf's is_valid_aggregation for the kinds implemented here.
void check_supported(int32_t src, int32_t kind)
{
  if (src == MIG_TYPE_DECIMAL128)
    fail(MIG_ERROR_DATA_TYPE, std::string("DECIMAL128 values are not supported (aggregation ") + kind_name(kind) + ")");
  if (kind == MIG_AGG_PRODUCT && is_decimal_type(src))  // sum_tests.cpp:199-202: a logic_error
    fail(MIG_ERROR_LOGIC, "PRODUCT of fixed-point values is not supported");
  bool ok = false;
  switch (kind) {
    case MIG_AGG_COUNT_VALID:
    case MIG_AGG_COUNT_ALL: ok = src != MIG_TYPE_EMPTY; break;
    case MIG_AGG_MIN:
    case MIG_AGG_MAX:
    case MIG_AGG_ARGMIN:
    case MIG_AGG_ARGMAX:
      ok = is_numeric_type(src) || is_chrono_type(src) || is_decimal_type(src) || src == MIG_TYPE_STRING;
      break;
    case MIG_AGG_SUM:
    case MIG_AGG_MEAN: ok = is_numeric_type(src) || is_decimal_type(src); break;
    case MIG_AGG_SUM_OF_SQUARES:
    case MIG_AGG_PRODUCT:
    case MIG_AGG_VARIANCE:
    case MIG_AGG_STD: ok = is_numeric_type(src); break;
    case MIG_AGG_NUNIQUE:
    case MIG_AGG_NTH_ELEMENT:
      ok = is_numeric_type(src) || is_chrono_type(src) || is_decimal_type(src) || src == MIG_TYPE_STRING;
      break;
    // arithmetic values only (group_quantiles.cu quantiles_functor)
    case MIG_AGG_MEDIAN:
    case MIG_AGG_QUANTILE: ok = is_numeric_type(src); break;
    default:
      fail(MIG_ERROR_LOGIC, std::string("aggregation ") + kind_name(kind) +
                              " is not supported by the hash group-by");
  }
  if (!ok)
    fail(MIG_ERROR_DATA_TYPE, std::string("unsupported type ") + std::to_string(src) +
                                " for aggregation " + kind_name(kind));
}

// Per distinct value column: which accumulators it needs.
struct value_plan {
  mig_column_view view{};
  int32_t storage   = ST_NONE;
  bool has_nulls    = false;
  bool want_sum = false, want_sumsq = false, want_min = false, want_max = false;
  bool want_count = false, want_m2 = false, want_argmin = false, want_argmax = false, want_prod = false;
  int32_t f_sum = -1, f_sumsq = -1, f_min = -1, f_max = -1, f_count = -1, f_m2 = -1;
  int32_t f_argmin = -1, f_argmax = -1, f_prod = -1;
  bool want_nunique[2] = {false, false};  // [null_handling == INCLUDE]
  int32_t f_nunique[2] = {-1, -1};
  std::vector<std::array<int32_t, 3>> nth;  // NTH_ELEMENT: (n, include nulls, field)
  int32_t nth_field(int32_t n, bool include) const
  {
    for (auto const& e : nth)
      if (e[0] == n && e[1] == int32_t(include)) return e[2];
    return -1;
  }
};

int32_t sum_field_kind(int32_t st)
{
  if (st == ST_F32) return FK_SUM_F32;
  if (st == ST_F64) return FK_SUM_F64;
  return FK_SUM_I64;
}
int32_t sumsq_field_kind(int32_t st)
{
  if (st == ST_F32) return FK_SUMSQ_F32;
  if (st == ST_F64) return FK_SUMSQ_F64;
  return FK_SUMSQ_I64;
}
int32_t minmax_field_kind(int32_t st, bool is_min)
{
  switch (st) {
    case ST_I8:
    case ST_I16:
    case ST_I32: return is_min ? FK_MIN_S32 : FK_MAX_S32;
    case ST_I64: return is_min ? FK_MIN_S64 : FK_MAX_S64;
    case ST_U8:
    case ST_U16:
    case ST_U32:
    case ST_BOOL: return is_min ? FK_MIN_U32 : FK_MAX_U32;
    case ST_U64: return is_min ? FK_MIN_U64 : FK_MAX_U64;
    case ST_F32: return is_min ? FK_MIN_F32 : FK_MAX_F32;
    default: return is_min ? FK_MIN_F64 : FK_MAX_F64;
  }
}
int32_t field_width(int32_t fk) { return field_kind_width(fk); }

uint64_t next_pow2(uint64_t x)
{
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Distinct values D of a population from `d` distinct among `n` samples,
// assuming equally frequent values: solve D (1 - exp(-n / D)) = d.
double estimate_distinct(double n, double d, double population)
{
  if (d >= n) return population;
  double lo = d, hi = population;
  if (hi * (1.0 - std::exp(-n / hi)) <= d) return population;
  for (int it = 0; it < 100; it++) {
    double mid = 0.5 * (lo + hi);
    double f   = mid * (1.0 - std::exp(-n / mid));
    if (f < d) lo = mid;
    else hi = mid;
  }
  return hi;
}

// The larger of the equal-frequency estimate and Chao's lower bound
// d + f1 (f1 - 1) / (2 (f2 + 1)) from the values seen once (f1) and twice
// (f2) in the sample (A. Chao, Scand. J. Statist. 11 (1984) 265-270).  The
// two agree for equally frequent values; under skew (Zipf-like keys) the hot
// values fill the sample and the equal-frequency model reads the tail as
// absent, where the singletons still show it.
double estimate_distinct_skewed(double n, double d, double f1, double f2, double population, bool* lower_bound)
{
  double const chao    = d + f1 * std::max(f1 - 1.0, 0.0) / (2.0 * (f2 + 1.0));
  double const uniform = estimate_distinct(n, d, population);
  if (lower_bound) *lower_bound = chao > 1.5 * uniform;  // skewed: the tail may hold many more
  return std::min(population, std::max(uniform, chao));
}

namespace {

owned_column make_owned_fixed(device_allocator& A, int32_t type_id, int64_t n, bool with_mask)
{
  owned_column oc;
  oc.col.type_id    = type_id;
  oc.col.size       = int32_t(n);
  oc.data_bytes     = size_t(type_width(type_id)) * size_t(n);
  oc.col.data       = n > 0 ? A.allocate(oc.data_bytes) : nullptr;
  oc.col.data_bytes = int64_t(oc.data_bytes);
  if (with_mask) {
    oc.mask_bytes    = std::max<size_t>(8, size_t((n + 63) / 64) * 8);  // whole 64-bit words
    oc.col.null_mask = static_cast<uint32_t*>(A.allocate(oc.mask_bytes));
    if (n == 0) MIG_HIP_TRY(hipMemsetAsync(oc.col.null_mask, 0, oc.mask_bytes, A.stream));
  }
  return oc;
}

owned_column make_empty_column(device_allocator& A, int32_t type_id)
{
  owned_column oc;
  oc.col.type_id = type_id;
  if (type_id == MIG_TYPE_STRING) {
    oc.col.offsets_type_id = MIG_TYPE_INT32;
    oc.offsets_bytes       = sizeof(int32_t);
    oc.col.offsets         = A.allocate(oc.offsets_bytes);
    MIG_HIP_TRY(hipMemsetAsync(oc.col.offsets, 0, oc.offsets_bytes, A.stream));
  }
  return oc;
}

}  // namespace

// ---------------------------------------------------------------------------
void hash_rows(mig_column_view const* cols, int32_t ncols, uint32_t seed, uint32_t* out,
               hipStream_t stream)
{
  MIG_EXPECTS(ncols >= 0 && ncols <= MAX_KEYS, "too many columns to hash");
  if (ncols == 0) return;
  int64_t const n = cols[0].size;
  for (int i = 0; i < ncols; i++) {
    MIG_EXPECTS(cols[i].size == n, "all columns must have the same number of rows");
    MIG_EXPECTS(cols[i].size >= 0 && cols[i].offset >= 0, "negative column size or offset");
    if (storage_of(cols[i].type_id) == ST_NONE)
      fail(MIG_ERROR_DATA_TYPE, "unsupported column type " + std::to_string(cols[i].type_id) + " for hashing");
  }
  if (n == 0) return;
  dcol h[MAX_KEYS];
  for (int i = 0; i < ncols; i++) h[i] = make_dcol(cols[i], null_count_of(cols[i], stream) > 0);
  scratch S(stream);
  dcol* d = S.get<dcol>(ncols);
  void* pin = thread_pinned().get(sizeof(h));
  std::memcpy(pin, h, sizeof(dcol) * ncols);
  MIG_HIP_TRY(hipMemcpyAsync(d, pin, sizeof(dcol) * ncols, hipMemcpyHostToDevice, stream));
  launch_hash_rows(d, ncols, n, seed, out, stream);
  MIG_HIP_TRY(hipGetLastError());
  MIG_HIP_TRY(hipStreamSynchronize(stream));
}

namespace {

// ---------------------------------------------------------------------------
// One group-by call.  The phases are methods: validation, the empty-input
// shortcut, key and accumulator planning (the reference's flattening), the
// sampled group estimate, then one aggregation strategy — runs of pre-sorted
// keys, the radix-partitioned LDS path or the device-wide table — and the
// output null counts.  Memory comes from the library's scratch cache and the
// caller's memory resource, or, for the allocation-free entry point, from a
// fixed device arena and the caller's output buffers (`call_memory`).
class groupby_call {
 public:
  groupby_call(mig_column_view const* keys, int32_t num_keys, int32_t null_policy,
               mig_aggregation_request const* requests, int32_t num_requests, hipStream_t stream,
               mig_memory_resource const* mr, bool keys_sorted, call_memory const* mem)
    : keys_(keys),
      num_keys_(num_keys),
      null_policy_(null_policy),
      requests_(requests),
      num_requests_(num_requests),
      stream_(stream),
      mr_(mr),
      keys_sorted_(keys_sorted),
      mem_(mem),
      stats_(thread_stats())
  {
  }
  std::unique_ptr<mig_groupby_result> run();

 private:
  struct string_out {  // string MIN / MAX results: (request, aggregation, value column, group rows)
    int r, a, vc;
    int32_t* rows;
  };

  mig_column_view const* keys_;
  int32_t num_keys_;
  int32_t null_policy_;
  mig_aggregation_request const* requests_;
  int32_t num_requests_;
  hipStream_t stream_;
  mig_memory_resource const* mr_;
  bool keys_sorted_;
  call_memory const* mem_;
  double* stats_;
  // created once the request is validated (both touch the device)
  std::optional<scratch> S;
  std::optional<phase_timer> T;

  int64_t N = 0;
  bool needs_sort_ = false;
  std::unique_ptr<mig_groupby_result> res_;
  std::unique_ptr<gb_params> P_ptr_;
  gb_params* Pp_ = nullptr;
  gb_params* dP = nullptr;
  uint8_t* pin = nullptr;
  std::vector<value_plan> vals_;
  std::vector<int> vcol_of_request_;
  int nf_        = 0;
  bool any_arg_  = false;
  double est_    = 0.0;  // estimated groups
  double pop_    = 0.0;  // rows the group-by keeps (estimated)
  uint32_t max_mult_ = 0;  // most sampled rows sharing one key hash
  double sampled_    = 0.0;  // sampled rows the group-by keeps
  bool est_low_      = false;  // the estimate is a lower bound (skewed keys)
  // A key on at least 1/512 of the sampled rows among at most 2M groups: one
  // bin of the partitioned path would hold that key's rows (a straggler
  // workgroup), while the device-wide table of that many groups stays in the
  // last-level cache and workgroup-local tables absorb the hot key.
  bool heavy_hitters() const
  {
    return sampled_ > 0.0 && double(max_mult_) * 512.0 >= sampled_ && max_mult_ >= 16 &&
           est_ <= double(1 << 21);
  }
  int n_str_keys_ = 0;
  std::vector<string_out> string_outs_;
  int borrowed_next_ = 0;  // allocation-free calls: next caller output column

  gb_params& P() { return *Pp_; }
  device_allocator& A() { return res_->alloc; }
  // Nulls of a view: its null_count when known, else counted on the device
  // (in this call's scratch).
  int32_t nulls_of(mig_column_view const& v)
  {
    if (v.null_mask == nullptr) return 0;
    if (v.null_count >= 0 || !mem_) return null_count_of(v, stream_);
    uint32_t* d = S->get<uint32_t>(1);
    MIG_HIP_TRY(hipMemsetAsync(d, 0, sizeof(uint32_t), stream_));
    launch_count_nulls(make_dcol(v, true), v.size, d, stream_);
    uint32_t* h = reinterpret_cast<uint32_t*>(host_staging());
    MIG_HIP_TRY(hipMemcpyAsync(h, d, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipStreamSynchronize(stream_));
    return v.size - int32_t(*h);
  }
  void validate();
  void empty_result();
  void plan_keys();
  void plan_accumulators();
  void upload(int which);
  void estimate_groups();
  owned_column fixed_output(int32_t type_id, int64_t rows_cap, int64_t G, bool with_mask, bool zero_masks);
  void alloc_outputs(int64_t rows_cap, int64_t G, bool zero_masks);
  void release_outputs();
  void gather_general_keys(uint32_t const* rep_rows, int64_t G);
  int64_t aggregate_sorted_runs();
  int64_t aggregate_partitioned();
  int64_t aggregate_global();
  void string_results(int64_t G);
  void finish(int64_t G);
};

// groupby.cu:210-214 and :175-192
void groupby_call::validate()
{
  MIG_EXPECTS(num_keys_ >= 0 && num_keys_ <= MAX_KEYS, "number of key columns out of range");
  MIG_EXPECTS(num_requests_ >= 0, "negative number of requests");
  MIG_EXPECTS(null_policy_ == MIG_NULL_EXCLUDE || null_policy_ == MIG_NULL_INCLUDE, "invalid null policy");
  N = num_keys_ > 0 ? keys_[0].size : 0;
  for (int i = 0; i < num_keys_; i++) {
    MIG_EXPECTS(keys_[i].size >= 0 && keys_[i].offset >= 0, "negative column size or offset");
    MIG_EXPECTS(keys_[i].size == N, "key columns must have the same number of rows");
    int32_t st = storage_of(keys_[i].type_id);
    if (st == ST_NONE)
      fail(MIG_ERROR_DATA_TYPE, "unsupported key column type " + std::to_string(keys_[i].type_id));
  }
  for (int r = 0; r < num_requests_; r++) {
    MIG_EXPECTS(requests_[r].values.size == N,
                "Size mismatch between request values and groupby keys.");
    MIG_EXPECTS(requests_[r].values.offset >= 0, "negative column offset");
    MIG_EXPECTS(requests_[r].num_aggregations >= 0, "negative number of aggregations");
    for (int a = 0; a < requests_[r].num_aggregations; a++) {
      check_supported(requests_[r].values.type_id, requests_[r].aggregations[a].kind);
      // a sort-based aggregation (groupby.cu:78-83): NUNIQUE and NTH_ELEMENT
      // run on the runs of pre-sorted keys, or else (and MEDIAN / QUANTILE
      // always) on the sort-based group-by
      int32_t const k = requests_[r].aggregations[a].kind;
      if (k == MIG_AGG_QUANTILE) {
        mig_aggregation const& ag = requests_[r].aggregations[a];
        MIG_EXPECTS(ag.num_quantiles >= 0 && (ag.num_quantiles == 0 || ag.quantiles != nullptr),
                    "invalid quantile list");
        if (ag.interpolation < MIG_INTERP_LINEAR || ag.interpolation > MIG_INTERP_NEAREST)
          fail(MIG_ERROR_LOGIC, "Invalid interpolation operation for quantiles.");
      }
      needs_sort_ |= k == MIG_AGG_MEDIAN || k == MIG_AGG_QUANTILE ||
                     ((k == MIG_AGG_NUNIQUE || k == MIG_AGG_NTH_ELEMENT) && !keys_sorted_);
    }
  }
  if (mem_) {  // the allocation-free entry point: fixed-width columns on the device-wide table
    MIG_EXPECTS(!needs_sort_ && !keys_sorted_,
                "the allocation-free group-by takes hash aggregations only (no sort-based kinds)");
    for (int i = 0; i < num_keys_; i++)
      MIG_EXPECTS(storage_of(keys_[i].type_id) != ST_STRING, "the allocation-free group-by takes no string keys");
    for (int r = 0; r < num_requests_; r++)
      MIG_EXPECTS(storage_of(requests_[r].values.type_id) != ST_STRING,
                  "the allocation-free group-by takes no string values");
  }
}

// groupby.cu:150-172: typed empty columns
void groupby_call::empty_result()
{
  if (mem_) {
    int c = 0;
    for (int i = 0; i < num_keys_; i++, c++) {
      mig_column& o = mem_->out_keys[i];
      o.type_id     = keys_[i].type_id;
      o.size        = 0;
      o.null_count  = 0;
      o.null_mask   = nullptr;
    }
    for (int r = 0; r < num_requests_; r++)
      for (int a = 0; a < requests_[r].num_aggregations; a++) {
        mig_column& o = mem_->out_results[borrowed_next_++];
        o.type_id     = target_type_of(requests_[r].values.type_id, requests_[r].aggregations[a].kind);
        o.size        = 0;
        o.null_count  = 0;
        o.null_mask   = nullptr;
      }
    (void)c;
    return;
  }
  for (int i = 0; i < num_keys_; i++) res_->keys.push_back(make_empty_column(A(), keys_[i].type_id));
  for (int r = 0; r < num_requests_; r++) {
    std::vector<owned_column> cols;
    for (int a = 0; a < requests_[r].num_aggregations; a++)
      cols.push_back(make_empty_column(
        A(), target_type_of(requests_[r].values.type_id, requests_[r].aggregations[a].kind)));
    res_->results.push_back(std::move(cols));
  }
  MIG_HIP_TRY(hipStreamSynchronize(stream_));
}

void groupby_call::plan_keys()
{
  bool any_key_nulls = false;
  for (int i = 0; i < num_keys_; i++) {
    bool has_nulls = nulls_of(keys_[i]) > 0;
    any_key_nulls |= has_nulls;
    P().keys[i] = make_dcol(keys_[i], has_nulls);
  }
  P().skip_null_keys = (null_policy_ == MIG_NULL_EXCLUDE) && any_key_nulls;
  bool const inline_keys = num_keys_ == 1 && storage_of(keys_[0].type_id) != ST_STRING;
  P().key_mode    = inline_keys ? KM_INLINE : KM_GENERAL;
  P().key_storage = inline_keys ? storage_of(keys_[0].type_id) : ST_NONE;
}

// Flattens the requests into per-column accumulators (hash/groupby.cu:113-181,
// 351-383): MEAN and VARIANCE/STD reduce to SUM + COUNT_VALID (+ M2), string
// MIN / MAX to ARGMIN / ARGMAX, duplicates share one accumulator.
void groupby_call::plan_accumulators()
{
  gb_params& p = P();
  vcol_of_request_.assign(size_t(num_requests_), 0);
  for (int r = 0; r < num_requests_; r++) {
    mig_column_view const& v = requests_[r].values;
    int idx = -1;
    for (size_t j = 0; j < vals_.size(); j++)
      if (same_view(vals_[j].view, v)) idx = int(j);
    if (idx < 0) {
      value_plan vp;
      vp.view      = v;
      vp.storage   = storage_of(v.type_id);
      vp.has_nulls = nulls_of(v) > 0;
      vals_.push_back(vp);
      idx = int(vals_.size()) - 1;
    }
    vcol_of_request_[r] = idx;
    value_plan& vp = vals_[idx];
    for (int a = 0; a < requests_[r].num_aggregations; a++) {
      switch (requests_[r].aggregations[a].kind) {
        case MIG_AGG_SUM: vp.want_sum = true; vp.want_count |= vp.has_nulls; break;
        case MIG_AGG_SUM_OF_SQUARES: vp.want_sumsq = true; vp.want_count |= vp.has_nulls; break;
        // string MIN / MAX through ARGMIN / ARGMAX and a gather (hash/groupby.cu:113-181)
        case MIG_AGG_MIN:
          if (vp.storage == ST_STRING) vp.want_argmin = true;
          else { vp.want_min = true; vp.want_count |= vp.has_nulls; }
          break;
        case MIG_AGG_MAX:
          if (vp.storage == ST_STRING) vp.want_argmax = true;
          else { vp.want_max = true; vp.want_count |= vp.has_nulls; }
          break;
        case MIG_AGG_COUNT_VALID: vp.want_count = true; break;
        case MIG_AGG_COUNT_ALL: p.need_count_all = 1; break;
        case MIG_AGG_MEAN: vp.want_sum = vp.want_count = true; break;
        case MIG_AGG_VARIANCE:
        case MIG_AGG_STD: vp.want_sum = vp.want_count = vp.want_m2 = true; break;
        case MIG_AGG_PRODUCT: vp.want_prod = true; vp.want_count |= vp.has_nulls; break;
        case MIG_AGG_ARGMIN: vp.want_argmin = true; break;
        case MIG_AGG_ARGMAX: vp.want_argmax = true; break;
        case MIG_AGG_NUNIQUE:
          vp.want_nunique[requests_[r].aggregations[a].null_handling == MIG_NULL_INCLUDE] = true;
          break;
        case MIG_AGG_NTH_ELEMENT: {
          int32_t const n    = requests_[r].aggregations[a].n;
          bool const include = requests_[r].aggregations[a].null_handling == MIG_NULL_INCLUDE;
          if (std::none_of(vp.nth.begin(), vp.nth.end(),
                           [&](auto const& e) { return e[0] == n && e[1] == int32_t(include); }))
            vp.nth.push_back({n, int32_t(include), -1});
          break;
        }
        default: break;
      }
    }
  }
  MIG_EXPECTS(vals_.size() <= size_t(MAX_VALS), "too many distinct value columns");

  // A column without nulls counts its valid values with the group row count.
  int const COUNT_FROM_ROWS = -2;
  auto add_field = [&](int vc, int32_t kind) {
    MIG_EXPECTS(nf_ < MAX_FIELDS, "too many accumulators");
    field_desc& F = p.fields[nf_];
    F.kind        = kind;
    F.vcol        = vc;
    F.width       = field_width(kind);
    F.stride      = F.width;
    F.aux_sum     = -1;
    F.aux_count   = -1;
    return nf_++;
  };
  for (size_t j = 0; j < vals_.size(); j++) {
    value_plan& vp = vals_[j];
    p.val_field_begin[j] = nf_;
    p.vals[j]            = make_dcol(vp.view, vp.has_nulls);
    if (vp.want_sum) vp.f_sum = add_field(int(j), sum_field_kind(vp.storage));
    if (vp.want_sumsq) vp.f_sumsq = add_field(int(j), sumsq_field_kind(vp.storage));
    if (vp.want_min) vp.f_min = add_field(int(j), minmax_field_kind(vp.storage, true));
    if (vp.want_max) vp.f_max = add_field(int(j), minmax_field_kind(vp.storage, false));
    for (int i = 0; i < 2; i++)
      if (vp.want_nunique[i]) vp.f_nunique[i] = add_field(int(j), i ? FK_NUNIQUE_NULLS : FK_NUNIQUE);
    for (auto& e : vp.nth) {  // the field carries n and the null policy in its aux slots
      e[2]                     = add_field(int(j), FK_NTH);
      p.fields[e[2]].aux_sum   = e[0];
      p.fields[e[2]].aux_count = e[1];
    }
    if (vp.want_count) {
      if (vp.has_nulls) vp.f_count = add_field(int(j), FK_COUNT);
      else {
        vp.f_count       = COUNT_FROM_ROWS;
        p.need_count_all = 1;
      }
    }
    if (vp.want_prod)
      vp.f_prod = add_field(int(j), vp.storage == ST_F32 ? FK_PROD_F32 : (vp.storage == ST_F64 ? FK_PROD_F64 : FK_PROD_I64));
    if (vp.want_argmin) vp.f_argmin = add_field(int(j), FK_ARGMIN);
    if (vp.want_argmax) vp.f_argmax = add_field(int(j), FK_ARGMAX);
    if (vp.want_m2) {
      vp.f_m2                     = add_field(int(j), FK_M2);
      p.fields[vp.f_m2].aux_sum   = vp.f_sum;
      p.fields[vp.f_m2].aux_count = vp.f_count;
      p.need_second_pass          = 1;
    }
  }
  p.val_field_begin[vals_.size()] = nf_;
  p.num_vals   = int32_t(vals_.size());
  p.num_fields = nf_;
  // ARGMIN / ARGMAX (a row index updated by compare-and-swap against the
  // values column) and PRODUCT (compare-and-swap multiply): only the
  // device-wide table path implements them
  for (int f = 0; f < nf_; f++) {
    int32_t const k = p.fields[f].kind;
    any_arg_ |= k == FK_ARGMIN || k == FK_ARGMAX || k == FK_PROD_I64 || k == FK_PROD_F32 || k == FK_PROD_F64;
  }
}

void groupby_call::upload(int which)
{
  void* h = pin + which * sizeof(gb_params);
  std::memcpy(h, Pp_, sizeof(gb_params));
  MIG_HIP_TRY(hipMemcpyAsync(dP, h, sizeof(gb_params), hipMemcpyHostToDevice, stream_));
}

// Groups estimated from a sample of rows hashed into a small device table.
void groupby_call::estimate_groups()
{
  est_ = double(N);
  pop_ = double(N);
  if (N > 4096 && !keys_sorted_) {
    int64_t const ns = std::min<int64_t>(N, SAMPLE_ROWS);
    uint32_t const tslots = uint32_t(next_pow2(uint64_t(2 * ns)));
    auto* table    = S->get<uint64_t>(tslots);
    auto* counters = S->get<uint32_t>(5);
    MIG_HIP_TRY(hipMemsetAsync(table, 0, sizeof(uint64_t) * tslots, stream_));
    MIG_HIP_TRY(hipMemsetAsync(counters, 0, sizeof(uint32_t) * 5, stream_));
    upload(0);
    launch_sample_hash(dP, N, ns, table, tslots - 1, counters, stream_);
    auto* h = reinterpret_cast<uint32_t*>(pin + 2 * sizeof(gb_params));
    MIG_HIP_TRY(hipMemcpyAsync(h, counters, 20, hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipStreamSynchronize(stream_));
    double const sampled = double(h[0]), d = double(h[1]);
    max_mult_ = h[2];
    sampled_  = sampled;
    // population: rows that are not dropped, scaled from the sample
    pop_ = double(N) * (sampled / double(ns));
    if (ns == N) est_ = d;
    else if (sampled <= 0) est_ = 1.0;
    else est_ = estimate_distinct_skewed(sampled, d, h[3], h[4], std::max(pop_, d), &est_low_);
  }
  stats_[STAT_EST_GROUPS] = est_;
}

// A fixed-width output column with room for `rows_cap` rows that reports `G`
// rows.  Allocation-free calls take the caller's next buffer instead, which
// must hold `rows_cap` elements and, for a mask, ceil(rows_cap / 64) 64-bit
// words.
owned_column groupby_call::fixed_output(int32_t type_id, int64_t rows_cap, int64_t G, bool with_mask,
                                        bool zero_masks)
{
  owned_column oc;
  if (mem_) {
    int const i = borrowed_next_++;
    mig_column& c = i < num_keys_ ? mem_->out_keys[i] : mem_->out_results[i - num_keys_];
    size_t const need = size_t(type_width(type_id)) * size_t(rows_cap);
    size_t const need_mask = size_t((rows_cap + 63) / 64) * 8;
    MIG_EXPECTS(c.data != nullptr || rows_cap == 0, "output column without a data buffer");
    if (c.data_bytes < 0 || size_t(c.data_bytes) < need || (with_mask && (c.null_mask == nullptr ||
                                                                          c.mask_bytes < 0 ||
                                                                          size_t(c.mask_bytes) < need_mask)))
      fail(MIG_ERROR_OVERFLOW, "output column " + std::to_string(i) + " holds fewer than the " +
                                 std::to_string(rows_cap) + " rows (and validity words) the result needs");
    oc.col            = c;
    oc.col.type_id    = type_id;
    oc.col.size       = int32_t(G);
    oc.col.null_count = 0;
    if (!with_mask) oc.col.null_mask = nullptr;
    else if (zero_masks && rows_cap > 0) MIG_HIP_TRY(hipMemsetAsync(oc.col.null_mask, 0, need_mask, stream_));
    return oc;  // owns nothing: the *_bytes members stay zero
  }
  oc          = make_owned_fixed(A(), type_id, rows_cap, with_mask);
  oc.col.size = int32_t(G);
  if (with_mask && zero_masks && rows_cap > 0)
    MIG_HIP_TRY(hipMemsetAsync(oc.col.null_mask, 0, oc.mask_bytes, stream_));
  return oc;
}

// Output columns: keys, then the results in request order (duplicates
// computed again from the same accumulators, as extract_results copies them).
// Fixed-width columns get room for `rows_cap` rows and report `G` rows (the
// partitioned path allocates before the exact group count is known).  Masks
// are zeroed when `zero_masks` (they are then set bit by bit).
void groupby_call::alloc_outputs(int64_t rows_cap, int64_t G, bool zero_masks)
{
  gb_params& p = P();
  res_->keys.clear();
  res_->results.clear();
  string_outs_.clear();
  n_str_keys_    = 0;
  borrowed_next_ = 0;
  for (int i = 0; i < num_keys_; i++) {
    bool const with_mask = keys_[i].null_mask != nullptr;
    if (storage_of(keys_[i].type_id) == ST_STRING) {
      owned_column oc;
      oc.col.type_id         = MIG_TYPE_STRING;
      oc.col.size            = int32_t(G);
      oc.col.offsets_type_id = MIG_TYPE_INT32;
      oc.offsets_bytes       = sizeof(int32_t) * size_t(G + 1);
      oc.col.offsets         = A().allocate(oc.offsets_bytes);
      if (with_mask) {
        oc.mask_bytes    = std::max<size_t>(4, size_t((G + 31) / 32) * 4);
        oc.col.null_mask = static_cast<uint32_t*>(A().allocate(oc.mask_bytes));
      }
      res_->keys.push_back(oc);
      n_str_keys_++;
    } else {
      res_->keys.push_back(fixed_output(keys_[i].type_id, rows_cap, G, with_mask, zero_masks));
      if (is_decimal_type(keys_[i].type_id)) res_->keys.back().col.scale = keys_[i].scale;
    }
    p.out_keys[i]     = storage_of(keys_[i].type_id) == ST_STRING ? nullptr : res_->keys.back().col.data;
    p.out_key_mask[i] = res_->keys.back().col.null_mask;
  }
  int no = 0;
  for (int r = 0; r < num_requests_; r++) {
    value_plan const& vp = vals_[vcol_of_request_[r]];
    std::vector<owned_column> cols;
    for (int a = 0; a < requests_[r].num_aggregations; a++) {
      int32_t const kind = requests_[r].aggregations[a].kind;
      int32_t const tt   = target_type_of(vp.view.type_id, kind);
      MIG_EXPECTS(no < MAX_OUTS, "too many aggregation results");
      out_desc& O   = p.outs[no++];
      O             = out_desc{};
      O.count_field = -1;
      O.src_storage = vp.storage;
      O.ddof        = requests_[r].aggregations[a].ddof;
      O.out_width   = type_width(tt);
      bool with_mask = false;
      switch (kind) {
        case MIG_AGG_SUM:
        case MIG_AGG_SUM_OF_SQUARES:
        case MIG_AGG_PRODUCT:
          O.field = kind == MIG_AGG_SUM ? vp.f_sum : (kind == MIG_AGG_PRODUCT ? vp.f_prod : vp.f_sumsq);
          O.kind  = tt == MIG_TYPE_FLOAT32 ? OK_COPY_F32 : (tt == MIG_TYPE_FLOAT64 ? OK_COPY_F64 : OK_COPY_I64);
          if (tt == MIG_TYPE_DECIMAL32) O.kind = OK_SUM_I32;  // the representation's sum, as int32
          if (vp.has_nulls) { O.count_field = vp.f_count; with_mask = true; }
          break;
        case MIG_AGG_MIN:
        case MIG_AGG_MAX:
          if (vp.storage == ST_STRING) {
            // the group's ARGMIN / ARGMAX rows into scratch; the strings are
            // gathered from them once the rows are known
            O.kind    = OK_ARG;
            O.field   = kind == MIG_AGG_MIN ? vp.f_argmin : vp.f_argmax;
            O.data    = S->get<int32_t>(std::max<int64_t>(rows_cap, 1));
            O.mask    = nullptr;
            owned_column oc;
            oc.col.type_id         = MIG_TYPE_STRING;
            oc.col.size            = int32_t(G);
            oc.col.offsets_type_id = MIG_TYPE_INT32;
            oc.offsets_bytes       = sizeof(int32_t) * size_t(G + 1);
            oc.col.offsets         = A().allocate(oc.offsets_bytes);
            if (vp.has_nulls) {
              oc.mask_bytes    = std::max<size_t>(8, size_t((G + 63) / 64) * 8);
              oc.col.null_mask = static_cast<uint32_t*>(A().allocate(oc.mask_bytes));
            }
            string_outs_.push_back({r, int(cols.size()), vcol_of_request_[r], static_cast<int32_t*>(O.data)});
            cols.push_back(oc);
            continue;
          }
          O.field = kind == MIG_AGG_MIN ? vp.f_min : vp.f_max;
          O.kind  = OK_MINMAX;
          if (vp.has_nulls) { O.count_field = vp.f_count; with_mask = true; }
          break;
        case MIG_AGG_COUNT_VALID:
          if (vp.f_count >= 0) { O.kind = OK_COUNT; O.field = vp.f_count; }
          else O.kind = OK_COUNT_ALL;
          break;
        case MIG_AGG_COUNT_ALL: O.kind = OK_COUNT_ALL; break;
        case MIG_AGG_MEAN:
          O.kind        = is_decimal_type(tt) ? OK_MEAN_INT : OK_MEAN;
          O.field       = vp.f_sum;
          O.count_field = vp.f_count;
          with_mask     = vp.has_nulls;
          break;
        case MIG_AGG_ARGMIN:
        case MIG_AGG_ARGMAX:
          O.kind    = OK_ARG;
          O.field   = kind == MIG_AGG_ARGMIN ? vp.f_argmin : vp.f_argmax;
          with_mask = vp.has_nulls;
          break;
        case MIG_AGG_NTH_ELEMENT:
          O.kind    = OK_NTH;
          O.field   = vp.nth_field(requests_[r].aggregations[a].n,
                                   requests_[r].aggregations[a].null_handling == MIG_NULL_INCLUDE);
          with_mask = true;
          if (vp.storage == ST_STRING) {
            // the group's n-th row into scratch; the strings are gathered
            // from it once the rows are known (null: no such row, or a null)
            O.kind = OK_ARG;
            O.data = S->get<int32_t>(std::max<int64_t>(rows_cap, 1));
            O.mask = nullptr;
            owned_column oc;
            oc.col.type_id         = MIG_TYPE_STRING;
            oc.col.size            = int32_t(G);
            oc.col.offsets_type_id = MIG_TYPE_INT32;
            oc.offsets_bytes       = sizeof(int32_t) * size_t(G + 1);
            oc.col.offsets         = A().allocate(oc.offsets_bytes);
            oc.mask_bytes          = std::max<size_t>(8, size_t((G + 63) / 64) * 8);
            oc.col.null_mask       = static_cast<uint32_t*>(A().allocate(oc.mask_bytes));
            string_outs_.push_back({r, int(cols.size()), vcol_of_request_[r], static_cast<int32_t*>(O.data)});
            cols.push_back(oc);
            continue;
          }
          break;
        case MIG_AGG_NUNIQUE:
          O.kind  = OK_COUNT;
          O.field = vp.f_nunique[requests_[r].aggregations[a].null_handling == MIG_NULL_INCLUDE];
          break;
        case MIG_AGG_VARIANCE:
        case MIG_AGG_STD:
          O.kind        = kind == MIG_AGG_VARIANCE ? OK_VAR : OK_STD;
          O.field       = vp.f_m2;
          O.count_field = vp.f_count;
          with_mask     = true;
          break;
        default: fail(MIG_ERROR_LOGIC, "unreachable aggregation kind");
      }
      owned_column oc = fixed_output(tt, rows_cap, G, with_mask, zero_masks);
      if (is_decimal_type(tt)) oc.col.scale = vp.view.scale;
      O.data          = oc.col.data;
      O.mask          = oc.col.null_mask;
      cols.push_back(oc);
    }
    res_->results.push_back(std::move(cols));
  }
  p.num_outs = no;
}

void groupby_call::release_outputs()
{
  for (auto& k : res_->keys) {
    A().deallocate(k.col.data, k.data_bytes);
    A().deallocate(k.col.null_mask, k.mask_bytes);
    A().deallocate(k.col.offsets, k.offsets_bytes);
  }
  for (auto& rc : res_->results)
    for (auto& c : rc) {
      A().deallocate(c.col.data, c.data_bytes);
      A().deallocate(c.col.null_mask, c.mask_bytes);
      A().deallocate(c.col.offsets, c.offsets_bytes);
    }
  res_->keys.clear();
  res_->results.clear();
}

// Key columns of general keys from each group's representative input row:
// fixed-width keys and validity by k_gather_keys, string keys by sizes, a
// scan of the offsets and a character gather.
void groupby_call::gather_general_keys(uint32_t const* rep_rows, int64_t G)
{
  int32_t* str_sizes = n_str_keys_ > 0 ? S->get<int32_t>(int64_t(n_str_keys_) * G + 1) : nullptr;
  int64_t* str_begin = n_str_keys_ > 0 ? S->get<int64_t>(int64_t(n_str_keys_) * G + 1) : nullptr;
  launch_gather_keys(dP, P(), rep_rows, G, str_sizes, str_begin, stream_);
  int si = 0;
  for (int i = 0; i < num_keys_; i++) {
    if (storage_of(keys_[i].type_id) != ST_STRING) continue;
    owned_column& oc = res_->keys[i];
    auto* offs       = static_cast<int32_t*>(oc.col.offsets);
    if (G == 0) {
      MIG_HIP_TRY(hipMemsetAsync(offs, 0, sizeof(int32_t), stream_));
      si++;
      continue;
    }
    MIG_HIP_TRY(hipMemcpyAsync(offs, str_sizes + int64_t(si) * G, sizeof(int32_t) * G,
                               hipMemcpyDeviceToDevice, stream_));
    launch_exclusive_scan_i32(offs, G, offs + G, stream_);
    auto* h = reinterpret_cast<int32_t*>(pin + 2 * sizeof(gb_params));
    MIG_HIP_TRY(hipMemcpyAsync(h, offs + G, 4, hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipStreamSynchronize(stream_));
    int32_t const total_chars = *h;
    MIG_EXPECTS(total_chars >= 0, "unique string keys exceed 2^31 characters");
    oc.data_bytes     = size_t(total_chars);
    oc.col.data       = A().allocate(oc.data_bytes);
    oc.col.data_bytes = total_chars;
    launch_gather_chars_from(P().keys[i].data, str_begin + int64_t(si) * G, G, offs,
                             static_cast<uint8_t*>(oc.col.data), stream_);
    si++;
  }
}

// Pre-sorted keys: groups are the runs of equal adjacent key rows (the
// reference's sort path with keys_pre_sorted, sort/sort_helper.cu:55-72,
// 133-181); groups come out in input order.
int64_t groupby_call::aggregate_sorted_runs()
{
  gb_params& p = P();
  int const num_cus     = props().num_cus;
  p.key_mode            = KM_GENERAL;  // keys are gathered from each group's first row
  int64_t const ntiles  = sorted_tiles(N);
  uint32_t* tile_counts = S->get<uint32_t>(ntiles + 1);
  upload(0);
  launch_sorted_count(dP, N, tile_counts, stream_);
  launch_scan_blocks(tile_counts, ntiles, tile_counts + ntiles, stream_);
  auto* h_groups = reinterpret_cast<uint32_t*>(pin + 2 * sizeof(gb_params));
  MIG_HIP_TRY(hipMemcpyAsync(h_groups, tile_counts + ntiles, 4, hipMemcpyDeviceToHost, stream_));
  MIG_HIP_TRY(hipStreamSynchronize(stream_));
  int64_t const G = int64_t(*h_groups);
  // accumulators of G groups (slots 0..G-1; the two special slots unused)
  int32_t* d_flags = S->get<int32_t>(4);
  p.cap            = uint64_t(G);
  p.cap_mask       = 0;
  p.overflow       = d_flags;
  p.special        = reinterpret_cast<uint32_t*>(d_flags + 1);
  uint64_t const total = uint64_t(G) + 2;
  p.slot_rows      = S->get<uint64_t>(total);
  p.slot_count     = p.need_count_all ? S->get<uint32_t>(total) : nullptr;
  p.count_stride   = 4;
  for (int f = 0; f < nf_; f++) {
    p.fields[f].ptr    = p.fields[f].width == 8 ? static_cast<void*>(S->get<uint64_t>(total))
                                                : static_cast<void*>(S->get<uint32_t>(total));
    p.fields[f].stride = p.fields[f].width;
  }
  upload(0);
  launch_init_state(dP, p, stream_);
  uint32_t* gid = S->get<uint32_t>(N);
  bool const any_nth = std::any_of(p.fields, p.fields + nf_, [](field_desc const& F) { return F.kind == FK_NTH; });
  uint32_t* group_last = any_nth ? S->get<uint32_t>(std::max<int64_t>(G, 1)) : nullptr;
  launch_sorted_assign(dP, N, tile_counts, gid, group_last, stream_);
  T->mark(2);
  sorted_stats_cover const cover = sorted_stats_plan(p);
  uint8_t* m2_scratch =
    cover.m2_columns > 0 ? S->get<uint8_t>(sorted_stats_scratch_bytes(G, cover.m2_columns)) : nullptr;
  launch_sorted_stats(dP, p, N, G, gid, num_cus, m2_scratch, stream_);
  launch_sorted_aggregate(dP, p, N, gid, num_cus, &cover, stream_);
  launch_sorted_string_arg(dP, p, N, gid, stream_);
  T->mark(3);
  T->mark(4);
  if (p.need_second_pass) {
    launch_sorted_second_pass(dP, p, N, gid, num_cus, &cover, stream_);
    launch_sorted_stats_m2(dP, p, N, G, gid, num_cus, m2_scratch, stream_);
  }
  for (int f = 0; f < nf_; f++) {
    if (p.fields[f].kind != FK_NUNIQUE && p.fields[f].kind != FK_NUNIQUE_NULLS) continue;
    launch_sorted_nunique(dP, f, N, G, gid, S->get<uint32_t>(2 * N), stream_);
  }
  for (int f = 0; f < nf_; f++)
    if (p.fields[f].kind == FK_NTH)
      launch_sorted_nth(dP, p, f, G, group_last, S->get<uint32_t>(std::max<int64_t>(G, 1) + 1), stream_);
  T->mark(5);
  uint32_t* map = S->get<uint32_t>(std::max<int64_t>(G, 1));
  launch_iota_u32(map, G, stream_);
  T->mark(6);
  alloc_outputs(G, G, false);
  uint32_t* rep_rows = S->get<uint32_t>(std::max<int64_t>(G, 1));
  upload(1);
  launch_outputs(dP, p, map, G, rep_rows, stream_);
  gather_general_keys(rep_rows, G);
  string_results(G);
  MIG_HIP_TRY(hipGetLastError());
  stats_[STAT_STRATEGY] = 3;  // runs of pre-sorted keys
  stats_[STAT_CAPACITY] = double(G);
  return G;
}

// Radix-partitioned LDS aggregation for high cardinality: returns the number
// of groups, or -1 when the path does not apply (or its bins overflowed three
// times) and the device-wide table takes the call.
int64_t groupby_call::aggregate_partitioned()
{
  gb_params& p    = P();
  int const forced = g_strategy.load();
  if (mem_ || forced == MIG_STRATEGY_GLOBAL) return -1;
  bool rpa_ok   = !any_arg_ && (p.key_mode == KM_INLINE || p.key_mode == KM_GENERAL);
  int n_flagged = 0;
  for (int v = 0; v < p.num_vals && rpa_ok; v++) {
    if (p.val_field_begin[v] == p.val_field_begin[v + 1]) continue;
    if (vals_[v].storage == ST_STRING) rpa_ok = false;
    if (vals_[v].has_nulls) n_flagged++;
  }
  rpa_ok = rpa_ok && n_flagged <= RPA_MAX_FLAGGED;
  {  // the aggregate kernel keeps its output plan in registers
    int n_outs = 0;
    for (int r = 0; r < num_requests_; r++) n_outs += requests_[r].num_aggregations;
    rpa_ok = rpa_ok && n_outs <= RPA_PLAN_OUTS;
  }
  // the aggregate kernel keeps its accumulator plan in registers
  {
    int ncols = 0;
    for (int v = 0; v < p.num_vals; v++)
      if (p.val_field_begin[v] < p.val_field_begin[v + 1]) ncols++;
    rpa_ok = rpa_ok && ncols <= RPA_PLAN_COLS && nf_ <= RPA_PLAN_FIELDS;
  }
  // LDS of the aggregating workgroup for S hash slots: keys, S + 2 rows of
  // every accumulator, the compaction map and the shared counters.
  auto agg_lds_bytes = [&](int32_t slots, bool wide) {
    int32_t bytes = slots * 8;
    for (int f = 0; f < nf_; f++) bytes += ((slots + 2) * rpa_lds_width(p.fields[f].kind, wide) + 7) & ~7;
    if (p.need_count_all) bytes += ((slots + 2) * 4 + 7) & ~7;
    bytes += ((slots + 2) * 2 + 7) & ~7;
    return bytes + 128;
  };
  auto slots_for = [&](bool wide) {
    int32_t slots = RPA_LDS_SLOTS;
    while (slots > 64 && agg_lds_bytes(slots, wide) > RPA_LDS_BUDGET) slots /= 2;
    return slots;
  };
  // float32 sums in float64 where that costs no hash slots
  int32_t S_slots   = slots_for(false);
  bool const wide   = slots_for(true) == S_slots;
  rpa_ok = rpa_ok && agg_lds_bytes(S_slots, wide) <= RPA_LDS_BUDGET;
  int32_t const max_fill = int32_t(0.85 * S_slots);
  bool use_rpa = false;
  if (forced == MIG_STRATEGY_PARTITIONED || forced == MIG_STRATEGY_PARTITIONED_EXACT) use_rpa = rpa_ok;
  else if (forced == MIG_STRATEGY_AUTO)  // and the estimate fits the largest partition
    use_rpa = rpa_ok && N >= (1 << 20) && est_ >= 65536 && !heavy_hitters() &&
              est_ <= double(int64_t(1) << RPA_MAX_LOG_BINS) * max_fill;
  if (!use_rpa) return -1;

  int const num_cus = props().num_cus;
  rpa_params R{};
  // LDS offsets (8-byte aligned)
  int32_t off = S_slots * 8;
  R.wide_f32 = wide ? 1 : 0;
  for (int f = 0; f < nf_; f++) {
    p.fields[f].lds_off = off;
    off += ((S_slots + 2) * rpa_lds_width(p.fields[f].kind, wide) + 7) & ~7;
  }
  R.count_all_off = -1;
  if (p.need_count_all) {
    R.count_all_off = off;
    off += ((S_slots + 2) * 4 + 7) & ~7;
  }
  R.slot_of_j_off = off;
  off += ((S_slots + 2) * 2 + 7) & ~7;
  R.misc_off  = off;
  off += 128;
  R.lds_bytes  = off;
  R.lds_slots  = S_slots;
  R.max_fill   = max_fill;
  R.plan_ncols = 0;
  R.plan_fb[0] = 0;
  for (int v = 0; v < p.num_vals; v++) {
    if (p.val_field_begin[v] == p.val_field_begin[v + 1]) continue;
    int const i      = R.plan_ncols++;
    R.plan_col[i]    = v;
    R.plan_fb[i + 1] = p.val_field_begin[v + 1];
    if (i == 0) R.plan_fb[0] = p.val_field_begin[v];
  }
  for (int f = 0; f < nf_; f++)
    R.plan_fd[f] = uint32_t(p.fields[f].kind) | (uint32_t(p.fields[f].lds_off) << 8);
  // LDS offset of a count: a count field, or the group row counts (-2)
  auto count_off = [&](int32_t f) { return f == -2 ? R.count_all_off : (f >= 0 ? p.fields[f].lds_off : -1); };
  R.m2_n = 0;
  for (int f = 0; f < nf_; f++) {
    if (p.fields[f].kind != FK_M2) continue;
    int const j      = R.m2_n++;
    field_desc const& SF = p.fields[p.fields[f].aux_sum];
    R.m2_off[j]     = p.fields[f].lds_off;
    R.m2_sum[j]     = uint32_t(SF.kind) | (uint32_t(SF.lds_off) << 8);
    R.m2_cnt_off[j] = count_off(p.fields[f].aux_count);
    R.m2_storage[j] = p.vals[p.fields[f].vcol].storage;
  }
  // the output plan (fixed by the requests; the buffers come from P.outs)
  auto fill_out_plan = [&]() {
    R.out_n = p.num_outs;
    for (int o = 0; o < p.num_outs; o++) {
      out_desc const& O = p.outs[o];
      int32_t const fk  = O.field >= 0 ? p.fields[O.field].kind : 0;
      R.out_op[o]   = uint32_t(O.kind) | (uint32_t(fk) << 8) | (uint32_t(O.out_width) << 16);
      R.out_foff[o] = O.kind == OK_COUNT_ALL ? R.count_all_off : (O.field >= 0 ? p.fields[O.field].lds_off : -1);
      R.out_coff[o] = O.count_field == -1 ? -1 : count_off(O.count_field);
      R.out_ddof[o] = O.ddof;
    }
  };
  // General keys travel in hashed form (compared by 64-bit words, verified
  // exactly afterwards); the form that compares key rows while probing is
  // the fallback when a verification finds two different keys.
  bool hashed = p.key_mode == KM_GENERAL && forced != MIG_STRATEGY_PARTITIONED_EXACT;
  R.general_hashed = hashed ? 1 : 0;
  // scatter staging layout: a tile of keys, copied values, flags and bins
  int32_t rec = 8 + 1 + 2 + (hashed ? 4 : 0);
  for (int v = 0; v < p.num_vals; v++)
    if (p.val_field_begin[v] < p.val_field_begin[v + 1]) rec += storage_width(vals_[v].storage);
  // the largest multiple of the block whose staging fits the budget
  R.tile_rows = std::max(1, std::min(RPA_TILE / RPA_SCATTER_BLOCK, RPA_STAGE_BUDGET / ((rec + 4) * RPA_SCATTER_BLOCK))) *
                RPA_SCATTER_BLOCK;
  {
    int32_t so = R.tile_rows * 8;
    for (int v = 0; v < p.num_vals; v++) {
      if (p.val_field_begin[v] == p.val_field_begin[v + 1]) continue;
      R.stage_val_off[v] = so;
      so += (R.tile_rows * storage_width(vals_[v].storage) + 7) & ~7;
    }
    R.stage_flags_off = so;
    so += (R.tile_rows + 7) & ~7;
    R.stage_bin_off = so;
    so += (R.tile_rows * 2 + 7) & ~7;
    R.stage_row_off = so;
    if (hashed) so += R.tile_rows * 4;
    R.stage_bytes = so;
  }
  R.hist_blocks = num_cus * 2;
  int flag_bit = 0;
  for (int v = 0; v < p.num_vals; v++)
    R.val_flag_bit[v] = (p.val_field_begin[v] < p.val_field_begin[v + 1] && vals_[v].has_nulls) ? flag_bit++ : -1;
  // bins average 0.6·S groups; the extra 10% absorbs the sampling error of
  // the estimate (about 7% at 10M groups from 65,536 sampled rows), so an
  // estimate a little high does not double the bins
  double const per_bin = 0.66 * double(S_slots);
  int log_bins = 1;
  while (log_bins < RPA_MAX_LOG_BINS && double(int64_t(1) << log_bins) * per_bin < est_) log_bins++;
  // the last doubling only for estimates that would fill 2^15 bins past
  // 0.78·S (measured: 2^15 fuller bins beat 2^16 half-empty ones by ~0.8 ms
  // at 25M groups)
  if (log_bins == RPA_MAX_LOG_BINS && est_ <= double(int64_t(1) << (RPA_MAX_LOG_BINS - 1)) * 0.78 * S_slots)
    log_bins = RPA_MAX_LOG_BINS - 1;
  // Bins are laid out either exactly, from a histogram pass over the keys,
  // or speculatively, at fixed strides sized from the sampled estimate so
  // that no histogram pass is needed.  Rows of one key all land in one
  // bin, so a key heavy enough to fill a bin defeats the speculative
  // layout: it is used only when no key repeats 8 times in the sample
  // (then the heaviest key holds under ~2e-4 of the rows with high
  // probability); a bin that still fills up is flagged and the call is
  // redone with the exact layout.
  bool speculative = max_mult_ < 8 && pop_ >= double(1 << 20) && est_ >= 65536;
  int bin_attempts = 0;
  int64_t G        = -1;
  while (G < 0 && bin_attempts < 3) {
    R.log_bins  = log_bins;
    R.two_level = log_bins > RPA_MAX_LOG_FANOUT;
    R.log_f1    = R.two_level ? (log_bins + 1) / 2 : log_bins;
    int64_t const B  = int64_t(1) << log_bins;
    int64_t const F1 = int64_t(1) << R.log_f1;
    int64_t const HUGE_CAP = int64_t(1) << 62;
    R.cap1 = R.cap2 = HUGE_CAP;
    int64_t rowsA = N, rowsB = N;  // records each partition buffer holds
    if (speculative) {
      // rows per bin: mean mu, variance ~ mu * (1 + rows per group)
      double const m   = pop_ / std::max(est_, 1.0);
      double const mu2 = pop_ / double(B), mu1 = pop_ / double(F1);
      R.cap2 = int64_t(std::ceil(1.2 * mu2 + 6.0 * std::sqrt(mu2 * (1.0 + m)) + 256.0));
      R.cap1 = R.two_level ? int64_t(std::ceil(1.1 * mu1 + 6.0 * std::sqrt(mu1 * (1.0 + m)) + 4096.0)) : R.cap2;
      rowsB  = B * R.cap2;
      rowsA  = R.two_level ? F1 * R.cap1 : 0;
    }
    R.hist         = speculative ? nullptr : S->get<uint32_t>(B);
    R.hist_partial = speculative ? nullptr : S->get<uint32_t>(size_t(R.hist_blocks) * B);
    R.bin_start    = S->get<int64_t>(B + 1);
    R.cursor1      = S->get<unsigned long long>(F1 * RPA_CURSOR_STRIDE);
    R.cursor2      = S->get<unsigned long long>(B * RPA_CURSOR_STRIDE);
    R.tile_start1  = S->get<int64_t>(F1 + 1);
    R.first_start  = S->get<int64_t>(F1);
    int32_t* d_small = S->get<int32_t>(4);
    R.overflow       = d_small;
    R.cap_overflow   = d_small + 1;
    R.group_counter  = reinterpret_cast<unsigned long long*>(S->get<uint64_t>(1));
    rpa_params* dR   = S->get<rpa_params>(1);
    if (!speculative) MIG_HIP_TRY(hipMemsetAsync(R.hist, 0, sizeof(uint32_t) * B, stream_));
    MIG_HIP_TRY(hipMemsetAsync(d_small, 0, sizeof(int32_t) * 4, stream_));
    MIG_HIP_TRY(hipMemsetAsync(R.group_counter, 0, sizeof(uint64_t), stream_));
    upload(0);
    MIG_HIP_TRY(hipMemcpyAsync(dR, &R, sizeof(rpa_params), hipMemcpyHostToDevice, stream_));
    T->mark(8);
    if (!speculative) launch_rpa_hist(dP, dR, R, N, num_cus, stream_);
    T->mark(9);
    // The output columns are allocated while the first kernels run: a bin
    // yields at most min(its rows, max_fill + 2) groups, so
    // B·(max_fill + 2) rows, and at most N, bound the groups before the bin
    // sizes are known.
    int64_t const rows_cap = std::min<int64_t>(N, B * int64_t(max_fill + 2));
    alloc_outputs(rows_cap, p.key_mode == KM_GENERAL ? rows_cap : 0, true);
    // bin layout (starts, cursors) computed on the device
    if (speculative) launch_rpa_spec_layout(dR, R, stream_);
    else launch_rpa_layout(dR, stream_);
    // partition buffers
    size_t part_bytes = 0;
    R.keysB  = S->get<uint64_t>(rowsB);
    R.flagsB = S->get<uint8_t>(rowsB);
    part_bytes += 9 * size_t(rowsB);
    if (R.two_level) {
      R.keysA  = S->get<uint64_t>(rowsA);
      R.flagsA = S->get<uint8_t>(rowsA);
      part_bytes += 9 * size_t(rowsA);
    }
    R.general_hashed = hashed ? 1 : 0;
    R.rowsA = R.rowsB = nullptr;
    R.rep_rows = nullptr;
    R.rec_slot = nullptr;
    R.rec_count = 0;
    R.verified = nullptr;
    R.verify_failed = nullptr;
    R.rep_stride = 0;
    if (hashed) {
      R.rowsB = S->get<uint32_t>(rowsB);
      if (R.two_level) R.rowsA = S->get<uint32_t>(rowsA);
      R.rep_stride    = (int64_t(S_slots) + 2 + 31) & ~int64_t(31);  // whole lines per bin
      R.rep_rows      = S->get<uint32_t>(size_t(B) * size_t(R.rep_stride));
      R.rec_count     = N;  // rec_slot is indexed by input row
      R.rec_slot      = S->get<uint32_t>(size_t(N));
      R.verified      = S->get<unsigned long long>(1);
      R.verify_failed = S->get<int32_t>(1);
      MIG_HIP_TRY(hipMemsetAsync(R.rec_slot, 0xff, sizeof(uint32_t) * size_t(N), stream_));
      MIG_HIP_TRY(hipMemsetAsync(R.verified, 0, sizeof(unsigned long long), stream_));
      MIG_HIP_TRY(hipMemsetAsync(R.verify_failed, 0, sizeof(int32_t), stream_));
      part_bytes += 4 * size_t(rowsB) * 2 + (R.two_level ? 4 * size_t(rowsA) : 0) + 4 * size_t(B) * R.rep_stride;
    }
    for (int v = 0; v < p.num_vals; v++) {
      if (p.val_field_begin[v] == p.val_field_begin[v + 1]) continue;
      size_t w   = size_t(storage_width(vals_[v].storage));
      R.valsB[v] = S->get<uint8_t>(w * rowsB);
      part_bytes += w * rowsB;
      if (R.two_level) {
        R.valsA[v] = S->get<uint8_t>(w * rowsA);
        part_bytes += w * rowsA;
      }
    }
    stats_[STAT_STATE_BYTES] = double(part_bytes);
    R.out_capacity = uint64_t(rows_cap);
    R.out_rep      = p.key_mode == KM_GENERAL ? S->get<uint32_t>(std::max<int64_t>(1, rows_cap)) : nullptr;
    fill_out_plan();
    upload(0);
    MIG_HIP_TRY(hipMemcpyAsync(dR, &R, sizeof(rpa_params), hipMemcpyHostToDevice, stream_));
    T->mark(10);
    launch_rpa_scatter1(dP, dR, R, N, num_cus, stream_);
    T->mark(11);
    if (R.two_level) {
      launch_rpa_tiles(dR, stream_);
      launch_rpa_scatter2(dP, dR, R, num_cus, 0, -1, stream_);
    }
    T->mark(12);
    launch_rpa_aggregate(dP, dR, R, p.key_mode == KM_GENERAL ? (hashed ? 2 : 1) : 0, 0, B, stream_);
    T->mark(13);
    MIG_HIP_TRY(hipGetLastError());
    // the hashed form compares every record with its representative before
    // the results are read back (a no-op launch for the other forms)
    if (hashed) launch_rpa_verify(dP, dR, R.rec_count, stream_);
    MIG_HIP_TRY(hipGetLastError());
    auto* h = reinterpret_cast<uint64_t*>(pin + 2 * sizeof(gb_params));
    h[2] = 0;
    h[3] = 0;
    MIG_HIP_TRY(hipMemcpyAsync(h, R.group_counter, 8, hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipMemcpyAsync(h + 1, d_small, 8, hipMemcpyDeviceToHost, stream_));
    if (hashed) {
      MIG_HIP_TRY(hipMemcpyAsync(h + 2, R.verified, 8, hipMemcpyDeviceToHost, stream_));
      MIG_HIP_TRY(hipMemcpyAsync(h + 3, R.verify_failed, 4, hipMemcpyDeviceToHost, stream_));
    }
    MIG_HIP_TRY(hipStreamSynchronize(stream_));
    uint32_t const groups_overflow = uint32_t(h[1]), bins_overflow = uint32_t(h[1] >> 32);
    bool const verify_failed = hashed && !bins_overflow && groups_overflow == 0 && uint32_t(h[3]) != 0;
    stats_[STAT_VERIFIED_PAIRS] = double(h[2]);
    if (verify_failed) {  // two different keys met with equal words: compare key rows instead
      release_outputs();
      hashed = false;
      stats_[STAT_RETRIES] += 1;
    } else if (bins_overflow) {  // a speculative bin filled up: the same bins, laid out exactly
      release_outputs();
      speculative = false;
      stats_[STAT_RETRIES] += 1;
    } else if (groups_overflow == 0) {
      G = int64_t(h[0]);
      for (auto& k : res_->keys) k.col.size = int32_t(G);
      for (auto& rc : res_->results)
        for (auto& c : rc) c.col.size = int32_t(G);
      if (p.key_mode == KM_GENERAL) gather_general_keys(R.out_rep, G);
      stats_[STAT_STRATEGY] = MIG_STRATEGY_PARTITIONED;
      stats_[STAT_CAPACITY] = double(B);
      stats_[STAT_RETRIES] += bin_attempts;
      stats_[STAT_RPA_SPECULATIVE] = speculative ? 1.0 : 0.0;
    } else {
      release_outputs();
      log_bins = std::min(RPA_MAX_LOG_BINS, log_bins + 2);
      bin_attempts++;
    }
  }
  return G;
}

// The device-wide hash table (the reference's design,
// hash/groupby.cu:548-622), growing on overflow.  Low cardinality with one
// fixed-width key: per-workgroup LDS partial tables merged into the global
// table (one global atomic per workgroup, group and field instead of one per
// row and field).
int64_t groupby_call::aggregate_global()
{
  gb_params& p       = P();
  int const num_cus  = props().num_cus;
  int const forced   = g_strategy.load();
  uint64_t const max_cap = next_pow2(uint64_t(2 * N + 2));
  int64_t lds_blocks = 0;
  p.lds_bytes        = 0;
  // workgroup-local tables for few groups, or for hot keys among many
  // groups (rows of keys that find no room go to the device-wide table)
  bool const hot = heavy_hitters() && est_ > double(LDS_PREAGG_MAX_GROUPS);
  if (p.key_mode == KM_INLINE && !any_arg_ && (est_ <= double(LDS_PREAGG_MAX_GROUPS) || hot) &&
      forced != MIG_STRATEGY_PARTITIONED) {
    int32_t S_l = 256;
    while (S_l < LDS_PREAGG_SLOTS && double(S_l) * 0.75 < est_ * 1.5) S_l *= 2;
    int32_t off = S_l * 8, m2_off = S_l * 8;
    for (int f = 0; f < nf_; f++) {
      field_desc& F = p.fields[f];
      if (F.kind == FK_M2) {
        p.lds_m2_off[f] = m2_off;
        m2_off += S_l * 8;
      } else {
        F.lds_off = off;
        off += ((S_l + 2) * F.width + 7) & ~7;
      }
    }
    p.lds_count_off = -1;
    if (p.need_count_all) {
      p.lds_count_off = off;
      off += ((S_l + 2) * 4 + 7) & ~7;
    }
    int32_t const bytes = std::max(off, m2_off);
    if (bytes <= LDS_PREAGG_BUDGET) {
      p.lds_slots = S_l;
      p.lds_fill  = S_l * 3 / 4;
      p.lds_bytes = bytes;
      int64_t const per_cu = std::max<int64_t>(1, std::min<int64_t>(8, (160 * 1024) / bytes));
      // at least ~16 rows per group and workgroup, so that merging pays off
      // (hot keys: the tables hold the hot keys, the rest goes through)
      int64_t const rows_per_wg = hot ? 4096 : std::max<int64_t>(4096, int64_t(16.0 * std::max(est_, 1.0)));
      lds_blocks = std::max<int64_t>(1, std::min<int64_t>(int64_t(num_cus) * per_cu,
                                                           (N + rows_per_wg - 1) / rows_per_wg));
    }
  }
  stats_[STAT_LDS_PREAGG] = lds_blocks > 0 ? 1.0 : 0.0;
  // a lower-bound estimate (skewed keys) starts 4x larger: a table that
  // overflows costs the rows processed before the rebuild
  uint64_t cap = std::max<uint64_t>(64, next_pow2(uint64_t(est_ * (est_low_ ? 10.4 : 2.6)) + 1));
  cap = std::min(cap, max_cap);
  int32_t* d_flags = S->get<int32_t>(4);  // [0] overflow, [1..2] special occupancy
  uint32_t* block_counts = nullptr;
  int retries = 0;
  int64_t G   = -1;
  size_t const table_mark = S->mark();
  for (;;) {
    S->release_to(table_mark);  // an arena reuses the previous table's memory
    p.cap         = cap;
    p.cap_mask    = cap - 1;
    p.probe_limit = uint32_t(cap >= max_cap ? std::min<uint64_t>(cap, 0xffffffffull)
                                            : std::min<uint64_t>(cap, 4096));
    p.overflow    = d_flags;
    p.special     = reinterpret_cast<uint32_t*>(d_flags + 1);
    size_t state_bytes = 0;
    uint64_t const total = cap + 2;
    if (p.key_mode == KM_INLINE) p.slot_keys = S->get<uint64_t>(cap);
    else p.slot_rows = S->get<uint64_t>(cap);
    state_bytes += 8 * cap;
    // one array per accumulator (structure of arrays): the atomics of a
    // row then go to different lines, which the memory system serves in
    // parallel (a shared row of all accumulators measured slower)
    p.slot_count   = p.need_count_all ? S->get<uint32_t>(total) : nullptr;
    p.count_stride = 4;
    if (p.slot_count) state_bytes += 4 * total;
    for (int f = 0; f < nf_; f++) {
      p.fields[f].ptr = p.fields[f].width == 8 ? static_cast<void*>(S->get<uint64_t>(total))
                                               : static_cast<void*>(S->get<uint32_t>(total));
      p.fields[f].stride = p.fields[f].width;
      state_bytes += size_t(p.fields[f].width) * total;
    }
    stats_[STAT_STATE_BYTES] = double(state_bytes);
    upload(0);
    launch_init_state(dP, p, stream_);
    T->mark(2);
    if (lds_blocks > 0) launch_aggregate_lds(dP, p, lds_blocks, stream_);
    else launch_aggregate(dP, p, num_cus, stream_);
    T->mark(3);
    MIG_HIP_TRY(hipGetLastError());

    int64_t nb   = (int64_t(cap) + 2 + SCAN_TILE - 1) / SCAN_TILE;
    block_counts = S->get<uint32_t>(nb + 1);
    launch_count_occupied(dP, p, block_counts, stream_);
    launch_scan_blocks(block_counts, nb, block_counts + nb, stream_);
    auto* h_flags = reinterpret_cast<uint32_t*>(pin + 2 * sizeof(gb_params));
    MIG_HIP_TRY(hipMemcpyAsync(h_flags, d_flags, 4, hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipMemcpyAsync(h_flags + 1, block_counts + nb, 4, hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipStreamSynchronize(stream_));
    if (h_flags[0] == 0) {
      G = h_flags[1];
      break;
    }
    MIG_EXPECTS(cap < max_cap, "hash table overflow at maximum capacity");
    cap = std::min(max_cap, cap * 4);
    retries++;
  }
  stats_[STAT_STRATEGY] = MIG_STRATEGY_GLOBAL;
  stats_[STAT_CAPACITY] = double(cap);
  stats_[STAT_RETRIES]  = retries;

  T->mark(4);
  if (p.need_second_pass) {
    if (lds_blocks > 0) launch_second_pass_lds(dP, p, lds_blocks, stream_);
    else launch_second_pass(dP, p, num_cus, stream_);
  }
  T->mark(5);
  uint32_t* map = S->get<uint32_t>(std::max<int64_t>(G, 1));
  launch_compact(dP, p, block_counts, map, stream_);
  T->mark(6);

  alloc_outputs(G, G, false);
  uint32_t* rep_rows = p.key_mode == KM_GENERAL ? S->get<uint32_t>(std::max<int64_t>(G, 1)) : nullptr;
  upload(1);  // stream-ordered after the kernels that read the previous block
  launch_outputs(dP, p, map, G, rep_rows, stream_);
  if (p.key_mode == KM_GENERAL) gather_general_keys(rep_rows, G);
  string_results(G);
  return G;
}

// string MIN / MAX: lengths and validity, offsets, then the characters
void groupby_call::string_results(int64_t G)
{
  for (auto const& so : string_outs_) {
    owned_column& oc = res_->results[so.r][so.a];
    auto* offs       = static_cast<int32_t*>(oc.col.offsets);
    launch_str_arg_sizes(P().vals[so.vc], so.rows, G, offs, oc.col.null_mask, stream_);
    if (G == 0) {
      MIG_HIP_TRY(hipMemsetAsync(offs, 0, sizeof(int32_t), stream_));
      continue;
    }
    launch_exclusive_scan_i32(offs, G, offs + G, stream_);
    auto* h = reinterpret_cast<int32_t*>(pin + 2 * sizeof(gb_params));
    MIG_HIP_TRY(hipMemcpyAsync(h, offs + G, 4, hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipStreamSynchronize(stream_));
    int32_t const total = *h;
    MIG_EXPECTS(total >= 0, "string MIN / MAX results exceed 2^31 characters");
    oc.data_bytes     = size_t(total);
    oc.col.data       = A().allocate(oc.data_bytes);
    oc.col.data_bytes = total;
    launch_gather_string_chars(P().vals[so.vc], reinterpret_cast<uint32_t const*>(so.rows), G, offs,
                               static_cast<uint8_t*>(oc.col.data), stream_);
  }
}

// Null counts of the outputs, computed on the device and read back; phase
// times; allocation-free calls hand the columns back in the caller's structs.
void groupby_call::finish(int64_t G)
{
  MIG_HIP_TRY(hipGetLastError());
  stats_[STAT_KEY_MODE] = P().key_mode;
  stats_[STAT_GROUPS]   = double(G);
  res_->num_groups      = int32_t(G);

  std::vector<mig_column*> masked;
  for (auto& k : res_->keys)
    if (k.col.null_mask) masked.push_back(&k.col);
  for (auto& rc : res_->results)
    for (auto& c : rc)
      if (c.col.null_mask) masked.push_back(&c.col);
  uint32_t* d_valid = masked.empty() ? nullptr : S->get<uint32_t>(masked.size());
  if (d_valid) {
    MIG_HIP_TRY(hipMemsetAsync(d_valid, 0, 4 * masked.size(), stream_));
    std::vector<uint32_t const*> masks(masked.size());
    for (size_t m = 0; m < masked.size(); m++) masks[m] = masked[m]->null_mask;
    launch_count_valid_list(masks.data(), int(masks.size()), G, d_valid, stream_);
    auto* hv = reinterpret_cast<uint32_t*>(pin + 2 * sizeof(gb_params));
    MIG_HIP_TRY(hipMemcpyAsync(hv, d_valid, 4 * masked.size(), hipMemcpyDeviceToHost, stream_));
    MIG_HIP_TRY(hipStreamSynchronize(stream_));
    for (size_t m = 0; m < masked.size(); m++) masked[m]->null_count = int32_t(G - hv[m]);
  }
  T->mark(7);
  MIG_HIP_TRY(hipStreamSynchronize(stream_));
  stats_[STAT_T_ESTIMATE]      = T->between(0, 1);
  stats_[STAT_T_INIT]          = T->between(1, 2);
  stats_[STAT_T_ACCUMULATE]    = T->between(2, 3);
  stats_[STAT_T_SECOND_PASS]   = T->between(4, 5);
  stats_[STAT_T_COMPACT]       = T->between(3, 4) + T->between(5, 6);
  stats_[STAT_T_OUTPUTS]       = T->between(6, 7) + T->between(13, 7);
  stats_[STAT_T_RPA_HIST]      = T->between(8, 9);
  stats_[STAT_T_RPA_SCATTER1]  = T->between(10, 11);
  stats_[STAT_T_RPA_SCATTER2]  = T->between(11, 12);
  stats_[STAT_T_RPA_AGGREGATE] = T->between(12, 13);
  if (mem_) {
    for (int i = 0; i < num_keys_; i++) mem_->out_keys[i] = res_->keys[size_t(i)].col;
    int c = 0;
    for (auto& rc : res_->results)
      for (auto& col : rc) mem_->out_results[c++] = col.col;
  }
}

std::unique_ptr<mig_groupby_result> groupby_call::run()
{
  std::fill(stats_, stats_ + STAT_COUNT, 0.0);
  (void)hipGetLastError();  // a launch error reported below belongs to this call
  validate();
  S.emplace(stream_, mem_ ? mem_->arena : nullptr, mem_ ? mem_->arena_bytes : 0);
  T.emplace(stream_, mem_ == nullptr);
  res_ = std::make_unique<mig_groupby_result>();
  res_->alloc.stream   = stream_;
  res_->alloc.borrowed = mem_ != nullptr;
  if (mr_) {
    res_->alloc.mr     = *mr_;
    res_->alloc.has_mr = true;
  }
  if (N == 0) {
    empty_result();
    return std::move(res_);
  }
  if (needs_sort_)
    return sort_groupby(keys_, num_keys_, null_policy_, requests_, num_requests_, stream_, mr_, keys_sorted_);

  P_ptr_ = std::make_unique<gb_params>();
  Pp_    = P_ptr_.get();
  std::memset(Pp_, 0, sizeof(gb_params));
  P().num_rows = N;
  P().num_keys = num_keys_;
  plan_keys();
  plan_accumulators();

  dP  = S->get<gb_params>(1);
  pin = mem_ ? host_staging() : static_cast<uint8_t*>(thread_pinned().get(HOST_STAGING_BYTES));
  T->mark(0);
  {
    func_range range("groupby: estimate groups");
    estimate_groups();
  }
  T->mark(1);

  int64_t G = -1;
  if (keys_sorted_) {
    func_range range("groupby: sorted runs");
    G = aggregate_sorted_runs();
  }
  if (G < 0) {
    func_range range("groupby: partitioned");
    G = aggregate_partitioned();
  }
  if (G < 0) {
    func_range range("groupby: device-wide table");
    G = aggregate_global();
  }
  func_range range("groupby: finish");
  finish(G);
  return std::move(res_);
}

}  // namespace

std::unique_ptr<mig_groupby_result> groupby_aggregate(mig_column_view const* keys, int32_t num_keys,
                                                      int32_t null_policy,
                                                      mig_aggregation_request const* requests,
                                                      int32_t num_requests, hipStream_t stream,
                                                      mig_memory_resource const* mr, bool keys_sorted)
{
  func_range range("cudf::groupby::groupby::aggregate");
  groupby_call call(keys, num_keys, null_policy, requests, num_requests, stream, mr, keys_sorted, nullptr);
  return call.run();
}

int32_t groupby_aggregate_into(mig_column_view const* keys, int32_t num_keys, int32_t null_policy,
                               mig_aggregation_request const* requests, int32_t num_requests, hipStream_t stream,
                               call_memory const& mem)
{
  func_range range("cudf::groupby::groupby::aggregate (allocation-free)");
  groupby_call call(keys, num_keys, null_policy, requests, num_requests, stream, nullptr, false, &mem);
  auto res = call.run();
  return res->num_groups;
}

}  // namespace mig

mig_groupby_result::~mig_groupby_result()
{
  auto release = [&](mig::owned_column& c) {
    alloc.deallocate(c.col.data, c.data_bytes);
    alloc.deallocate(c.col.null_mask, c.mask_bytes);
    alloc.deallocate(c.col.offsets, c.offsets_bytes);
    c = mig::owned_column{};
  };
  for (auto& k : keys) release(k);
  for (auto& r : results)
    for (auto& c : r) release(c);
}
