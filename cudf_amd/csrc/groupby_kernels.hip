// gfx950 kernels of the hash group-by.
//
// Row hashing follows cudf/table/experimental/row_operators.cuh:1858-1883 and
// the MurmurHash3_x86_32 element hash; equality follows row_operators.cuh:
// 1352-1364, 1429-1448 (nulls equal, NaN equal NaN, -0.0 equal 0.0).  The
// accumulate step replaces compute_single_pass_aggs_fn
// (groupby/hash/groupby_kernels.cuh:98-105) and the variance pass replaces
// var_hash_functor (groupby/hash/multi_pass_kernels.cuh:74-111).
#include "device_ops.hpp"

#include <map>
#include <mutex>
#include <type_traits>

namespace mig {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------
// public row hash: cudf::hashing::murmurhash3_x86_32 (src/hash/murmurhash3_x86_32.cu:34-57)
__global__ void __launch_bounds__(BLOCK)
  k_hash_rows(dcol const* __restrict__ cols, int ncols, int64_t n, uint32_t seed, uint32_t* out)
{
  for (int64_t r = int64_t(blockIdx.x) * BLOCK + threadIdx.x; r < n; r += int64_t(gridDim.x) * BLOCK)
    out[r] = row_hash(cols, ncols, r, seed);
}

void launch_hash_rows(dcol const* d_cols, int ncols, int64_t n, uint32_t seed, uint32_t* out,
                      hipStream_t s)
{
  if (n <= 0) return;
  int64_t blocks = (n + BLOCK - 1) / BLOCK;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_hash_rows, dim3(blocks), dim3(BLOCK), 0, s, d_cols, ncols, n, seed, out);
}

// ---------------------------------------------------------------------------
// cardinality sample: row hashes of `nsample` evenly strided rows
// A pseudo-random permutation of [0, n): a 4-round Feistel network on the
// smallest even number of bits covering n (round function: MurmurHash3's
// finaliser), walking the cycle until the value falls below n.
__device__ __forceinline__ int64_t sample_row(int64_t i, int64_t n)
{
  int bits = 2;
  while ((int64_t(1) << bits) < n) bits += 2;
  int const half      = bits / 2;
  uint64_t const mask = (uint64_t(1) << half) - 1;
  uint64_t x          = uint64_t(i);
  for (int walk = 0; walk < 256; walk++) {
    uint64_t l = x >> half, r = x & mask;
    for (uint32_t round = 0; round < 4; round++) {
      uint64_t const f = mm_fmix32(uint32_t(r) * 0x9E3779B9u + round * 0x85EBCA6Bu + uint32_t(r >> 32));
      uint64_t const t = l ^ (f & mask);
      l                = r;
      r                = t;
    }
    x = (l << half) | r;
    if (x < uint64_t(n)) return int64_t(x);
  }
  return (i * 7919) % n;  // not reached in practice: the cycle through i < n returns below n
}

__global__ void __launch_bounds__(BLOCK)
  k_sample_hash(gb_params const* __restrict__ Pp, int64_t n, int64_t nsample, uint64_t* table,
                uint32_t table_mask, uint32_t* counters)
{
  gb_params const& P = *Pp;
  int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (i >= nsample) return;
  // every row when the sample covers the input; otherwise a uniformly random
  // subset (sample index -> row through a pseudo-random permutation of
  // [0, n)): evenly strided rows would all fall in different runs of sorted
  // or clustered keys and look like all-distinct keys, and rows drawn with
  // replacement would repeat rows and look like repeated keys
  int64_t row = i;
  if (nsample < n) row = sample_row(i, n);
  // rows the group-by drops (null keys under EXCLUDE) are not part of the
  // population whose distinct keys are estimated
  if (P.skip_null_keys && any_key_null(P, row)) return;
  aadd(counters, 1u);
  // a slot holds (rows sharing the hash << 32) | hash; 0 = empty
  uint32_t const h = row_hash(P.keys, P.num_keys, row, 0);
  uint32_t s       = mm_fmix32(h) & table_mask;
  for (uint32_t k = 0; k <= table_mask; k++) {
    uint64_t cur = aload(table + s);
    if (cur == 0) {
      uint64_t const old = acas64(table + s, 0ull, (1ull << 32) | h);
      if (old == 0) {
        aadd(counters + 1, 1u);
        amax(counters + 2, 1u);
        return;
      }
      cur = old;
    }
    if (uint32_t(cur) == h) {
      uint64_t const before = __hip_atomic_fetch_add(table + s, 1ull << 32, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
      amax(counters + 2, uint32_t(before >> 32) + 1u);
      return;
    }
    s = (s + 1) & table_mask;
  }
}

// Frequencies of frequencies of the sample: counters[3] = hashes seen once,
// counters[4] = hashes seen twice.
__global__ void __launch_bounds__(BLOCK)
  k_sample_freqs(uint64_t const* __restrict__ table, uint32_t slots, uint32_t* counters)
{
  __shared__ uint32_t red[BLOCK / 64];
  uint32_t once = 0, twice = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < slots; i += gridDim.x * BLOCK) {
    uint32_t const c = uint32_t(table[i] >> 32);
    once += c == 1u;
    twice += c == 2u;
  }
  uint32_t const t1 = block_sum_u32<BLOCK>(once, red);
  uint32_t const t2 = block_sum_u32<BLOCK>(twice, red);
  if (threadIdx.x == 0) {
    if (t1) aadd(counters + 3, t1);
    if (t2) aadd(counters + 4, t2);
  }
}

void launch_sample_hash(gb_params const* dP, int64_t n, int64_t nsample, uint64_t* table,
                        uint32_t table_mask, uint32_t* counters, hipStream_t s)
{
  int64_t blocks = (nsample + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL(k_sample_hash, dim3(blocks), dim3(BLOCK), 0, s, dP, n, nsample, table, table_mask,
                     counters);
  uint32_t const slots = table_mask + 1u;
  hipLaunchKernelGGL(k_sample_freqs, dim3(std::min<uint32_t>(256u, (slots + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s,
                     table, slots, counters);
}

// ---------------------------------------------------------------------------
// state initialisation: table, row counts, special flags and accumulator identities

__global__ void __launch_bounds__(BLOCK) k_init_state(gb_params const* __restrict__ Pp)
{
  gb_params const& P = *Pp;
  int64_t const total = int64_t(P.cap) + 2;
  int const which = blockIdx.y;  // 0 = table, 1 = row counts, 2.. = fields
  int64_t const stride = int64_t(gridDim.x) * BLOCK;
  int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (which == 0) {
    if (P.key_mode == KM_INLINE) {
      for (int64_t j = i; j < int64_t(P.cap); j += stride) P.slot_keys[j] = EMPTY_KEY;
    } else {
      for (int64_t j = i; j < int64_t(P.cap); j += stride) P.slot_rows[j] = EMPTY_ROW;
    }
    if (i == 0) {
      P.special[0] = 0;
      P.special[1] = 0;
      *P.overflow  = 0;
    }
  } else if (which == 1) {
    if (P.slot_count)
      for (int64_t j = i; j < total; j += stride) *count_at(P, j) = 0;
  } else {
    field_desc const& F = P.fields[which - 2];
    uint64_t const id = field_identity(F.kind);
    if (F.width == 8) {
      for (int64_t j = i; j < total; j += stride) *reinterpret_cast<uint64_t*>(field_at(F, j)) = id;
    } else {
      for (int64_t j = i; j < total; j += stride) *reinterpret_cast<uint32_t*>(field_at(F, j)) = uint32_t(id);
    }
  }
}

void launch_init_state(gb_params const* dP, gb_params const& hP, hipStream_t s)
{
  int64_t total = int64_t(hP.cap) + 2;
  int64_t blocks = (total + BLOCK - 1) / BLOCK;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_init_state, dim3(blocks, 2 + hP.num_fields), dim3(BLOCK), 0, s, dP);
}

// ---------------------------------------------------------------------------
// hash table probing
__device__ __forceinline__ void mark_special(gb_params const& P, int which)
{
  if (aload(&P.special[which]) == 0u) astore(&P.special[which], 1u);
}

// Insert-or-find of an inline key; returns the slot, or -1 on overflow.
__device__ __forceinline__ int64_t insert_inline(gb_params const& P, uint64_t key, uint32_t h)
{
  if (key == EMPTY_KEY) {
    mark_special(P, 0);
    return int64_t(P.cap);
  }
  uint64_t s = h & P.cap_mask;
  for (uint32_t i = 0; i < P.probe_limit; i++) {
    if ((i & 63u) == 63u && aload(P.overflow) != 0) return -1;  // the table is being rebuilt
    uint64_t cur = aload(&P.slot_keys[s]);
    if (cur == key) return int64_t(s);
    if (cur == EMPTY_KEY) {
      uint64_t old = acas64(&P.slot_keys[s], EMPTY_KEY, key);
      if (old == EMPTY_KEY || old == key) return int64_t(s);
    }
    s = (s + 1) & P.cap_mask;
  }
  astore(P.overflow, 1);
  return -1;
}

__device__ __forceinline__ int64_t find_inline(gb_params const& P, uint64_t key, uint32_t h)
{
  if (key == EMPTY_KEY) return int64_t(P.cap);
  uint64_t s = h & P.cap_mask;
  for (uint32_t i = 0; i < P.probe_limit; i++) {
    uint64_t cur = P.slot_keys[s];
    if (cur == key) return int64_t(s);
    if (cur == EMPTY_KEY) return -1;
    s = (s + 1) & P.cap_mask;
  }
  return -1;
}

__device__ __forceinline__ int64_t insert_general(gb_params const& P, int64_t row, uint32_t h)
{
  uint64_t const mine = (uint64_t(h) << 32) | uint32_t(row);
  uint64_t s = h & P.cap_mask;
  for (uint32_t i = 0; i < P.probe_limit; i++) {
    if ((i & 63u) == 63u && aload(P.overflow) != 0) return -1;  // the table is being rebuilt
    uint64_t cur = aload(&P.slot_rows[s]);
    if (cur == EMPTY_ROW) {
      uint64_t old = acas64(&P.slot_rows[s], EMPTY_ROW, mine);
      if (old == EMPTY_ROW) return int64_t(s);
      cur = old;
    }
    if (uint32_t(cur >> 32) == h && rows_equal(P, int64_t(uint32_t(cur)), row)) return int64_t(s);
    s = (s + 1) & P.cap_mask;
  }
  astore(P.overflow, 1);
  return -1;
}

__device__ __forceinline__ int64_t find_general(gb_params const& P, int64_t row, uint32_t h)
{
  uint64_t s = h & P.cap_mask;
  for (uint32_t i = 0; i < P.probe_limit; i++) {
    uint64_t cur = P.slot_rows[s];
    if (cur == EMPTY_ROW) return -1;
    if (uint32_t(cur >> 32) == h && rows_equal(P, int64_t(uint32_t(cur)), row)) return int64_t(s);
    s = (s + 1) & P.cap_mask;
  }
  return -1;
}

// Slot of `row`'s group (inserting it when INSERT), or -1 when the row is
// skipped (null key under EXCLUDE) or the table overflowed.
template <int KM, bool INSERT>
__device__ __forceinline__ int64_t locate(gb_params const& P, int64_t row)
{
  if constexpr (KM == KM_INLINE) {
    dcol const& kc = P.keys[0];
    if (!dcol_valid(kc, row)) {
      if (P.skip_null_keys) return -1;
      if (INSERT) mark_special(P, 1);
      return int64_t(P.cap) + 1;
    }
    uint64_t const bits = load_fixed_bits(kc, row);
    uint32_t const h    = hash_combine(0u, hash_fixed_bits(kc.storage, bits, 0u));
    return INSERT ? insert_inline(P, bits, h) : find_inline(P, bits, h);
  } else {
    if (P.skip_null_keys && any_key_null(P, row)) return -1;
    uint32_t const h = row_hash(P.keys, P.num_keys, row, 0u);
    return INSERT ? insert_general(P, row, h) : find_general(P, row, h);
  }
}

// ---------------------------------------------------------------------------
// accumulate one row into the group's accumulators (apply_field: device_ops.hpp)

// ARGMIN / ARGMAX: is row r's value a better choice than row o's?  The
// reference keeps the first of equal values to reach the target
// (aggregation.cuh:440-490); here equal values take the smaller row, and a
// NaN only wins when both are NaN (NaN compares false in the reference, so a
// number there is never replaced by a NaN), which makes the result
// deterministic.
__device__ __forceinline__ bool arg_better(dcol const& c, int32_t kind, int64_t r, int64_t o)
{
  int cmp = 0;
  if (c.storage == ST_STRING) {  // bytewise, then by length (string_view::compare)
    int64_t rb, re, ob, oe;
    str_span(c, r, rb, re);
    str_span(c, o, ob, oe);
    int64_t const lr = re - rb, lo = oe - ob, m = lr < lo ? lr : lo;
    auto const d     = gp(c.data);
    for (int64_t i = 0; i < m && cmp == 0; i++) {
      uint8_t const x = d[rb + i], y = d[ob + i];
      cmp             = x < y ? -1 : (x > y ? 1 : 0);
    }
    if (cmp == 0) cmp = lr < lo ? -1 : (lr > lo ? 1 : 0);
    if (cmp != 0) return kind == FK_ARGMIN ? cmp < 0 : cmp > 0;
    return r < o;
  }
  loaded_value const a = load_value(c, r), b = load_value(c, o);
  if (c.storage == ST_F32 || c.storage == ST_F64) {
    bool const an = isnan(a.d), bn = isnan(b.d);
    if (an != bn) return bn;  // a number beats a NaN
    if (!an) cmp = a.d < b.d ? -1 : (a.d > b.d ? 1 : 0);
  } else if (is_unsigned_storage_dev(c.storage)) {
    cmp = uint64_t(a.i) < uint64_t(b.i) ? -1 : (uint64_t(a.i) > uint64_t(b.i) ? 1 : 0);
  } else {
    cmp = a.i < b.i ? -1 : (a.i > b.i ? 1 : 0);
  }
  if (cmp != 0) return kind == FK_ARGMIN ? cmp < 0 : cmp > 0;
  return r < o;
}

// Compare-and-swap loop on the group's row index (aggregation.cuh:440-490).
__device__ __forceinline__ void arg_update(field_desc const& F, int64_t slot, dcol const& c, int64_t row)
{
  auto* a      = reinterpret_cast<unsigned int*>(field_at(F, slot));
  uint32_t cur = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (cur == 0xFFFFFFFFu || arg_better(c, F.kind, row, int64_t(cur))) {
    uint32_t const prev = atomicCAS(a, cur, uint32_t(row));
    if (prev == cur) break;
    cur = prev;
  }
}

__device__ __forceinline__ void update_row(gb_params const& P, int64_t slot, int64_t row)
{
  if (P.need_count_all) aadd(count_at(P, slot), 1u);
  for (int vc = 0; vc < P.num_vals; vc++) {
    int const fb = P.val_field_begin[vc], fe = P.val_field_begin[vc + 1];
    if (fb == fe) continue;
    dcol const& c = P.vals[vc];
    if (!dcol_valid(c, row)) continue;
    loaded_value const v = load_value(c, row);
    for (int f = fb; f < fe; f++) {
      field_desc const& F = P.fields[f];
      if (F.kind == FK_ARGMIN || F.kind == FK_ARGMAX) arg_update(F, slot, c, row);
      else apply_field(F, slot, v);
    }
  }
}

template <int KM>
__global__ void __launch_bounds__(BLOCK) k_aggregate_global(gb_params const* __restrict__ Pp)
{
  gb_params const& P = *Pp;
  int64_t const n = P.num_rows;
  for (int64_t row = int64_t(blockIdx.x) * BLOCK + threadIdx.x; row < n;
       row += int64_t(gridDim.x) * BLOCK) {
    if (aload(P.overflow) != 0) break;  // the call rebuilds a larger table: stop early
    int64_t const slot = locate<KM, true>(P, row);
    if (slot < 0) continue;
    update_row(P, slot, row);
  }
}

static int64_t row_grid(int64_t n, int num_cus, int per_cu)
{
  int64_t blocks = (n + BLOCK - 1) / BLOCK;
  int64_t cap = int64_t(num_cus) * per_cu;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  return blocks;
}

template <int KM>
__global__ void __launch_bounds__(BLOCK) k_aggregate_global_combined(gb_params const* __restrict__ Pp);

// Whether every accumulator of the plan combines partials (the wave-combined
// form of k_aggregate_global) and slot numbers fit its 26-bit sort keys.
static bool combinable(gb_params const& hP)
{
  if (hP.cap + 2 >= (uint64_t(1) << 26)) return false;
  for (int f = 0; f < hP.num_fields; f++) {
    switch (hP.fields[f].kind) {
      case FK_ARGMIN:
      case FK_ARGMAX:
      case FK_NUNIQUE:
      case FK_NUNIQUE_NULLS:
      case FK_NTH: return false;
      default: break;
    }
  }
  return true;
}

void launch_aggregate(gb_params const* dP, gb_params const& hP, int num_cus, hipStream_t s)
{
  if (hP.num_rows <= 0) return;
  int64_t blocks = row_grid(hP.num_rows, num_cus, 16);
  if (combinable(hP)) {
    if (hP.key_mode == KM_INLINE)
      hipLaunchKernelGGL(k_aggregate_global_combined<KM_INLINE>, dim3(blocks), dim3(BLOCK), 0, s, dP);
    else
      hipLaunchKernelGGL(k_aggregate_global_combined<KM_GENERAL>, dim3(blocks), dim3(BLOCK), 0, s, dP);
    return;
  }
  if (hP.key_mode == KM_INLINE)
    hipLaunchKernelGGL(k_aggregate_global<KM_INLINE>, dim3(blocks), dim3(BLOCK), 0, s, dP);
  else
    hipLaunchKernelGGL(k_aggregate_global<KM_GENERAL>, dim3(blocks), dim3(BLOCK), 0, s, dP);
}

// ---------------------------------------------------------------------------
// Group labels for the sort-based path: every row's slot in the device-wide
// table (UINT32_MAX for a dropped row), then each slot's rank among the
// occupied slots (the compaction order), then every row's group number.
template <int KM>
__global__ void __launch_bounds__(BLOCK) k_label_rows(gb_params const* __restrict__ Pp, uint32_t* __restrict__ slot_of_row)
{
  gb_params const& P = *Pp;
  int64_t const n    = P.num_rows;
  for (int64_t row = int64_t(blockIdx.x) * BLOCK + threadIdx.x; row < n; row += int64_t(gridDim.x) * BLOCK) {
    int64_t const slot = locate<KM, true>(P, row);
    gp(slot_of_row)[row] = slot < 0 ? 0xFFFFFFFFu : uint32_t(slot);
  }
}

void launch_label_rows(gb_params const* dP, gb_params const& hP, int num_cus, uint32_t* slot_of_row, hipStream_t s)
{
  if (hP.num_rows <= 0) return;
  int64_t blocks = row_grid(hP.num_rows, num_cus, 16);
  if (hP.key_mode == KM_INLINE)
    hipLaunchKernelGGL(k_label_rows<KM_INLINE>, dim3(blocks), dim3(BLOCK), 0, s, dP, slot_of_row);
  else
    hipLaunchKernelGGL(k_label_rows<KM_GENERAL>, dim3(blocks), dim3(BLOCK), 0, s, dP, slot_of_row);
}

__global__ void __launch_bounds__(BLOCK)
  k_rank_slots(uint32_t const* __restrict__ map, int64_t G, uint32_t* __restrict__ rank_of_slot)
{
  int64_t const j = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (j < G) gp(rank_of_slot)[gp(map)[j]] = uint32_t(j);
}

__global__ void __launch_bounds__(BLOCK)
  k_gid_of_rows(uint32_t const* __restrict__ slot_of_row, uint32_t const* __restrict__ rank_of_slot, int64_t n,
                uint32_t dropped, uint32_t* __restrict__ gid)
{
  int64_t const r = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (r >= n) return;
  uint32_t const s = gp(slot_of_row)[r];
  gp(gid)[r]       = s == 0xFFFFFFFFu ? dropped : gp(rank_of_slot)[s];
}

void launch_group_numbers(uint32_t const* map, int64_t G, uint32_t const* slot_of_row, int64_t n,
                          uint32_t* rank_of_slot, uint32_t* gid, hipStream_t s)
{
  if (G > 0) hipLaunchKernelGGL(k_rank_slots, dim3((G + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, map, G, rank_of_slot);
  if (n > 0)
    hipLaunchKernelGGL(k_gid_of_rows, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, slot_of_row, rank_of_slot, n,
                       uint32_t(G), gid);
}

// ---------------------------------------------------------------------------
// second pass: sum of squared deviations from the group mean (VARIANCE / STD)
__device__ __forceinline__ double field_as_double(field_desc const& F, int64_t slot)
{
  switch (F.kind) {
    case FK_PROD_I64: return double(*reinterpret_cast<int64_t const*>(field_at(F, slot)));
    case FK_PROD_F32: return double(*reinterpret_cast<float const*>(field_at(F, slot)));
    case FK_PROD_F64: return *reinterpret_cast<double const*>(field_at(F, slot));
    case FK_SUM_I64: return double(*reinterpret_cast<int64_t const*>(field_at(F, slot)));
    case FK_SUM_F32: return double(*reinterpret_cast<float const*>(field_at(F, slot)));
    case FK_SUM_F64: return *reinterpret_cast<double const*>(field_at(F, slot));
    case FK_COUNT: return double(*reinterpret_cast<uint32_t const*>(field_at(F, slot)));
    default: return 0.0;
  }
}

// Count accumulator `f`, or the group row count when f == -2 (columns
// without nulls count their values with the row count).
__device__ __forceinline__ uint32_t count_of(gb_params const& P, int32_t f, int64_t slot)
{
  return f == -2 ? *count_at(P, slot) : *reinterpret_cast<uint32_t const*>(field_at(P.fields[f], slot));
}

template <int KM>
__global__ void __launch_bounds__(BLOCK) k_second_pass(gb_params const* __restrict__ Pp)
{
  gb_params const& P = *Pp;
  int64_t const n = P.num_rows;
  for (int64_t row = int64_t(blockIdx.x) * BLOCK + threadIdx.x; row < n;
       row += int64_t(gridDim.x) * BLOCK) {
    int64_t slot = -2;
    for (int vc = 0; vc < P.num_vals; vc++) {
      int const fb = P.val_field_begin[vc], fe = P.val_field_begin[vc + 1];
      bool has_m2 = false;
      for (int f = fb; f < fe; f++) has_m2 |= (P.fields[f].kind == FK_M2);
      if (!has_m2) continue;
      dcol const& c = P.vals[vc];
      if (!dcol_valid(c, row)) continue;
      if (slot == -2) slot = locate<KM, false>(P, row);
      if (slot < 0) break;
      loaded_value const v = load_value(c, row);
      for (int f = fb; f < fe; f++) {
        field_desc const& F = P.fields[f];
        if (F.kind != FK_M2) continue;
        double const cnt  = double(count_of(P, F.aux_count, slot));
        double const mean = field_as_double(P.fields[F.aux_sum], slot) / cnt;
        double const d    = v.d - mean;
        gadd_f64(reinterpret_cast<double*>(field_at(F, slot)), d * d);
      }
    }
  }
}

void launch_second_pass(gb_params const* dP, gb_params const& hP, int num_cus, hipStream_t s)
{
  if (hP.num_rows <= 0) return;
  int64_t blocks = row_grid(hP.num_rows, num_cus, 16);
  if (hP.key_mode == KM_INLINE)
    hipLaunchKernelGGL(k_second_pass<KM_INLINE>, dim3(blocks), dim3(BLOCK), 0, s, dP);
  else
    hipLaunchKernelGGL(k_second_pass<KM_GENERAL>, dim3(blocks), dim3(BLOCK), 0, s, dP);
}

// ---------------------------------------------------------------------------
// Low cardinality (one fixed-width key, few groups): per-workgroup partial
// tables in LDS merged into the global table once per workgroup, so a hot
// group costs one global atomic per workgroup and field instead of one per
// row.  Rows whose key finds the LDS table full go straight to the global
// table.  The variance pass keys an LDS table by group slot.
// Workgroup of the LDS-table kernels: their tables (up to 96 KB) allow two
// workgroups per CU, so the waves per CU come from the workgroup size
// (measured at 40M rows / 1,000 keys: 256 threads 1.03 ms, 512 0.61 ms,
// 1024 0.63 ms; 1M rows: 0.19, 0.11, 0.07 ms).
constexpr int LBLOCK   = 1024;
constexpr int LDS_ROWS = 4;  // rows per thread in flight in k_aggregate_lds

__global__ void __launch_bounds__(LBLOCK) k_aggregate_lds(gb_params const* __restrict__ Pp)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ uint32_t fill_sh, present_sh[2];
  gb_params const& P = *Pp;
  constexpr int WG   = __HIP_MEMORY_SCOPE_WORKGROUP;
  int const S        = P.lds_slots;
  auto* tkeys        = reinterpret_cast<uint64_t*>(lds);
  for (int i = threadIdx.x; i < S; i += LBLOCK) tkeys[i] = EMPTY_KEY;
  for (int f = 0; f < P.num_fields; f++) {
    field_desc const& F = P.fields[f];
    if (F.kind == FK_M2) continue;
    uint64_t const idv = field_identity(F.kind);
    if (F.width == 8)
      for (int i = threadIdx.x; i < S + 2; i += LBLOCK) *lds_at<uint64_t>(lds, F.lds_off, i) = idv;
    else
      for (int i = threadIdx.x; i < S + 2; i += LBLOCK) *lds_at<uint32_t>(lds, F.lds_off, i) = uint32_t(idv);
  }
  if (P.lds_count_off >= 0)
    for (int i = threadIdx.x; i < S + 2; i += LBLOCK) *lds_at<uint32_t>(lds, P.lds_count_off, i) = 0;
  if (threadIdx.x == 0) {
    fill_sh       = 0;
    present_sh[0] = 0;
    present_sh[1] = 0;
  }
  __syncthreads();
  dcol const& kc = P.keys[0];
  // LDS_ROWS rows per thread at a time: their key loads, then their value
  // loads column by column, are in flight together (two workgroups per CU
  // leave too few waves to hide memory latency one row at a time)
  constexpr int U      = LDS_ROWS;
  int64_t const stride = int64_t(gridDim.x) * LBLOCK * U;
  for (int64_t r0 = int64_t(blockIdx.x) * LBLOCK * U + threadIdx.x; r0 < P.num_rows; r0 += stride) {
    if (aload(P.overflow) != 0) break;  // the call rebuilds a larger table: stop early
    bool kv[U];
    uint64_t bits[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int64_t const row = r0 + int64_t(u) * LBLOCK;
      int64_t const rc  = row < P.num_rows ? row : P.num_rows - 1;  // in bounds; unused past the end
      kv[u]             = dcol_valid(kc, rc);
      bits[u]           = load_fixed_bits(kc, rc);
    }
    int id[U];  // LDS slot, -1: the device-wide table, -2: no row / dropped
#pragma unroll
    for (int u = 0; u < U; u++) {
      int64_t const row = r0 + int64_t(u) * LBLOCK;
      id[u]             = -2;
      if (row >= P.num_rows) continue;
      if (!kv[u]) {
        if (P.skip_null_keys) continue;
        id[u]         = S + 1;
        present_sh[1] = 1;
      } else if (bits[u] == EMPTY_KEY) {
        id[u]         = S;
        present_sh[0] = 1;
      } else {
        uint32_t const h = hash_combine(0u, hash_fixed_bits(kc.storage, bits[u], 0u));
        id[u]            = lds_insert(tkeys, &fill_sh, bits[u], h, EMPTY_KEY, S, uint32_t(P.lds_fill));
      }
    }
    if (P.lds_count_off >= 0) {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (id[u] >= 0)
          __hip_atomic_fetch_add(lds_at<uint32_t>(lds, P.lds_count_off, id[u]), 1u, __ATOMIC_RELAXED, WG);
    }
    for (int vc = 0; vc < P.num_vals; vc++) {
      int const fb = P.val_field_begin[vc], fe = P.val_field_begin[vc + 1];
      if (fb == fe) continue;
      dcol const& c = P.vals[vc];
      loaded_value v[U];
      uint32_t ok = 0;
#pragma unroll
      for (int u = 0; u < U; u++) {
        int64_t const row = r0 + int64_t(u) * LBLOCK;
        v[u]              = loaded_value{0, 0.0, 0.0f, 0};
        if (id[u] >= 0 && dcol_valid(c, row)) {
          v[u] = load_value(c, row);
          ok |= 1u << u;
        }
      }
      for (int f = fb; f < fe; f++) {
        field_desc const& F = P.fields[f];
        if (F.kind == FK_M2) continue;
#pragma unroll
        for (int u = 0; u < U; u++)
          if ((ok >> u) & 1u) lds_apply(F.kind, lds + F.lds_off + int64_t(id[u]) * F.width, v[u]);
      }
    }
    // keys that found no room in the LDS table update the device-wide table
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (id[u] != -1) continue;
      int64_t const row  = r0 + int64_t(u) * LBLOCK;
      int64_t const slot = insert_inline(P, bits[u], hash_combine(0u, hash_fixed_bits(kc.storage, bits[u], 0u)));
      if (slot >= 0) update_row(P, slot, row);
    }
  }
  __syncthreads();
  if (aload(P.overflow) != 0) return;  // this table is discarded
  // merge the partials into the global table
  for (int i = threadIdx.x; i < S + 2; i += LBLOCK) {
    int64_t slot;
    if (i < S) {
      uint64_t const k = tkeys[i];
      if (k == EMPTY_KEY) continue;
      slot = insert_inline(P, k, hash_combine(0u, hash_fixed_bits(kc.storage, k, 0u)));
    } else if (i == S) {
      if (!present_sh[0]) continue;
      slot = insert_inline(P, EMPTY_KEY, 0u);
    } else {
      if (!present_sh[1]) continue;
      mark_special(P, 1);
      slot = int64_t(P.cap) + 1;
    }
    if (slot < 0) continue;
    if (P.lds_count_off >= 0) aadd(count_at(P, slot), *lds_at<uint32_t>(lds, P.lds_count_off, i));
    for (int f = 0; f < P.num_fields; f++) {
      field_desc const& F = P.fields[f];
      if (F.kind == FK_M2) continue;
      uint64_t const part = F.width == 8 ? *lds_at<uint64_t>(lds, F.lds_off, i)
                                         : uint64_t(*lds_at<uint32_t>(lds, F.lds_off, i));
      // an identity partial changes nothing (float sums start at +0.0)
      if (part != field_identity(F.kind))
        acc_combine(F.kind, field_at(F, slot), part);
    }
  }
}

__global__ void __launch_bounds__(LBLOCK) k_second_pass_lds(gb_params const* __restrict__ Pp)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ uint32_t fill_sh;
  gb_params const& P = *Pp;
  constexpr int WG   = __HIP_MEMORY_SCOPE_WORKGROUP;
  constexpr uint64_t NO = ~0ull;  // free slot of the LDS table of group slots
  int const S        = P.lds_slots;
  auto* tslots       = reinterpret_cast<uint64_t*>(lds);
  for (int i = threadIdx.x; i < S; i += LBLOCK) tslots[i] = NO;
  for (int f = 0; f < P.num_fields; f++) {
    if (P.fields[f].kind != FK_M2) continue;
    for (int i = threadIdx.x; i < S; i += LBLOCK) *lds_at<double>(lds, P.lds_m2_off[f], i) = 0.0;
  }
  if (threadIdx.x == 0) fill_sh = 0;
  __syncthreads();
  for (int64_t row = int64_t(blockIdx.x) * LBLOCK + threadIdx.x; row < P.num_rows;
       row += int64_t(gridDim.x) * LBLOCK) {
    int64_t slot = -2;
    int id       = -1;
    for (int vc = 0; vc < P.num_vals; vc++) {
      int const fb = P.val_field_begin[vc], fe = P.val_field_begin[vc + 1];
      bool has_m2  = false;
      for (int f = fb; f < fe; f++) has_m2 |= (P.fields[f].kind == FK_M2);
      if (!has_m2) continue;
      dcol const& c = P.vals[vc];
      if (!dcol_valid(c, row)) continue;
      if (slot == -2) {
        slot = locate<KM_INLINE, false>(P, row);
        if (slot >= 0) id = lds_insert(tslots, &fill_sh, uint64_t(slot), mm_fmix32(uint32_t(slot)), NO, S,
                                       uint32_t(P.lds_fill));
      }
      if (slot < 0) break;
      loaded_value const v = load_value(c, row);
      for (int f = fb; f < fe; f++) {
        field_desc const& F = P.fields[f];
        if (F.kind != FK_M2) continue;
        double const cnt  = double(count_of(P, F.aux_count, slot));
        double const mean = field_as_double(P.fields[F.aux_sum], slot) / cnt;
        double const d    = v.d - mean;
        if (id >= 0) __hip_atomic_fetch_add(lds_at<double>(lds, P.lds_m2_off[f], id), d * d, __ATOMIC_RELAXED, WG);
        else gadd_f64(reinterpret_cast<double*>(field_at(F, slot)), d * d);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < S; i += LBLOCK) {
    uint64_t const slot = tslots[i];
    if (slot == NO) continue;
    for (int f = 0; f < P.num_fields; f++) {
      if (P.fields[f].kind != FK_M2) continue;
      double const part = *lds_at<double>(lds, P.lds_m2_off[f], i);
      if (part != 0.0) gadd_f64(reinterpret_cast<double*>(field_at(P.fields[f], int64_t(slot))), part);
    }
  }
}

void allow_dynamic_lds(void const* fn, int bytes)
{
  if (bytes <= 64 * 1024) return;
  // Raised once per kernel and device (and only upwards): setting a function
  // attribute is not a stream operation, but other threads may be capturing
  // graphs, so it is kept off the per-call path.
  static std::mutex mu;
  static std::map<std::pair<void const*, int>, int> done;
  int dev = 0;
  MIG_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  int& cur = done[{fn, dev}];
  if (bytes <= cur) return;
  MIG_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  cur = bytes;
}

void launch_aggregate_lds(gb_params const* dP, gb_params const& hP, int64_t blocks, hipStream_t s)
{
  if (hP.num_rows <= 0) return;
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_aggregate_lds), hP.lds_bytes);
  hipLaunchKernelGGL(k_aggregate_lds, dim3(blocks), dim3(LBLOCK), hP.lds_bytes, s, dP);
}

void launch_second_pass_lds(gb_params const* dP, gb_params const& hP, int64_t blocks, hipStream_t s)
{
  if (hP.num_rows <= 0) return;
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_second_pass_lds), hP.lds_bytes);
  hipLaunchKernelGGL(k_second_pass_lds, dim3(blocks), dim3(LBLOCK), hP.lds_bytes, s, dP);
}

// ---------------------------------------------------------------------------
// compaction of occupied slots -> dense group ids
__device__ __forceinline__ bool slot_occupied(gb_params const& P, uint64_t slot)
{
  if (slot < P.cap) {
    return P.key_mode == KM_INLINE ? (P.slot_keys[slot] != EMPTY_KEY)
                                   : (P.slot_rows[slot] != EMPTY_ROW);
  }
  return P.special[slot - P.cap] != 0u;
}


__global__ void __launch_bounds__(BLOCK)
  k_count_occupied(gb_params const* __restrict__ Pp, uint32_t* block_counts)
{
  __shared__ uint32_t red[BLOCK / 64];
  gb_params const& P = *Pp;
  uint64_t const total = P.cap + 2;
  uint64_t const base  = uint64_t(blockIdx.x) * SCAN_TILE;
  uint32_t c = 0;
  for (int k = threadIdx.x; k < SCAN_TILE; k += BLOCK) {
    uint64_t slot = base + k;
    if (slot < total && slot_occupied(P, slot)) c++;
  }
  c = block_sum_u32(c, red);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = c;
}

void launch_count_occupied(gb_params const* dP, gb_params const& hP, uint32_t* block_counts,
                           hipStream_t s)
{
  int64_t nb = (int64_t(hP.cap) + 2 + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_count_occupied, dim3(nb), dim3(BLOCK), 0, s, dP, block_counts);
}

// Exclusive scan of `n` u32 values in place by one block of 1024 threads.
__global__ void __launch_bounds__(1024) k_scan_small(uint32_t* data, int64_t n, uint32_t* total)
{
  __shared__ uint32_t part[1024];
  int64_t const per = (n + 1023) / 1024;
  int64_t const b   = int64_t(threadIdx.x) * per;
  int64_t const e   = (b + per < n) ? b + per : n;
  uint32_t s = 0;
  for (int64_t i = b; i < e; i++) s += data[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
  for (int64_t i = b; i < e; i++) {
    uint32_t v = data[i];
    data[i]    = run;
    run += v;
  }
  if (threadIdx.x == 1023 && total) *total = part[1023];
}

void launch_scan_blocks(uint32_t* block_counts, int64_t nblocks, uint32_t* total, hipStream_t s)
{
  hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, block_counts, nblocks, total);
}

// Writes the gather map: rep_rows[g] = slot of dense group g (slot order).
__global__ void __launch_bounds__(BLOCK)
  k_compact(gb_params const* __restrict__ Pp, uint32_t const* block_offsets, uint32_t* map)
{
  __shared__ uint32_t wave_tot[BLOCK / 64];
  gb_params const& P = *Pp;
  uint64_t const total = P.cap + 2;
  uint64_t const base  = uint64_t(blockIdx.x) * SCAN_TILE;
  uint32_t running     = block_offsets[blockIdx.x];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = 0; k < SCAN_TILE; k += BLOCK) {
    uint64_t const slot = base + k + threadIdx.x;
    bool const occ     = slot < total && slot_occupied(P, slot);
    uint64_t const bal = __ballot(occ);
    uint32_t const below = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(bal);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < BLOCK / 64; w++) {
      if (w < wave) before += wave_tot[w];
      all += wave_tot[w];
    }
    if (occ) map[running + before + below] = uint32_t(slot);
    running += all;
    __syncthreads();
  }
}

void launch_compact(gb_params const* dP, gb_params const& hP, uint32_t const* block_offsets,
                    uint32_t* map, hipStream_t s)
{
  int64_t nb = (int64_t(hP.cap) + 2 + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_compact, dim3(nb), dim3(BLOCK), 0, s, dP, block_offsets, map);
}

// ---------------------------------------------------------------------------
// dense outputs: keys and aggregation results for group g = blockIdx*BLOCK+tid
__device__ __forceinline__ uint64_t field_bits(field_desc const& F, int64_t slot)
{
  return F.width == 8 ? *reinterpret_cast<uint64_t const*>(field_at(F, slot))
                      : *reinterpret_cast<uint32_t const*>(field_at(F, slot));
}


// Writes the validity bit of group g for the whole wave with one ballot.
__device__ __forceinline__ void store_valid_bits(uint32_t* mask, int64_t g, int64_t G, bool valid)
{
  uint64_t const bal = __ballot(g < G && valid);
  int const lane     = threadIdx.x & 63;
  int64_t const g0   = g - lane;  // multiple of 64
  if (lane == 0 && g0 < G) mask[g0 >> 5] = uint32_t(bal);
  if (lane == 32 && g0 + 32 < G) mask[(g0 >> 5) + 1] = uint32_t(bal >> 32);
}

template <int KM>
__global__ void __launch_bounds__(BLOCK)
  k_outputs(gb_params const* __restrict__ Pp, uint32_t const* __restrict__ map, int64_t G,
            uint32_t* rep_rows)
{
  gb_params const& P = *Pp;
  int64_t const g    = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  bool const live    = g < G;
  int64_t const slot = live ? int64_t(map[g]) : 0;

  // keys
  if constexpr (KM == KM_INLINE) {
    uint64_t bits = 0;
    bool kvalid   = true;
    if (live) {
      if (slot < int64_t(P.cap)) bits = P.slot_keys[slot];
      else if (slot == int64_t(P.cap)) bits = EMPTY_KEY;
      else kvalid = false;
      store_width(P.out_keys[0], g, storage_width(P.key_storage), kvalid ? bits : 0);
    }
    if (P.out_key_mask[0]) store_valid_bits(P.out_key_mask[0], g, G, kvalid);
  } else {
    if (live) rep_rows[g] = uint32_t(P.slot_rows[slot]);
  }

  // aggregation results
  for (int o = 0; o < P.num_outs; o++) {
    out_desc const& O = P.outs[o];
    bool valid        = true;
    if (live) {
      uint32_t cnt = 0;
      if (O.count_field != -1) {
        cnt   = count_of(P, O.count_field, slot);
        valid = cnt > 0;
      }
      switch (O.kind) {
        case OK_COPY_I64:
        case OK_COPY_F64: reinterpret_cast<uint64_t*>(O.data)[g] = field_bits(P.fields[O.field], slot); break;
        case OK_COPY_F32: reinterpret_cast<uint32_t*>(O.data)[g] = uint32_t(field_bits(P.fields[O.field], slot)); break;
        case OK_COUNT:
          reinterpret_cast<int32_t*>(O.data)[g] =
            int32_t(*reinterpret_cast<uint32_t const*>(field_at(P.fields[O.field], slot)));
          valid = true;
          break;
        case OK_COUNT_ALL:
          reinterpret_cast<int32_t*>(O.data)[g] = int32_t(*count_at(P, slot));
          valid = true;
          break;
        case OK_MINMAX: {
          field_desc const& F = P.fields[O.field];
          uint64_t b          = field_bits(F, slot);
          if (F.kind == FK_MIN_F32 || F.kind == FK_MAX_F32) b = unord32(uint32_t(b));
          else if (F.kind == FK_MIN_F64 || F.kind == FK_MAX_F64) b = unord64(b);
          store_width(O.data, g, O.out_width, b);
          break;
        }
        case OK_MEAN: {
          double const sum = field_as_double(P.fields[O.field], slot);
          reinterpret_cast<double*>(O.data)[g] = valid ? sum / double(cnt) : 0.0;
          break;
        }
        case OK_SUM_I32:
          reinterpret_cast<int32_t*>(O.data)[g] = int32_t(uint32_t(field_bits(P.fields[O.field], slot)));
          break;
        case OK_MEAN_INT: {  // fixed-point mean: the representation's sum / count, truncated
          int64_t sum = int64_t(field_bits(P.fields[O.field], slot));
          if (O.out_width == 4) sum = int64_t(int32_t(uint32_t(sum)));
          store_width(O.data, g, O.out_width, uint64_t(valid ? sum / int64_t(cnt) : 0));
          break;
        }
        case OK_ARG: {
          uint32_t const r = *reinterpret_cast<uint32_t const*>(field_at(P.fields[O.field], slot));
          valid            = r != 0xFFFFFFFFu;
          reinterpret_cast<int32_t*>(O.data)[g] = valid ? int32_t(r) : -1;
          break;
        }
        case OK_NTH: {  // gather of the value at the group's n-th row
          field_desc const& F = P.fields[O.field];
          dcol const& c       = P.vals[F.vcol];
          uint32_t const r    = *reinterpret_cast<uint32_t const*>(field_at(F, slot));
          valid               = r != 0xFFFFFFFFu && dcol_valid(c, r);
          uint64_t b          = 0;
          if (valid) {
            auto const d = gp(c.data);
            switch (O.out_width) {
              case 1: b = d[r]; break;
              case 2: b = reinterpret_cast<gmem<uint16_t const>*>(d)[r]; break;
              case 4: b = reinterpret_cast<gmem<uint32_t const>*>(d)[r]; break;
              default: b = reinterpret_cast<gmem<uint64_t const>*>(d)[r]; break;
            }
          }
          store_width(O.data, g, O.out_width, b);
          break;
        }
        case OK_VAR:
        case OK_STD: {
          int64_t const dn = int64_t(cnt) - O.ddof;
          valid            = cnt > 0 && dn > 0;
          double var       = 0.0;
          if (valid) var = *reinterpret_cast<double const*>(field_at(P.fields[O.field], slot)) / double(dn);
          reinterpret_cast<double*>(O.data)[g] = (O.kind == OK_STD) ? sqrt(var) : var;
          break;
        }
        default: break;
      }
    }
    if (O.mask) store_valid_bits(O.mask, g, G, valid);
  }
}

void launch_outputs(gb_params const* dP, gb_params const& hP, uint32_t const* map, int64_t G,
                    uint32_t* rep_rows, hipStream_t s)
{
  if (G <= 0) return;
  int64_t blocks = (G + BLOCK - 1) / BLOCK;
  if (hP.key_mode == KM_INLINE)
    hipLaunchKernelGGL(k_outputs<KM_INLINE>, dim3(blocks), dim3(BLOCK), 0, s, dP, map, G, rep_rows);
  else
    hipLaunchKernelGGL(k_outputs<KM_GENERAL>, dim3(blocks), dim3(BLOCK), 0, s, dP, map, G, rep_rows);
}

// ---------------------------------------------------------------------------
// key gather for the general table (fixed-width keys and string lengths).
// With up to GATHER_KEYS_AT_ONCE key columns every column's validity word
// and value (or string offsets) is loaded before the first store, so a
// group costs one round of random reads instead of one or two per column.
constexpr int GATHER_KEYS_AT_ONCE = 4;

__device__ __forceinline__ void gather_key_column(gb_params const& P, int c, int64_t g, int64_t G, bool live,
                                                  bool v, uint64_t x, uint64_t e, int32_t* str_sizes,
                                                  int64_t* str_begin, int& nstr)
{
  dcol const& kc = P.keys[c];
  if (kc.storage == ST_STRING) {
    if (live) {
      str_sizes[int64_t(nstr) * G + g] = v ? int32_t(e - x) : 0;
      str_begin[int64_t(nstr) * G + g] = int64_t(x);
    }
    nstr++;
  } else if (live) {
    store_width(P.out_keys[c], g, storage_width(kc.storage), x);
  }
  if (P.out_key_mask[c]) store_valid_bits(P.out_key_mask[c], g, G, live && v);
}

__device__ __forceinline__ uint64_t raw_key(dcol const& kc, int64_t r)
{
  auto const d = gp(kc.data);
  switch (storage_width(kc.storage)) {
    case 1: return d[r];
    case 2: return reinterpret_cast<gmem<uint16_t const>*>(d)[r];
    case 4: return reinterpret_cast<gmem<uint32_t const>*>(d)[r];
    default: return reinterpret_cast<gmem<uint64_t const>*>(d)[r];
  }
}

__global__ void __launch_bounds__(BLOCK)
  k_gather_keys(gb_params const* __restrict__ Pp, uint32_t const* __restrict__ rep_rows, int64_t G,
                int32_t* str_sizes, int64_t* str_begin)
{
  gb_params const& P = *Pp;
  int64_t const g    = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  bool const live    = g < G;
  int64_t const r    = live ? int64_t(rep_rows[g]) : 0;
  int nstr           = 0;
  if (P.num_keys <= GATHER_KEYS_AT_ONCE) {
    bool v[GATHER_KEYS_AT_ONCE];
    uint64_t x[GATHER_KEYS_AT_ONCE], e[GATHER_KEYS_AT_ONCE];  // value, or string begin / end
#pragma unroll
    for (int c = 0; c < GATHER_KEYS_AT_ONCE; c++) {
      v[c] = false;
      x[c] = e[c] = 0;
      if (c >= P.num_keys) continue;
      dcol const& kc = P.keys[c];
      v[c]           = live && dcol_valid(kc, r);
      if (!live) continue;
      if (kc.storage == ST_STRING) {
        int64_t b0, e0;
        str_span(kc, r, b0, e0);
        x[c] = uint64_t(b0);
        e[c] = uint64_t(e0);
      } else {
        x[c] = raw_key(kc, r);
      }
    }
#pragma unroll
    for (int c = 0; c < GATHER_KEYS_AT_ONCE; c++)
      if (c < P.num_keys) gather_key_column(P, c, g, G, live, v[c], x[c], e[c], str_sizes, str_begin, nstr);
    return;
  }
  for (int c = 0; c < P.num_keys; c++) {
    dcol const& kc = P.keys[c];
    bool const v   = live && dcol_valid(kc, r);
    uint64_t x = 0, e = 0;
    if (live) {
      if (kc.storage == ST_STRING) {
        int64_t b0, e0;
        str_span(kc, r, b0, e0);
        x = uint64_t(b0);
        e = uint64_t(e0);
      } else {
        x = raw_key(kc, r);
      }
    }
    gather_key_column(P, c, g, G, live, v, x, e, str_sizes, str_begin, nstr);
  }
}

void launch_gather_keys(gb_params const* dP, gb_params const& hP, uint32_t const* rep_rows,
                        int64_t G, int32_t* str_sizes, int64_t* str_begin, hipStream_t s)
{
  if (G <= 0) return;
  int64_t blocks = (G + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL(k_gather_keys, dim3(blocks), dim3(BLOCK), 0, s, dP, rep_rows, G, str_sizes, str_begin);
}

// Exclusive scan of int32 values (three phases: tile sums, scan of sums, apply).
constexpr int I32_TILE = 4096;
__global__ void __launch_bounds__(BLOCK) k_tile_sums_i32(int32_t const* data, int64_t n, uint32_t* sums)
{
  __shared__ uint32_t red[BLOCK / 64];
  int64_t const base = int64_t(blockIdx.x) * I32_TILE;
  uint32_t s = 0;
  for (int k = threadIdx.x; k < I32_TILE; k += BLOCK)
    if (base + k < n) s += uint32_t(data[base + k]);
  s = block_sum_u32(s, red);
  if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

__global__ void __launch_bounds__(BLOCK)
  k_apply_scan_i32(int32_t* data, int64_t n, uint32_t const* tile_offsets)
{
  __shared__ uint32_t wave_tot[BLOCK / 64];
  int64_t const base = int64_t(blockIdx.x) * I32_TILE;
  uint32_t running   = tile_offsets[blockIdx.x];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = 0; k < I32_TILE; k += BLOCK) {
    int64_t const i = base + k + threadIdx.x;
    uint32_t v      = i < n ? uint32_t(data[i]) : 0u;
    // inclusive scan within the wave
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wave_tot[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < BLOCK / 64; w++) {
      if (w < wave) before += wave_tot[w];
      all += wave_tot[w];
    }
    if (i < n) data[i] = int32_t(running + before + x - v);
    running += all;
    __syncthreads();
  }
}

void launch_exclusive_scan_i32(int32_t* data, int64_t n, int32_t* total, hipStream_t s)
{
  // `total` must point at data[n] (one extra element) or be a separate slot.
  int64_t const nt = (n + I32_TILE - 1) / I32_TILE;
  uint32_t* sums   = nullptr;
  MIG_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&sums), sizeof(uint32_t) * (nt + 1), s));
  hipLaunchKernelGGL(k_tile_sums_i32, dim3(nt), dim3(BLOCK), 0, s, data, n, sums);
  hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, sums, nt, reinterpret_cast<uint32_t*>(total));
  hipLaunchKernelGGL(k_apply_scan_i32, dim3(nt), dim3(BLOCK), 0, s, data, n, sums);
  MIG_HIP_TRY(hipFreeAsync(sums, s));
}

// Chars of each group's representative string: GATHER_LANES lanes per group
// (strings are short; a wave serves 64 / GATHER_LANES groups).
constexpr int GATHER_LANES = 4;

__global__ void __launch_bounds__(BLOCK)
  k_gather_chars(dcol src, uint32_t const* __restrict__ rep_rows, int64_t G,
                 int32_t const* __restrict__ out_offsets, uint8_t* out_chars)
{
  int64_t const t = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  int64_t const g = t / GATHER_LANES;
  int const lane  = int(t % GATHER_LANES);
  if (g >= G) return;
  int64_t const o   = out_offsets[g];
  int64_t const len = int64_t(out_offsets[g + 1]) - o;
  if (len == 0) return;
  int64_t b, e;
  str_span(src, int64_t(rep_rows[g]), b, e);
  for (int64_t i = lane; i < len; i += GATHER_LANES) gp(out_chars)[o + i] = gp(src.data)[b + i];
}

// The same copy from source offsets k_gather_keys recorded (sequential reads
// instead of a second random read of the string offsets).
__global__ void __launch_bounds__(BLOCK)
  k_gather_chars_from(uint8_t const* __restrict__ src_chars, int64_t const* __restrict__ src_begin, int64_t G,
                      int32_t const* __restrict__ out_offsets, uint8_t* out_chars)
{
  int64_t const t = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  int64_t const g = t / GATHER_LANES;
  int const lane  = int(t % GATHER_LANES);
  if (g >= G) return;
  int64_t const o   = out_offsets[g];
  int64_t const len = int64_t(out_offsets[g + 1]) - o;
  if (len == 0) return;
  int64_t const b = src_begin[g];
  for (int64_t i = lane; i < len; i += GATHER_LANES) gp(out_chars)[o + i] = gp(src_chars)[b + i];
}

void launch_gather_chars_from(uint8_t const* src_chars, int64_t const* src_begin, int64_t G,
                              int32_t const* out_offsets, uint8_t* out_chars, hipStream_t s)
{
  if (G <= 0) return;
  int64_t blocks = (G * GATHER_LANES + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL(k_gather_chars_from, dim3(blocks), dim3(BLOCK), 0, s, src_chars, src_begin, G, out_offsets,
                     out_chars);
}

void launch_gather_string_chars(dcol const& src, uint32_t const* rep_rows, int64_t G,
                                int32_t const* out_offsets, uint8_t* out_chars, hipStream_t s)
{
  if (G <= 0) return;
  int64_t blocks = (G * GATHER_LANES + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL(k_gather_chars, dim3(blocks), dim3(BLOCK), 0, s, src, rep_rows, G, out_offsets,
                     out_chars);
}

// Number of valid rows among [0, n) of a column view: popcount of the mask
// words covering bits [mask_offset, mask_offset + n), the partial first and
// last words masked.
__global__ void __launch_bounds__(BLOCK) k_count_valid(dcol c, int64_t n, uint32_t* out)
{
  __shared__ uint32_t red[BLOCK / 64];
  if (c.mask == nullptr) {  // no mask: every row is valid
    if (blockIdx.x == 0 && threadIdx.x == 0) aadd(out, uint32_t(n));
    return;
  }
  int64_t const b0 = c.mask_offset, b1 = int64_t(c.mask_offset) + n;
  int64_t const w0 = b0 >> 5, w1 = (b1 + 31) >> 5;
  uint32_t cnt = 0;
  for (int64_t w = w0 + int64_t(blockIdx.x) * BLOCK + threadIdx.x; w < w1; w += int64_t(gridDim.x) * BLOCK) {
    uint32_t bits = c.mask[w];
    if (w == w0) bits &= ~0u << (b0 & 31);
    if (w == w1 - 1 && (b1 & 31) != 0) bits &= (1u << (b1 & 31)) - 1u;
    cnt += __popc(bits);
  }
  cnt = block_sum_u32(cnt, red);
  if (threadIdx.x == 0 && cnt != 0) aadd(out, cnt);
}

// Set bits among the first n bits of each of up to MASK_LIST masks (word
// aligned, as result masks are): blockIdx.y picks the mask.
constexpr int MASK_LIST = 32;
struct mask_list {
  uint32_t const* mask[MASK_LIST];
};

__global__ void __launch_bounds__(BLOCK) k_count_valid_list(mask_list L, int64_t n, uint32_t* out)
{
  __shared__ uint32_t red[BLOCK / 64];
  auto const mask    = gp(L.mask[blockIdx.y]);
  int64_t const w1   = (n + 31) >> 5;
  uint32_t cnt       = 0;
  for (int64_t w = int64_t(blockIdx.x) * BLOCK + threadIdx.x; w < w1; w += int64_t(gridDim.x) * BLOCK) {
    uint32_t bits = mask[w];
    if (w == w1 - 1 && (n & 31) != 0) bits &= (1u << (n & 31)) - 1u;
    cnt += __popc(bits);
  }
  cnt = block_sum_u32(cnt, red);
  if (threadIdx.x == 0 && cnt != 0) aadd(out + blockIdx.y, cnt);
}

void launch_count_valid_list(uint32_t const* const* masks, int count, int64_t n, uint32_t* d_counts, hipStream_t s)
{
  if (n <= 0) return;
  int64_t blocks = ((n + 31) / 32 + BLOCK - 1) / BLOCK;
  if (blocks > 256) blocks = 256;
  for (int i0 = 0; i0 < count; i0 += MASK_LIST) {
    int const k = count - i0 < MASK_LIST ? count - i0 : MASK_LIST;
    mask_list L{};
    for (int i = 0; i < k; i++) L.mask[i] = masks[i0 + i];
    hipLaunchKernelGGL(k_count_valid_list, dim3(blocks, k), dim3(BLOCK), 0, s, L, n, d_counts + i0);
  }
}

void launch_count_nulls(dcol const& c, int64_t n, uint32_t* d_count, hipStream_t s)
{
  if (n <= 0) return;
  int64_t const words = (int64_t(c.mask_offset & 31) + n + 31) / 32;
  int64_t blocks      = c.mask == nullptr ? 1 : (words + BLOCK - 1) / BLOCK;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_count_valid, dim3(blocks), dim3(BLOCK), 0, s, c, n, d_count);
}

// ---------------------------------------------------------------------------
// Pre-sorted keys (groupby(keys, null_policy, sorted::YES)): the groups are
// the runs of equal adjacent key rows, numbered in input order, as the
// reference's sort path finds them with keys_pre_sorted (sort/sort_helper.cu:
// 55-72, 133-181).  k_sorted_count counts run heads per tile, a scan gives
// each tile its first group number, k_sorted_assign writes every row's group
// (NO_GROUP for rows the null policy drops) and each group's first row, and
// the aggregation kernels reduce each 64-row chunk with a segmented scan
// across the wave, so a run costs one atomic per wave and field.
constexpr int SORTED_ROWS = 16;                    // rows per thread per tile
constexpr int SORTED_TILE = BLOCK * SORTED_ROWS;   // rows per tile
constexpr uint32_t NO_GROUP = 0xFFFFFFFFu;

__device__ __forceinline__ bool sorted_included(gb_params const& P, int64_t row)
{
  return !(P.skip_null_keys && any_key_null(P, row));
}

__device__ __forceinline__ bool sorted_head(gb_params const& P, int64_t row)
{
  if (!sorted_included(P, row)) return false;
  if (row == 0 || !sorted_included(P, row - 1)) return true;
  return !rows_equal(P, row, row - 1);
}

// Inclusion and run-head flags of a row (sorted_included, sorted_head).  With
// one fixed-width key column, the common case, the row's and its
// predecessor's validity and key bits are read once each.
__device__ __forceinline__ bool sorted_one_key(gb_params const& P)
{
  return P.num_keys == 1 && P.keys[0].storage != ST_STRING;
}

__device__ __forceinline__ void sorted_flags(gb_params const& P, bool one_key, int64_t row, bool& incl, bool& head)
{
  if (!one_key) {
    incl = sorted_included(P, row);
    head = incl && sorted_head(P, row);
    return;
  }
  dcol const& k = P.keys[0];
  bool const v  = dcol_valid(k, row);
  incl          = v || !P.skip_null_keys;
  head          = incl;
  if (incl && row > 0) {
    bool const pv = dcol_valid(k, row - 1);
    if (pv || !P.skip_null_keys)  // the predecessor is kept: a head when the keys differ (nulls equal)
      head = v != pv || (v && load_fixed_bits(k, row) != load_fixed_bits(k, row - 1));
  }
}

__global__ void __launch_bounds__(BLOCK) k_sorted_count(gb_params const* __restrict__ Pp, uint32_t* block_counts)
{
  __shared__ uint32_t red[BLOCK / 64];
  gb_params const& P = *Pp;
  bool const one_key = sorted_one_key(P);
  int64_t const base = int64_t(blockIdx.x) * SORTED_TILE;
  uint32_t heads = 0;
  for (int j = 0; j < SORTED_ROWS; j++) {
    int64_t const row = base + int64_t(j) * BLOCK + threadIdx.x;
    if (row < P.num_rows) {
      bool incl, head;
      sorted_flags(P, one_key, row, incl, head);
      heads += head ? 1u : 0u;
    }
  }
  uint32_t const t = block_sum_u32<BLOCK>(heads, red);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = t;
}

__global__ void __launch_bounds__(BLOCK)
  k_sorted_assign(gb_params const* __restrict__ Pp, uint32_t const* __restrict__ block_offsets,
                  uint32_t* __restrict__ gid, uint32_t* __restrict__ group_last)
{
  constexpr int NW = BLOCK / 64;
  __shared__ uint32_t wave_tot[NW];
  __shared__ uint64_t head_bal[NW], incl_bal[NW];
  gb_params const& P = *Pp;
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t const below = (1ull << lane) - 1ull;
  int64_t const base   = int64_t(blockIdx.x) * SORTED_TILE;
  uint32_t running     = block_offsets[blockIdx.x];  // heads before this iteration's rows
  // the block's last thread: its row's group when that row is included and
  // whether the group ends there is decided by the next iteration's first row
  uint32_t pending_g = NO_GROUP;
  uint32_t pending_row = 0;
  bool const one_key   = sorted_one_key(P);
  for (int j = 0; j < SORTED_ROWS; j++) {
    int64_t const row = base + int64_t(j) * BLOCK + threadIdx.x;
    bool const live   = row < P.num_rows;
    bool incl = false, head = false;
    if (live) sorted_flags(P, one_key, row, incl, head);
    uint64_t const bal  = __ballot(head);
    uint64_t const ibal = __ballot(incl);
    if (lane == 0) {
      wave_tot[wave] = uint32_t(__popcll(bal));
      head_bal[wave] = bal;
      incl_bal[wave] = ibal;
    }
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < NW; w++) {
      if (w < wave) before += wave_tot[w];
      all += wave_tot[w];
    }
    uint32_t const heads_before = running + before + uint32_t(__popcll(bal & below));
    uint32_t const g            = !incl ? NO_GROUP : (head ? heads_before : heads_before - 1u);
    // A group's rows are one run with no dropped row inside (a dropped row
    // ends a run), so its last row is the included row whose successor is
    // dropped, a head or past the end.  The successor's flags come from the
    // wave's ballots, the next wave's (LDS), or for the block's last row the
    // next iteration's first row.
    if (group_last != nullptr) {
      if (pending_g != NO_GROUP && ((incl_bal[0] & 1u) == 0 || (head_bal[0] & 1u) != 0))
        group_last[pending_g] = pending_row;  // the previous iteration's last row (block's last thread only)
      pending_g = NO_GROUP;
      if (incl) {
        bool ends = false;
        if (lane < 63) ends = ((ibal >> (lane + 1)) & 1u) == 0 || ((bal >> (lane + 1)) & 1u) != 0;
        else if (wave < NW - 1) ends = (incl_bal[wave + 1] & 1u) == 0 || (head_bal[wave + 1] & 1u) != 0;
        else if (j + 1 < SORTED_ROWS && row + 1 < P.num_rows) {
          pending_g   = g;
          pending_row = uint32_t(row);
        } else {  // the tile's last row: its successor starts the next tile
          int64_t const nr = row + 1;
          bool ni = false, nh = false;
          if (nr < P.num_rows) sorted_flags(P, one_key, nr, ni, nh);
          ends = !ni || nh;
        }
        if (ends) group_last[g] = uint32_t(row);
      }
    }
    if (live) {
      gid[row] = g;
      if (head) P.slot_rows[g] = uint64_t(row);
    }
    running += all;
    __syncthreads();
  }
}

// Value of one row in the accumulator encoding of field kind `kind`.
__device__ __forceinline__ uint64_t sorted_part(int32_t kind, loaded_value const& v)
{
  switch (kind) {
    case FK_SUM_I64: return uint64_t(v.i);
    case FK_SUM_F32: return __float_as_uint(v.f);
    case FK_SUM_F64: return uint64_t(__double_as_longlong(v.d));
    case FK_SUMSQ_I64: return uint64_t(v.i) * uint64_t(v.i);
    case FK_SUMSQ_F32: return __float_as_uint(v.f * v.f);
    case FK_SUMSQ_F64: return uint64_t(__double_as_longlong(v.d * v.d));
    case FK_MIN_S32:
    case FK_MAX_S32: return uint32_t(int32_t(v.i));
    case FK_MIN_S64:
    case FK_MAX_S64:
    case FK_MIN_U64:
    case FK_MAX_U64: return uint64_t(v.i);
    case FK_MIN_U32:
    case FK_MAX_U32: return uint32_t(v.i);
    case FK_MIN_F32:
    case FK_MAX_F32: return isnan(v.f) ? field_identity(kind) : ord32(uint32_t(v.b));
    case FK_MIN_F64:
    case FK_MAX_F64: return isnan(v.d) ? field_identity(kind) : ord64(v.b);
    case FK_COUNT: return 1;
    case FK_PROD_I64: return uint64_t(v.i);
    case FK_PROD_F32: return __float_as_uint(v.f);
    case FK_PROD_F64: return uint64_t(__double_as_longlong(v.d));
    default: return 0;
  }
}

// Combine two partials in the accumulator encoding.
__device__ __forceinline__ uint64_t sorted_combine(int32_t kind, uint64_t a, uint64_t b)
{
  switch (kind) {
    case FK_SUM_I64:
    case FK_SUMSQ_I64:
    case FK_COUNT: return a + b;
    case FK_SUM_F32:
    case FK_SUMSQ_F32: return __float_as_uint(__uint_as_float(uint32_t(a)) + __uint_as_float(uint32_t(b)));
    case FK_SUM_F64:
    case FK_SUMSQ_F64:
    case FK_M2:
      return uint64_t(__double_as_longlong(__longlong_as_double(int64_t(a)) + __longlong_as_double(int64_t(b))));
    case FK_MIN_S32: return uint32_t(min(int32_t(uint32_t(a)), int32_t(uint32_t(b))));
    case FK_MAX_S32: return uint32_t(max(int32_t(uint32_t(a)), int32_t(uint32_t(b))));
    case FK_MIN_S64: return uint64_t(min(int64_t(a), int64_t(b)));
    case FK_MAX_S64: return uint64_t(max(int64_t(a), int64_t(b)));
    case FK_MIN_U32:
    case FK_MIN_F32: return min(uint32_t(a), uint32_t(b));
    case FK_MAX_U32:
    case FK_MAX_F32: return max(uint32_t(a), uint32_t(b));
    case FK_MIN_U64:
    case FK_MIN_F64: return min(a, b);
    case FK_MAX_U64:
    case FK_MAX_F64: return max(a, b);
    case FK_PROD_I64: return a * b;
    case FK_PROD_F32: return __float_as_uint(__uint_as_float(uint32_t(a)) * __uint_as_float(uint32_t(b)));
    case FK_PROD_F64:
      return uint64_t(__double_as_longlong(__longlong_as_double(int64_t(a)) * __longlong_as_double(int64_t(b))));
    default: return a;
  }
}

// Combining operations of the accumulator kinds (sums and counts add in their
// type; min / max compare signed or unsigned: float min / max accumulators
// hold order-preserving unsigned encodings).
enum seg_op : int32_t {
  OP_ADD_I64 = 0, OP_ADD_F32, OP_ADD_F64, OP_MIN_S32, OP_MAX_S32, OP_MIN_S64, OP_MAX_S64,
  OP_MIN_U32, OP_MAX_U32, OP_MIN_U64, OP_MAX_U64, OP_MUL_I64, OP_MUL_F32, OP_MUL_F64
};

__host__ __device__ inline int32_t seg_op_of(int32_t kind)
{
  switch (kind) {
    case FK_SUM_F32:
    case FK_SUMSQ_F32: return OP_ADD_F32;
    case FK_SUM_F64:
    case FK_SUMSQ_F64:
    case FK_M2: return OP_ADD_F64;
    case FK_MIN_S32: return OP_MIN_S32;
    case FK_MAX_S32: return OP_MAX_S32;
    case FK_MIN_S64: return OP_MIN_S64;
    case FK_MAX_S64: return OP_MAX_S64;
    case FK_MIN_U32:
    case FK_MIN_F32: return OP_MIN_U32;
    case FK_MAX_U32:
    case FK_MAX_F32: return OP_MAX_U32;
    case FK_MIN_U64:
    case FK_MIN_F64: return OP_MIN_U64;
    case FK_MAX_U64:
    case FK_MAX_F64: return OP_MAX_U64;
    case FK_PROD_I64: return OP_MUL_I64;
    case FK_PROD_F32: return OP_MUL_F32;
    case FK_PROD_F64: return OP_MUL_F64;
    default: return OP_ADD_I64;  // integer sums, sums of squares, counts
  }
}

template <int OP>
__device__ __forceinline__ uint64_t op_combine(uint64_t a, uint64_t b)
{
  if constexpr (OP == OP_ADD_I64) return a + b;
  if constexpr (OP == OP_ADD_F32) return __float_as_uint(__uint_as_float(uint32_t(a)) + __uint_as_float(uint32_t(b)));
  if constexpr (OP == OP_ADD_F64)
    return uint64_t(__double_as_longlong(__longlong_as_double(int64_t(a)) + __longlong_as_double(int64_t(b))));
  if constexpr (OP == OP_MIN_S32) return uint32_t(min(int32_t(uint32_t(a)), int32_t(uint32_t(b))));
  if constexpr (OP == OP_MAX_S32) return uint32_t(max(int32_t(uint32_t(a)), int32_t(uint32_t(b))));
  if constexpr (OP == OP_MIN_S64) return uint64_t(min(int64_t(a), int64_t(b)));
  if constexpr (OP == OP_MAX_S64) return uint64_t(max(int64_t(a), int64_t(b)));
  if constexpr (OP == OP_MIN_U32) return min(uint32_t(a), uint32_t(b));
  if constexpr (OP == OP_MAX_U32) return max(uint32_t(a), uint32_t(b));
  if constexpr (OP == OP_MIN_U64) return min(a, b);
  if constexpr (OP == OP_MAX_U64) return max(a, b);
  if constexpr (OP == OP_MUL_I64) return a * b;
  if constexpr (OP == OP_MUL_F32) return __float_as_uint(__uint_as_float(uint32_t(a)) * __uint_as_float(uint32_t(b)));
  if constexpr (OP == OP_MUL_F64)
    return uint64_t(__double_as_longlong(__longlong_as_double(int64_t(a)) * __longlong_as_double(int64_t(b))));
  return a;
}

// One step of the segmented scan: lanes take the partial of the lane the DPP
// control names (row shift within 16-lane rows, or a broadcast of a row's last
// lane to the next rows) and combine it when that lane holds the same group.
// Lanes without a source keep `old`, whose group value matches no group.
template <int OP, int CTRL, int ROW_MASK>
__device__ __forceinline__ void seg_step(uint64_t& part, uint32_t g)
{
  constexpr int NOT_A_GROUP = int(0xFFFFFFFEu);
  uint32_t const olo = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(part)), CTRL, ROW_MASK, 0xF, false));
  uint32_t const ohi = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(part >> 32)), CTRL, ROW_MASK, 0xF, false));
  uint32_t const og  = uint32_t(__builtin_amdgcn_update_dpp(NOT_A_GROUP, int(g), CTRL, ROW_MASK, 0xF, false));
  if (og == g) part = op_combine<OP>(part, (uint64_t(ohi) << 32) | olo);
}

// Inclusive scan of `part` over the lanes of each run of equal `g` (runs
// are contiguous lanes because group numbers never decrease along rows):
// row shifts by 1, 2, 4, 8 inside each 16-lane row, then row 0's and row 2's
// last lanes into rows 1 and 3, then lane 31 into rows 2 and 3.  A lane only
// combines a partial of its own group, and groups are contiguous, so every
// lane between the source and the lane belongs to the same group.
template <int OP>
__device__ __forceinline__ uint64_t seg_scan(uint64_t part, uint32_t g)
{
  seg_step<OP, 0x111, 0xF>(part, g);  // row_shr:1
  seg_step<OP, 0x112, 0xF>(part, g);  // row_shr:2
  seg_step<OP, 0x114, 0xF>(part, g);  // row_shr:4
  seg_step<OP, 0x118, 0xF>(part, g);  // row_shr:8
  seg_step<OP, 0x142, 0xA>(part, g);  // row_bcast:15
  seg_step<OP, 0x143, 0xC>(part, g);  // row_bcast:31
  return part;
}

// One dispatch on the combining operation per scan.
__device__ __forceinline__ uint64_t segmented_scan(int32_t kind, uint64_t part, uint32_t g, int)
{
  switch (seg_op_of(kind)) {
    case OP_ADD_F32: return seg_scan<OP_ADD_F32>(part, g);
    case OP_ADD_F64: return seg_scan<OP_ADD_F64>(part, g);
    case OP_MIN_S32: return seg_scan<OP_MIN_S32>(part, g);
    case OP_MAX_S32: return seg_scan<OP_MAX_S32>(part, g);
    case OP_MIN_S64: return seg_scan<OP_MIN_S64>(part, g);
    case OP_MAX_S64: return seg_scan<OP_MAX_S64>(part, g);
    case OP_MIN_U32: return seg_scan<OP_MIN_U32>(part, g);
    case OP_MAX_U32: return seg_scan<OP_MAX_U32>(part, g);
    case OP_MIN_U64: return seg_scan<OP_MIN_U64>(part, g);
    case OP_MAX_U64: return seg_scan<OP_MAX_U64>(part, g);
    case OP_MUL_I64: return seg_scan<OP_MUL_I64>(part, g);
    case OP_MUL_F32: return seg_scan<OP_MUL_F32>(part, g);
    case OP_MUL_F64: return seg_scan<OP_MUL_F64>(part, g);
    default: return seg_scan<OP_ADD_I64>(part, g);
  }
}

// Ascending bitonic sort of one 32-bit key per lane across the wave.
__device__ __forceinline__ uint32_t wave_sort_u32(uint32_t key)
{
  int const lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint32_t const other = uint32_t(__shfl_xor(int(key), j));
      bool const keep_min  = ((lane & j) == 0) == ((lane & k) == 0);
      key                  = keep_min ? min(key, other) : max(key, other);
    }
  }
  return key;
}

// The device-wide table with equal groups of a wave combined before the
// atomics: each wave sorts its rows' (slot, lane) pairs, so rows of one group
// sit in adjacent lanes; every accumulator's row contributions are reduced
// over each run by the segmented scan, and the run's last lane applies the
// total with one atomic.  Hot keys then cost one atomic per wave and field
// instead of one per row.
template <int KM>
__global__ void __launch_bounds__(BLOCK) k_aggregate_global_combined(gb_params const* __restrict__ Pp)
{
  constexpr uint32_t NONE = (1u << 26) - 1u;  // sort key of a lane without a group
  gb_params const& P = *Pp;
  int64_t const n    = P.num_rows;
  int const lane     = threadIdx.x & 63;
  for (int64_t base = int64_t(blockIdx.x) * BLOCK; base < n; base += int64_t(gridDim.x) * BLOCK) {
    // the call rebuilds a larger table: stop early (one load of the flag per
    // wave, so the whole wave leaves together)
    if (__builtin_amdgcn_readfirstlane(int(aload(P.overflow))) != 0) break;
    int64_t const row  = base + threadIdx.x;
    int64_t const slot = row < n ? locate<KM, true>(P, row) : -1;
    uint32_t const key = wave_sort_u32(((slot < 0 ? NONE : uint32_t(slot)) << 6) | uint32_t(lane));
    uint32_t const g   = key >> 6;
    int64_t const srow = base + (threadIdx.x & ~63) + int64_t(key & 63u);  // the row this lane now holds
    bool const has     = g != NONE;
    uint32_t const gn  = uint32_t(__shfl_down(int(g), 1));
    bool const tail    = has && (lane == 63 || gn != g);
    if (P.need_count_all) {
      uint64_t const part = seg_scan<OP_ADD_I64>(has ? 1u : 0u, g);
      if (tail) aadd(count_at(P, g), uint32_t(part));
    }
    for (int vc = 0; vc < P.num_vals; vc++) {
      int const fb = P.val_field_begin[vc], fe = P.val_field_begin[vc + 1];
      if (fb == fe) continue;
      dcol const& c    = P.vals[vc];
      bool const valid = has && dcol_valid(c, srow);
      loaded_value v{0, 0.0, 0.0f, 0};
      if (valid) v = load_value(c, srow);
      for (int f = fb; f < fe; f++) {
        field_desc const& F = P.fields[f];
        if (F.kind == FK_M2) continue;  // the second pass
        uint64_t const ident = field_identity(F.kind);
        uint64_t const part  = segmented_scan(F.kind, valid ? sorted_part(F.kind, v) : ident, g, 0);
        if (tail && part != ident) acc_combine(F.kind, field_at(F, g), part);
      }
    }
  }
}

// Accumulators one pass over the rows reduces: up to SORTED_FPP fields
// (index -1: the group row count) reading up to SORTED_COLS value columns.
constexpr int SORTED_FPP  = 8;
constexpr int SORTED_COLS = 4;
struct sorted_pass {
  int32_t n;
  int32_t ncols;
  int32_t field[SORTED_FPP];
  int32_t col_of[SORTED_FPP];   // loaded column slot of each field (-1: the row count)
  int32_t kind[SORTED_FPP];     // accumulator kind (FK_COUNT for the row count)
  int32_t op[SORTED_FPP];       // combining operation (seg_op)
  uint64_t ident[SORTED_FPP];   // identity, in the scan domain (see scan_domain)
  int32_t xf[SORTED_FPP];       // accumulator encoding -> scan domain (scan_domain)
  int32_t col[SORTED_COLS];     // value column (index into P.vals) of each slot
};

// Element bits of a fixed-width column as stored (no float normalisation).
// A string column (width 0) is in a pass only for its validity (COUNT_VALID):
// its chars buffer is not indexed by row, so nothing is loaded.
__device__ __forceinline__ uint64_t raw_bits(dcol const& c, int64_t row)
{
  auto const d = gp(c.data);
  switch (storage_width(c.storage)) {
    case 1: return d[row];
    case 2: return reinterpret_cast<gmem<uint16_t const>*>(d)[row];
    case 4: return reinterpret_cast<gmem<uint32_t const>*>(d)[row];
    case 8: return reinterpret_cast<gmem<uint64_t const>*>(d)[row];
    default: return 0;
  }
}

// The loaded_value of element bits (as load_value builds it from memory).
__device__ __forceinline__ loaded_value value_of_bits(int32_t st, uint64_t u)
{
  loaded_value v{0, 0.0, 0.0f, 0};
  switch (st) {
    case ST_I8: v.i = int8_t(u); break;
    case ST_I16: v.i = int16_t(u); break;
    case ST_I32: v.i = int32_t(u); break;
    case ST_I64: v.i = int64_t(u); break;
    case ST_BOOL: v.i = (u & 0xff) != 0; break;
    case ST_U8: v.i = uint8_t(u); break;
    case ST_U16: v.i = uint16_t(u); break;
    case ST_U32: v.i = uint32_t(u); break;
    case ST_U64: v.i = int64_t(u); v.d = double(u); return v;
    case ST_F32: v.b = uint32_t(u); v.f = __uint_as_float(uint32_t(u)); v.d = double(v.f); return v;
    case ST_F64: v.b = u; v.d = __longlong_as_double(int64_t(u)); return v;
    default: return v;
  }
  v.d = double(v.i);
  return v;
}

__device__ __forceinline__ void sorted_item(gb_params const& P, int32_t f, int64_t g, int32_t& kind, int32_t& width,
                                            void*& ptr)
{
  if (f < 0) {
    kind  = FK_COUNT;
    width = 4;
    ptr   = count_at(P, g);
  } else {
    field_desc const& F = P.fields[f];
    kind  = F.kind;
    width = F.width;
    ptr   = field_at(F, g);
  }
}

// A finished group total: stored when no other wave holds rows of the group,
// combined atomically otherwise (the accumulators start at the identity).
__device__ __forceinline__ void sorted_flush(gb_params const& P, int32_t f, uint32_t g, uint64_t part, bool shared)
{
  int32_t kind, width;
  void* ptr;
  sorted_item(P, f, g, kind, width, ptr);
  if (shared) {
    if (kind == FK_M2) gadd_f64(static_cast<double*>(ptr), __longlong_as_double(int64_t(part)));
    else acc_combine(kind, ptr, part);
  } else if (width == 8) {
    *gp(static_cast<uint64_t*>(ptr)) = part;
  } else {
    *gp(static_cast<uint32_t*>(ptr)) = uint32_t(part);
  }
}

// Row contribution of accumulator `f` (kind `kind`, identity `ident`) from
// the row's loaded value (the value's square deviation from the group mean
// for FK_M2).
__device__ __forceinline__ uint64_t sorted_row_part(gb_params const& P, int32_t f, int32_t kind, uint64_t ident,
                                                    uint32_t g, bool valid, loaded_value const& v)
{
  if (f < 0) return g != NO_GROUP ? 1u : 0u;
  if (g == NO_GROUP || !valid) return ident;
  if (kind != FK_M2) return sorted_part(kind, v);
  field_desc const& F = P.fields[f];
  double const cnt    = double(count_of(P, F.aux_count, g));
  double const mean   = field_as_double(P.fields[F.aux_sum], g) / cnt;
  double const d      = v.d - mean;
  return uint64_t(__double_as_longlong(d * d));
}

// The runs reduction works in a "scan domain" with few combining operations:
// every MIN / MAX accumulator becomes a signed 64-bit MIN / MAX (32-bit
// signed values sign-extended, 32-bit unsigned values and float order codes
// zero-extended, 64-bit unsigned values and float order codes with the sign
// bit flipped, which maps unsigned order onto signed order).  Sums, counts
// and products keep their operation.  Fewer operations, fewer instantiated
// scans: the kernel's hot loop stays small for the instruction cache.
enum scan_xf : int32_t { XF_NONE = 0, XF_SEXT32 = 1, XF_ZEXT32 = 2, XF_FLIP64 = 3 };

__host__ __device__ inline void scan_domain(int32_t kind, int32_t& op, int32_t& xf)
{
  xf = XF_NONE;
  switch (kind) {
    case FK_MIN_S32: op = OP_MIN_S64; xf = XF_SEXT32; return;
    case FK_MAX_S32: op = OP_MAX_S64; xf = XF_SEXT32; return;
    case FK_MIN_U32:
    case FK_MIN_F32: op = OP_MIN_S64; xf = XF_ZEXT32; return;
    case FK_MAX_U32:
    case FK_MAX_F32: op = OP_MAX_S64; xf = XF_ZEXT32; return;
    case FK_MIN_U64:
    case FK_MIN_F64: op = OP_MIN_S64; xf = XF_FLIP64; return;
    case FK_MAX_U64:
    case FK_MAX_F64: op = OP_MAX_S64; xf = XF_FLIP64; return;
    default: op = seg_op_of(kind); return;
  }
}

__host__ __device__ inline uint64_t xf_fwd(int32_t xf, uint64_t u)
{
  switch (xf) {
    case XF_SEXT32: return uint64_t(int64_t(int32_t(uint32_t(u))));
    case XF_ZEXT32: return uint64_t(uint32_t(u));
    case XF_FLIP64: return u ^ 0x8000000000000000ull;
    default: return u;
  }
}

__host__ __device__ inline uint64_t xf_inv(int32_t xf, uint64_t t)
{
  switch (xf) {
    case XF_SEXT32:
    case XF_ZEXT32: return uint64_t(uint32_t(t));
    case XF_FLIP64: return t ^ 0x8000000000000000ull;
    default: return t;
  }
}

// Combine two partials of the scan domain.
__device__ __forceinline__ uint64_t domain_combine(int32_t op, uint64_t a, uint64_t b)
{
  switch (op) {
    case OP_ADD_F32: return op_combine<OP_ADD_F32>(a, b);
    case OP_ADD_F64: return op_combine<OP_ADD_F64>(a, b);
    case OP_MIN_S64: return op_combine<OP_MIN_S64>(a, b);
    case OP_MAX_S64: return op_combine<OP_MAX_S64>(a, b);
    case OP_MUL_I64: return op_combine<OP_MUL_I64>(a, b);
    case OP_MUL_F32: return op_combine<OP_MUL_F32>(a, b);
    case OP_MUL_F64: return op_combine<OP_MUL_F64>(a, b);
    default: return op_combine<OP_ADD_I64>(a, b);
  }
}

// Segmented scan of the scan domain, the operation chosen once per call.
__device__ __forceinline__ uint64_t segmented_scan_op(int32_t op, uint64_t part, uint32_t g)
{
  switch (op) {
    case OP_ADD_F32: return seg_scan<OP_ADD_F32>(part, g);
    case OP_ADD_F64: return seg_scan<OP_ADD_F64>(part, g);
    case OP_MIN_S64: return seg_scan<OP_MIN_S64>(part, g);
    case OP_MAX_S64: return seg_scan<OP_MAX_S64>(part, g);
    case OP_MUL_I64: return seg_scan<OP_MUL_I64>(part, g);
    case OP_MUL_F32: return seg_scan<OP_MUL_F32>(part, g);
    case OP_MUL_F64: return seg_scan<OP_MUL_F64>(part, g);
    default: return seg_scan<OP_ADD_I64>(part, g);
  }
}

// ARGMIN / ARGMAX on runs: a row's partial is (order key of its value, row),
// reduced lexicographically (smallest / largest key, then the smallest row),
// which is the device-wide path's rule (first row among equal values, NaN
// only in an all-NaN group; -0.0 equals 0.0).  Rows without a valid value
// carry the identity key and row 0xFFFFFFFF.
template <bool IS_MIN>
__device__ __forceinline__ void arg_pick(uint64_t& key, uint32_t& row, uint64_t okey, uint32_t orow)
{
  bool const take = IS_MIN ? (okey < key || (okey == key && orow < row)) : (okey > key || (okey == key && orow < row));
  if (take) {
    key = okey;
    row = orow;
  }
}

__device__ __forceinline__ uint64_t arg_order_key(bool is_min, int32_t st, loaded_value const& v)
{
  if (st == ST_F32 || st == ST_F64) {
    if (isnan(v.d)) return is_min ? ~0ull : 0ull;  // a NaN never beats a number
    double const d = v.d == 0.0 ? 0.0 : v.d;        // -0.0 equals 0.0
    return ord64(uint64_t(__double_as_longlong(d)));
  }
  if (is_unsigned_storage_dev(st)) return uint64_t(v.i);
  return uint64_t(v.i) ^ 0x8000000000000000ull;
}

template <bool IS_MIN, int CTRL, int ROW_MASK>
__device__ __forceinline__ void arg_step(uint64_t& key, uint32_t& row, uint32_t g)
{
  constexpr int NOT_A_GROUP = int(0xFFFFFFFEu);
  constexpr int ID_WORD     = IS_MIN ? -1 : 0;
  uint32_t const klo = uint32_t(__builtin_amdgcn_update_dpp(ID_WORD, int(uint32_t(key)), CTRL, ROW_MASK, 0xF, false));
  uint32_t const khi = uint32_t(__builtin_amdgcn_update_dpp(ID_WORD, int(uint32_t(key >> 32)), CTRL, ROW_MASK, 0xF, false));
  uint32_t const orow = uint32_t(__builtin_amdgcn_update_dpp(-1, int(row), CTRL, ROW_MASK, 0xF, false));
  uint32_t const og   = uint32_t(__builtin_amdgcn_update_dpp(NOT_A_GROUP, int(g), CTRL, ROW_MASK, 0xF, false));
  if (og == g) arg_pick<IS_MIN>(key, row, (uint64_t(khi) << 32) | klo, orow);
}

template <bool IS_MIN>
__device__ __forceinline__ void arg_scan(uint64_t& key, uint32_t& row, uint32_t g)
{
  arg_step<IS_MIN, 0x111, 0xF>(key, row, g);  // row_shr:1
  arg_step<IS_MIN, 0x112, 0xF>(key, row, g);  // row_shr:2
  arg_step<IS_MIN, 0x114, 0xF>(key, row, g);  // row_shr:4
  arg_step<IS_MIN, 0x118, 0xF>(key, row, g);  // row_shr:8
  arg_step<IS_MIN, 0x142, 0xA>(key, row, g);  // row_bcast:15
  arg_step<IS_MIN, 0x143, 0xC>(key, row, g);  // row_bcast:31
}

// A finished group's row: stored when no other wave holds rows of the group,
// otherwise merged by the device-wide path's compare-and-swap on the values.
__device__ __forceinline__ void arg_flush(gb_params const& P, int32_t f, uint32_t g, uint32_t row, bool shared)
{
  if (row == 0xFFFFFFFFu) return;  // no valid value: the accumulator keeps "none"
  field_desc const& F = P.fields[f];
  if (shared) arg_update(F, g, P.vals[F.vcol], int64_t(row));
  else *gp(reinterpret_cast<uint32_t*>(field_at(F, g))) = row;
}

// Each wave reduces a contiguous range of rows in 64-row chunks; the group
// still open at a chunk's last lane is carried into the next chunk.  Groups
// that start and end inside the range are stored once; the range's first and
// last groups may have rows in the neighbouring ranges and are combined
// atomically.  HAS_ARG: the pass holds ARGMIN / ARGMAX accumulators (their
// branch is compiled only into that variant, which keeps the common one lean).
template <bool HAS_ARG>
__global__ void __launch_bounds__(BLOCK)
  k_sorted_reduce(gb_params const* __restrict__ Pp, uint32_t const* __restrict__ gid, int64_t per, int64_t W,
                  sorted_pass pass)
{
  // the partials carried between chunks live in LDS (one row per wave): the
  // field loop stays rolled, which keeps the kernel small enough for the
  // instruction cache
  __shared__ uint64_t carry[BLOCK / 64][SORTED_FPP];
  __shared__ uint32_t carry_row[HAS_ARG ? BLOCK / 64 : 1][SORTED_FPP];  // ARGMIN / ARGMAX partials
  gb_params const& P = *Pp;
  int const lane     = threadIdx.x & 63;
  int const wave     = threadIdx.x >> 6;
  int64_t const w    = int64_t(blockIdx.x) * (BLOCK / 64) + wave;
  if (w >= W) return;
  int64_t const b = w * per;
  int64_t const e = b + per < P.num_rows ? b + per : P.num_rows;
  uint32_t const first_g = gid[b];  // may continue from the previous range
  uint32_t carry_g       = NO_GROUP;
  for (int64_t i0 = b; i0 < e; i0 += 64) {
    int64_t const row = i0 + lane;
    bool const live   = row < e;
    int64_t const rc  = live ? row : e - 1;  // every load of the chunk issues before any use
    uint32_t const gl = gid[rc];
    uint64_t bits[SORTED_COLS];
    bool vld[SORTED_COLS];
#pragma unroll
    for (int c = 0; c < SORTED_COLS; c++) {
      bits[c] = 0;
      vld[c]  = false;
      if (c < pass.ncols) {
        dcol const& d = P.vals[pass.col[c]];
        bits[c]       = raw_bits(d, rc);
        vld[c]        = dcol_valid(d, rc);
      }
    }
    // each column's value decoded once per chunk
    loaded_value lv[SORTED_COLS];
#pragma unroll
    for (int c = 0; c < SORTED_COLS; c++)
      lv[c] = c < pass.ncols ? value_of_bits(P.vals[pass.col[c]].storage, bits[c]) : loaded_value{0, 0.0, 0.0f, 0};
    uint32_t const g  = live ? gl : NO_GROUP;
    // group of the next lane (wave_shl:1; the last lane gets no group)
    uint32_t const gn = uint32_t(__builtin_amdgcn_update_dpp(int(NO_GROUP), int(g), 0x130, 0xF, 0xF, false));
    uint32_t const g0 = uint32_t(__builtin_amdgcn_readlane(int(g), 0));
    bool const carry_ends = carry_g != NO_GROUP && g0 != carry_g;
    bool const ends_here  = g != NO_GROUP && lane < 63 && gn != g;
#pragma unroll 1
    for (int k = 0; k < pass.n; k++) {
      int32_t const f    = pass.field[k];
      int32_t const cs   = pass.col_of[k];
      int32_t const kind = pass.kind[k];
      uint64_t const c   = carry[wave][k];
      // the field's column among the loaded ones (uniform select, no indexing)
      loaded_value const v  = cs <= 0 ? lv[0] : cs == 1 ? lv[1] : cs == 2 ? lv[2] : lv[3];
      bool const vv         = cs <= 0 ? vld[0] : cs == 1 ? vld[1] : cs == 2 ? vld[2] : vld[3];
      if (HAS_ARG && (kind == FK_ARGMIN || kind == FK_ARGMAX)) {
        bool const is_min  = kind == FK_ARGMIN;
        uint32_t const crow = carry_row[wave][k];
        if (carry_ends && lane == 0) arg_flush(P, f, carry_g, crow, carry_g == first_g);
        bool const valid = g != NO_GROUP && vv;
        uint64_t key     = valid ? arg_order_key(is_min, P.vals[pass.col[cs]].storage, v) : (is_min ? ~0ull : 0ull);
        uint32_t r       = valid ? uint32_t(row) : 0xFFFFFFFFu;
        if (lane == 0 && carry_g != NO_GROUP && g == carry_g) {
          if (is_min) arg_pick<true>(key, r, c, crow);
          else arg_pick<false>(key, r, c, crow);
        }
        if (is_min) arg_scan<true>(key, r, g);
        else arg_scan<false>(key, r, g);
        if (ends_here) arg_flush(P, f, g, r, g == first_g);
        uint64_t const klast = (uint64_t(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(key >> 32)), 63))) << 32) |
                               uint32_t(__builtin_amdgcn_readlane(int(uint32_t(key)), 63));
        uint32_t const rlast = uint32_t(__builtin_amdgcn_readlane(int(r), 63));
        if (lane == 0) {
          carry[wave][k]     = klast;
          carry_row[wave][k] = rlast;
        }
        continue;
      }
      int32_t const op = pass.op[k], xf = pass.xf[k];
      if (carry_ends && lane == 0) sorted_flush(P, f, carry_g, xf_inv(xf, c), carry_g == first_g);
      uint64_t part = sorted_row_part(P, f, kind, pass.ident[k], g, vv, v);
      if (f >= 0 && g != NO_GROUP && vv) part = xf_fwd(xf, part);  // the identity is already in the domain
      if (lane == 0 && carry_g != NO_GROUP && g == carry_g) part = domain_combine(op, part, c);
      part = segmented_scan_op(op, part, g);
      if (ends_here) sorted_flush(P, f, g, xf_inv(xf, part), g == first_g);
      uint64_t const last = (uint64_t(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(part >> 32)), 63))) << 32) |
                            uint32_t(__builtin_amdgcn_readlane(int(uint32_t(part)), 63));
      if (lane == 0) carry[wave][k] = last;
    }
    carry_g = uint32_t(__builtin_amdgcn_readlane(int(g), 63));
  }
  if (carry_g != NO_GROUP && lane == 0) {
    for (int k = 0; k < pass.n; k++) {  // may continue in the next range
      if (HAS_ARG && (pass.kind[k] == FK_ARGMIN || pass.kind[k] == FK_ARGMAX))
        arg_flush(P, pass.field[k], carry_g, carry_row[wave][k], true);
      else
        sorted_flush(P, pass.field[k], carry_g, xf_inv(pass.xf[k], carry[wave][k]), true);
    }
  }
}

__global__ void __launch_bounds__(BLOCK) k_iota_u32(uint32_t* out, int64_t n)
{
  int64_t const i = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (i < n) out[i] = uint32_t(i);
}

void launch_sorted_count(gb_params const* dP, int64_t n, uint32_t* block_counts, hipStream_t s)
{
  int64_t const nb = (n + SORTED_TILE - 1) / SORTED_TILE;
  hipLaunchKernelGGL(k_sorted_count, dim3(nb), dim3(BLOCK), 0, s, dP, block_counts);
}

void launch_sorted_assign(gb_params const* dP, int64_t n, uint32_t const* block_offsets, uint32_t* gid,
                          uint32_t* group_last, hipStream_t s)
{
  int64_t const nb = (n + SORTED_TILE - 1) / SORTED_TILE;
  hipLaunchKernelGGL(k_sorted_assign, dim3(nb), dim3(BLOCK), 0, s, dP, block_offsets, gid, group_last);
}

int64_t sorted_tiles(int64_t n) { return (n + SORTED_TILE - 1) / SORTED_TILE; }

// Wave ranges: enough waves to fill the device, at least 64 rows each.
static void sorted_ranges(int64_t n, int num_cus, int64_t& per, int64_t& W)
{
  int64_t const target = int64_t(num_cus) * 64;
  per = ((n + target - 1) / target + 63) / 64 * 64;
  if (per < 64) per = 64;
  W = (n + per - 1) / per;
}

// ---------------------------------------------------------------------------
// Column statistics on runs.  A value column whose accumulators are all among
// SUM, SUM_OF_SQUARES, COUNT, MIN, MAX and M2 (VARIANCE / STD), stored as int32, int64,
// float32 or float64, is reduced by a kernel compiled for its type: the combining
// operations are fixed at compile time, the segment masks of the six scan
// steps are computed once per chunk for all accumulators, and the partials
// carried between chunks stay in registers.  M2 comes from the sums in the
// same pass where that is exact or within the bound below; the groups where
// it is not (and the groups a wave shares with its neighbours) get the exact
// two-pass M2 (the reference's mean, then squared deviations:
// sort/group_std.cu:60-69 through sort/aggregate.cpp:286-306): groups of at most
// M2_WAVE_ROWS rows are listed and reduced by one wave each
// (k_sorted_stats_m2_list), longer ones are flagged for the row-parallel
// k_sorted_stats_m2, which returns at once when none is.
template <int ST>
struct stats_traits;
template <>
struct stats_traits<ST_I32> {
  using T   = int32_t;
  using Sum = int64_t;
  using Ord = int32_t;  // MIN / MAX domain
};
template <>
struct stats_traits<ST_I64> {
  using T   = int64_t;
  using Sum = int64_t;
  using Ord = int64_t;
};
template <>
struct stats_traits<ST_F64> {
  using T   = double;
  using Sum = double;
  using Ord = int64_t;
};
template <>
struct stats_traits<ST_F32> {
  using T   = float;
  using Sum = double;  // float32 sums and sums of squares accumulate in double, rounded once when stored
  using Ord = int32_t;
};

// Value of a row in the signed MIN / MAX domain (floats: the order code with
// the sign bit flipped, as scan_domain maps them; integers: themselves).
template <int ST>
__device__ __forceinline__ typename stats_traits<ST>::Ord stats_order(typename stats_traits<ST>::T x)
{
  if constexpr (ST == ST_F64) return int64_t(ord64(uint64_t(__double_as_longlong(x))) ^ 0x8000000000000000ull);
  else if constexpr (ST == ST_F32) return int32_t(ord32(__float_as_uint(x)) ^ 0x80000000u);
  else return x;
}

// The accumulator value of a MIN / MAX domain value.
template <int ST>
__device__ __forceinline__ typename stats_traits<ST>::Ord stats_order_inv(typename stats_traits<ST>::Ord t)
{
  if constexpr (ST == ST_F64) return int64_t(uint64_t(t) ^ 0x8000000000000000ull);
  else if constexpr (ST == ST_F32) return int32_t(uint32_t(t) ^ 0x80000000u);
  else return t;
}

// Same-group masks of the six scan steps (see seg_scan): lane l combines the
// partial the step's DPP control brings when that lane holds l's group.
struct seg_masks {
  bool s[6];
};

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ bool seg_same(uint32_t g)
{
  constexpr int NOT_A_GROUP = int(0xFFFFFFFEu);
  return uint32_t(__builtin_amdgcn_update_dpp(NOT_A_GROUP, int(g), CTRL, ROW_MASK, 0xF, false)) == g;
}

__device__ __forceinline__ seg_masks seg_masks_of(uint32_t g)
{
  return seg_masks{{seg_same<0x111, 0xF>(g), seg_same<0x112, 0xF>(g), seg_same<0x114, 0xF>(g),
                    seg_same<0x118, 0xF>(g), seg_same<0x142, 0xA>(g), seg_same<0x143, 0xC>(g)}};
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp64(uint64_t x)
{
  uint32_t const lo = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(x)), CTRL, ROW_MASK, 0xF, false));
  uint32_t const hi = uint32_t(__builtin_amdgcn_update_dpp(0, int(uint32_t(x >> 32)), CTRL, ROW_MASK, 0xF, false));
  return (uint64_t(hi) << 32) | lo;
}

// Segmented inclusive scan with precomputed masks; `F` combines two partials.
template <typename V, typename F>
__device__ __forceinline__ V mscan(V x, seg_masks const& m, F&& f)
{
  auto step = [&](auto ctrl, auto rmask, bool same) {
    constexpr int C = decltype(ctrl)::value, R = decltype(rmask)::value;
    V o;
    if constexpr (sizeof(V) == 8) {
      uint64_t const u = dpp64<C, R>(__builtin_bit_cast(uint64_t, x));
      o                = __builtin_bit_cast(V, u);
    } else {
      o = __builtin_bit_cast(V, uint32_t(__builtin_amdgcn_update_dpp(0, int(__builtin_bit_cast(uint32_t, x)), C, R,
                                                                     0xF, false)));
    }
    if (same) x = f(x, o);
  };
  step(std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xF>{}, m.s[0]);
  step(std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xF>{}, m.s[1]);
  step(std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xF>{}, m.s[2]);
  step(std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xF>{}, m.s[3]);
  step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{}, m.s[4]);
  step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{}, m.s[5]);
  return x;
}

template <typename V>
__device__ __forceinline__ V lane63(V x)
{
  if constexpr (sizeof(V) == 8) {
    uint64_t const u = __builtin_bit_cast(uint64_t, x);
    uint64_t const r = (uint64_t(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), 63))) << 32) |
                       uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), 63));
    return __builtin_bit_cast(V, r);
  } else {
    return __builtin_bit_cast(V, uint32_t(__builtin_amdgcn_readlane(int(__builtin_bit_cast(uint32_t, x)), 63)));
  }
}

// M2 of a whole group from its count n, sum s1 and sum of squares s2 (the
// rule of the partitioned path's m2_from_sums): integer sources while s2 <
// 2^53 (every partial then exact in double) through n·M2 = n·s2 − s1² in 128
// bits; float64 sources with at most 256 values when the rounding error of
// s2 − s1²/n, at most ~3·n·ε·s2, is below 1e-11 of the result.  False: the
// group goes to the exact pass.
template <int ST>
__device__ __forceinline__ bool stats_m2(uint32_t n, typename stats_traits<ST>::Sum s1, double s2, double& m2)
{
  if (n <= 1 && isfinite(s2)) {  // one finite value (or none): no deviation; NaN / inf give NaN below
    m2 = 0.0;
    return true;
  }
  if constexpr (ST == ST_F64 || ST == ST_F32) {  // float32 values: exact squares, sums in double
    if (n > 256) return false;
    double const v = s2 - s1 * (s1 / double(n));
    if (!(v > 0.0) || 3.0 * double(n) * 1.1102230246251565e-16 * s2 > 1e-11 * v) {
      m2 = 0.0;
      return v == 0.0 && s2 == 0.0;
    }
    m2 = v;
    return true;
  } else {
    if (!(s2 < 9007199254740992.0)) return false;
    if (double(n) * s2 < 9.2e18) {  // n·s2 < 2^63, and s1² <= n·s2: 64 bits hold both
      int64_t const num = int64_t(n) * int64_t(s2) - s1 * s1;
      m2                = double(num) / double(n);
      return true;
    }
    __int128 const num = __int128(n) * __int128(uint64_t(s2)) - __int128(s1) * __int128(s1);
    m2                 = double(num) / double(n);
    return true;
  }
}

struct sorted_stats_job {
  int32_t vcol;
  int32_t f_sum, f_count, f_min, f_max, f_m2;  // field index, -1 when not requested
  int32_t f_sumsq;
  int32_t count_rows;                          // also the group row count
  int64_t min_ident, max_ident;                // identities in the MIN / MAX domain
  // the accumulators of the fields above (arrays of G entries)
  uint32_t* p_rows;
  uint32_t* p_count;
  void* p_sum;
  void* p_min;
  void* p_max;
  void* p_sumsq;
  double* p_m2;
  uint8_t* m2_flag;                            // per group: 1 = M2 by the row-parallel exact pass
  uint32_t* m2_list;                           // groups whose M2 one wave each recomputes
  uint32_t* m2_counts;                         // [0] listed groups, [1] flagged groups
};

// Longest group whose exact M2 one wave recomputes; longer groups take the
// row-parallel pass.
constexpr int64_t M2_WAVE_ROWS = 4096;

template <int ST>
struct stats_part {
  uint32_t rows, cnt;
  typename stats_traits<ST>::Sum sum;
  typename stats_traits<ST>::Ord mn, mx;
  double sq;     // Σx² in double (M2; SUM_OF_SQUARES of float64)
  uint64_t isq;  // Σx² of integers, wrapping (SUM_OF_SQUARES of int32 / int64)
};

// A group's partials: stored when the group lies inside one wave's range,
// combined atomically when neighbouring ranges hold rows of it.
// `ends`: no later row of the group follows (this is the group's last flush),
// and `last_row` is then its last row.
template <int ST>
__device__ __forceinline__ void stats_flush(gb_params const& P, sorted_stats_job const& J, uint32_t g,
                                            stats_part<ST> const& a, bool shared, bool ends, int64_t last_row)
{
  using Sum = typename stats_traits<ST>::Sum;
  using Ord = typename stats_traits<ST>::Ord;
  if (J.count_rows) {
    if (shared) aadd(J.p_rows + g, a.rows);
    else *gp(J.p_rows + g) = a.rows;
  }
  if (J.f_count >= 0) {
    if (shared) aadd(J.p_count + g, a.cnt);
    else *gp(J.p_count + g) = a.cnt;
  }
  if (J.f_sum >= 0) {
    if constexpr (ST == ST_F32) {
      float* p = static_cast<float*>(J.p_sum) + g;
      if (shared) gadd_f32(p, float(a.sum));
      else *gp(p) = float(a.sum);
    } else if constexpr (ST == ST_F64) {
      Sum* p = static_cast<Sum*>(J.p_sum) + g;
      if (shared) gadd_f64(p, a.sum);
      else *gp(p) = a.sum;
    } else {
      Sum* p = static_cast<Sum*>(J.p_sum) + g;
      if (shared) aadd(reinterpret_cast<uint64_t*>(p), uint64_t(a.sum));
      else *gp(p) = a.sum;
    }
  }
  // MIN / MAX accumulators: int32 / int64 values, float order codes (unsigned)
  if (J.f_min >= 0) {
    Ord const u = stats_order_inv<ST>(a.mn);
    Ord* p      = static_cast<Ord*>(J.p_min) + g;
    if (!shared) *gp(p) = u;
    else if constexpr (ST == ST_F64) amin(reinterpret_cast<uint64_t*>(p), uint64_t(u));
    else if constexpr (ST == ST_F32) amin(reinterpret_cast<uint32_t*>(p), uint32_t(u));
    else amin(p, u);
  }
  if (J.f_max >= 0) {
    Ord const u = stats_order_inv<ST>(a.mx);
    Ord* p      = static_cast<Ord*>(J.p_max) + g;
    if (!shared) *gp(p) = u;
    else if constexpr (ST == ST_F64) amax(reinterpret_cast<uint64_t*>(p), uint64_t(u));
    else if constexpr (ST == ST_F32) amax(reinterpret_cast<uint32_t*>(p), uint32_t(u));
    else amax(p, u);
  }
  if (J.f_sumsq >= 0) {
    if constexpr (ST == ST_F32) {
      float* p = static_cast<float*>(J.p_sumsq) + g;
      if (shared) gadd_f32(p, float(a.sq));
      else *gp(p) = float(a.sq);
    } else if constexpr (ST == ST_F64) {
      double* p = static_cast<double*>(J.p_sumsq) + g;
      if (shared) gadd_f64(p, a.sq);
      else *gp(p) = a.sq;
    } else {
      uint64_t* p = static_cast<uint64_t*>(J.p_sumsq) + g;
      if (shared) aadd(p, a.isq);
      else *gp(p) = a.isq;
    }
  }
  if (J.f_m2 >= 0) {
    double m2        = 0.0;
    bool const exact = !shared && stats_m2<ST>(a.cnt, a.sum, a.sq, m2);  // a group not shared ends here
    if (exact) {
      if (m2 != 0.0) *gp(J.p_m2 + g) = m2;
      *gp(J.m2_flag + g) = 0;
    } else if (ends) {  // once per group: its last flush
      bool const large   = last_row - int64_t(gp(P.slot_rows)[g]) + 1 > M2_WAVE_ROWS;
      *gp(J.m2_flag + g) = large ? 1 : 0;
      if (large) aadd(J.m2_counts + 1, 1u);
      else *gp(J.m2_list + __hip_atomic_fetch_add(J.m2_counts, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) = g;
    }
  }
}

// Rows per lane in k_sorted_stats: a lane first combines its consecutive rows
// serially, so one cross-lane scan covers 64 * STATS_ROWS rows.
constexpr int STATS_ROWS = 4;

// Two partials of one group combined.
template <int ST>
__device__ __forceinline__ stats_part<ST> stats_combine(stats_part<ST> a, stats_part<ST> const& b)
{
  a.rows += b.rows;
  a.cnt += b.cnt;
  a.sum += b.sum;
  a.sq += b.sq;
  a.isq += b.isq;
  a.mn = a.mn < b.mn ? a.mn : b.mn;
  a.mx = a.mx > b.mx ? a.mx : b.mx;
  return a;
}

// One wave per contiguous row range (the range layout and the carry between
// chunks of k_sorted_reduce), chunks of 64 * STATS_ROWS rows with
// STATS_ROWS consecutive rows per lane.  A lane's rows are combined serially;
// a group that ends inside the lane (before its last row) is finished there
// with the scan total of the lanes before it, and the partial of the lane's
// last group takes part in the segmented scan across the wave.
template <int ST>
__global__ void __launch_bounds__(BLOCK)
  k_sorted_stats(gb_params const* __restrict__ Pp, uint32_t const* __restrict__ gid, int64_t per, int64_t W,
                 sorted_stats_job J)
{
  using T            = typename stats_traits<ST>::T;
  using Sum          = typename stats_traits<ST>::Sum;
  using Ord          = typename stats_traits<ST>::Ord;
  constexpr int K    = STATS_ROWS;
  gb_params const& P = *Pp;
  int const lane     = threadIdx.x & 63;
  int64_t const w    = int64_t(blockIdx.x) * (BLOCK / 64) + (threadIdx.x >> 6);
  if (w >= W) return;
  int64_t const b = w * per;
  int64_t const e = b + per < P.num_rows ? b + per : P.num_rows;
  dcol const c    = P.vals[J.vcol];
  auto const* data = reinterpret_cast<gmem<T const>*>(gp(c.data));
  constexpr bool FLOAT = ST == ST_F64 || ST == ST_F32;
  bool const want_sq  = J.f_m2 >= 0 || (FLOAT && J.f_sumsq >= 0);
  bool const want_isq = !FLOAT && J.f_sumsq >= 0;
  uint32_t const first_g = gid[b];  // may continue from the previous range
  uint32_t carry_g       = NO_GROUP;
  Ord const min_ident    = Ord(J.min_ident);
  Ord const max_ident    = Ord(J.max_ident);
  stats_part<ST> const none{0, 0, Sum(0), min_ident, max_ident, 0.0, 0};
  stats_part<ST> cy = none;
  for (int64_t i0 = b; i0 < e; i0 += 64 * K) {
    int64_t const r0 = i0 + int64_t(K) * lane;  // the lane's first row
    uint32_t gg[K];
    stats_part<ST> a[K];
#pragma unroll
    for (int u = 0; u < K; u++) {
      int64_t const row = r0 + u;
      bool const live   = row < e;
      int64_t const rc  = live ? row : e - 1;
      uint32_t const gl = gid[rc];
      T const x         = data[rc];
      bool const vv     = dcol_valid(c, rc);
      gg[u]             = live ? gl : NO_GROUP;
      bool const has    = gg[u] != NO_GROUP && vv;
      a[u].rows         = gg[u] != NO_GROUP ? 1u : 0u;
      a[u].cnt          = has ? 1u : 0u;
      a[u].sum          = has ? Sum(x) : Sum(0);
      a[u].sq           = has && want_sq ? double(x) * double(x) : 0.0;
      a[u].isq          = has && want_isq ? uint64_t(int64_t(x)) * uint64_t(int64_t(x)) : 0;
      bool ordered      = has;
      if constexpr (FLOAT) ordered = has && !isnan(x);  // NaN takes no part in MIN / MAX
      Ord const t = stats_order<ST>(x);
      a[u].mn     = ordered ? t : min_ident;
      a[u].mx     = ordered ? t : max_ident;
    }
    uint32_t const g0 = uint32_t(__builtin_amdgcn_readlane(int(gg[0]), 0));
    if (carry_g != NO_GROUP && g0 != carry_g) {
      if (lane == 0) stats_flush<ST>(P, J, carry_g, cy, carry_g == first_g, true, i0 - 1);
    } else if (carry_g != NO_GROUP && lane == 0) {  // the carried group continues at lane 0
      a[0] = stats_combine<ST>(a[0], cy);
    }
    // the lane's rows, serially: `first` holds the lane's first group when
    // that group ends inside the lane
    uint32_t cur = gg[0];
    stats_part<ST> R = a[0], first = none;
    bool split = false;
#pragma unroll
    for (int u = 1; u < K; u++) {
      if (gg[u] == cur) {
        R = stats_combine<ST>(R, a[u]);
      } else {
        if (!split) {
          first = R;
          split = true;
        } else if (cur != NO_GROUP) {  // a group wholly inside the lane (more than two rows per lane)
          stats_flush<ST>(P, J, cur, R, cur == first_g, true, r0 + u - 1);
        }
        cur = gg[u];
        R   = a[u];
      }
    }
    seg_masks const m = seg_masks_of(cur);
    auto add32        = [](uint32_t p, uint32_t q) { return p + q; };
    if (J.count_rows) R.rows = mscan(R.rows, m, add32);
    R.cnt = mscan(R.cnt, m, add32);
    R.sum = mscan(R.sum, m, [](Sum p, Sum q) { return p + q; });
    if (J.f_min >= 0) R.mn = mscan(R.mn, m, [](Ord p, Ord q) { return p < q ? p : q; });
    if (J.f_max >= 0) R.mx = mscan(R.mx, m, [](Ord p, Ord q) { return p > q ? p : q; });
    if (want_sq) R.sq = mscan(R.sq, m, [](double p, double q) { return p + q; });
    if (want_isq) R.isq = mscan(R.isq, m, [](uint64_t p, uint64_t q) { return p + q; });
    // a first group that ended inside the lane: its rows in the lanes before
    // (the previous lane's scan total when that lane's last group is the same)
    uint32_t const cprev = uint32_t(__builtin_amdgcn_update_dpp(int(NO_GROUP), int(cur), 0x138, 0xF, 0xF, false));
    stats_part<ST> prev;
    prev.rows = uint32_t(__builtin_amdgcn_update_dpp(0, int(R.rows), 0x138, 0xF, 0xF, false));
    prev.cnt  = uint32_t(__builtin_amdgcn_update_dpp(0, int(R.cnt), 0x138, 0xF, 0xF, false));
    prev.sum  = __builtin_bit_cast(Sum, dpp64<0x138, 0xF>(__builtin_bit_cast(uint64_t, R.sum)));
    if constexpr (sizeof(Ord) == 8) {
      prev.mn = __builtin_bit_cast(Ord, dpp64<0x138, 0xF>(__builtin_bit_cast(uint64_t, R.mn)));
      prev.mx = __builtin_bit_cast(Ord, dpp64<0x138, 0xF>(__builtin_bit_cast(uint64_t, R.mx)));
    } else {
      prev.mn = Ord(__builtin_amdgcn_update_dpp(0, int(R.mn), 0x138, 0xF, 0xF, false));
      prev.mx = Ord(__builtin_amdgcn_update_dpp(0, int(R.mx), 0x138, 0xF, 0xF, false));
    }
    prev.sq  = __builtin_bit_cast(double, dpp64<0x138, 0xF>(__builtin_bit_cast(uint64_t, R.sq)));
    prev.isq = dpp64<0x138, 0xF>(R.isq);
    if (split && gg[0] != NO_GROUP) {
      stats_part<ST> const tot = lane > 0 && cprev == gg[0] ? stats_combine<ST>(first, prev) : first;
      stats_flush<ST>(P, J, gg[0], tot, gg[0] == first_g, true, r0);
    }
    // the lane's last group ends at the lane's last row when the next lane
    // starts another group
    uint32_t const gnext = uint32_t(__builtin_amdgcn_update_dpp(int(NO_GROUP), int(gg[0]), 0x130, 0xF, 0xF, false));
    if (cur != NO_GROUP && lane < 63 && gnext != cur) stats_flush<ST>(P, J, cur, R, cur == first_g, true, r0 + K - 1);
    cy.rows = lane63(R.rows);
    cy.cnt  = lane63(R.cnt);
    cy.sum  = lane63(R.sum);
    cy.mn   = lane63(R.mn);
    cy.mx   = lane63(R.mx);
    cy.sq   = lane63(R.sq);
    cy.isq  = lane63(R.isq);
    carry_g = uint32_t(__builtin_amdgcn_readlane(int(cur), 63));
  }
  if (carry_g != NO_GROUP && lane == 0) {  // may continue in the next range
    bool const ends = e == P.num_rows || gid[e] != carry_g;
    stats_flush<ST>(P, J, carry_g, cy, true, ends, e - 1);
  }
}

// The exact M2 of the flagged (long) groups: Σ(x − mean)² with the mean from
// the finished sum and count, summed over each run of a chunk and added to
// the group's M2 once per chunk.  Chunks without a flagged row only read
// their group numbers and flags; with no flagged group every wave returns
// after reading the count.
template <int ST>
__global__ void __launch_bounds__(BLOCK)
  k_sorted_stats_m2(gb_params const* __restrict__ Pp, uint32_t const* __restrict__ gid, sorted_stats_job J)
{
  using T            = typename stats_traits<ST>::T;
  gb_params const& P = *Pp;
  if (aload(J.m2_counts + 1) == 0) return;
  int64_t const n    = P.num_rows;
  int const lane     = threadIdx.x & 63;
  dcol const c       = P.vals[J.vcol];
  auto const* data   = reinterpret_cast<gmem<T const>*>(gp(c.data));
  field_desc const& F = P.fields[J.f_m2];
  for (int64_t base = int64_t(blockIdx.x) * BLOCK + (threadIdx.x & ~63); base < n;
       base += int64_t(gridDim.x) * BLOCK) {
    int64_t const row = base + lane;
    uint32_t const g  = row < n ? gid[row] : NO_GROUP;
    bool const fl     = g != NO_GROUP && gp(J.m2_flag)[g] != 0;
    if (__builtin_amdgcn_ballot_w64(fl) == 0) continue;
    double d2 = 0.0;
    if (fl && dcol_valid(c, row)) {
      double const cnt  = double(count_of(P, F.aux_count, g));
      double const mean = field_as_double(P.fields[F.aux_sum], g) / cnt;
      double const d    = double(data[row]) - mean;
      d2                = d * d;
    }
    uint32_t const gk  = fl ? g : NO_GROUP;
    seg_masks const m  = seg_masks_of(gk);
    d2                 = mscan(d2, m, [](double p, double q) { return p + q; });
    uint32_t const gn  = uint32_t(__builtin_amdgcn_update_dpp(int(NO_GROUP), int(gk), 0x130, 0xF, 0xF, false));
    if (fl && (lane == 63 || gn != gk) && d2 != 0.0) gadd_f64(reinterpret_cast<double*>(field_at(F, g)), d2);
  }
}

// The exact M2 of the listed groups (at most M2_WAVE_ROWS rows each): one
// wave per group walks its rows from the first (the runs are contiguous),
// sums the squared deviations in its lanes, reduces them across the wave and
// stores the group's M2 (the only writer).
template <int ST>
__global__ void __launch_bounds__(BLOCK)
  k_sorted_stats_m2_list(gb_params const* __restrict__ Pp, uint32_t const* __restrict__ gid, sorted_stats_job J)
{
  using T             = typename stats_traits<ST>::T;
  gb_params const& P  = *Pp;
  int64_t const n     = P.num_rows;
  int const lane      = threadIdx.x & 63;
  dcol const c        = P.vals[J.vcol];
  auto const* data    = reinterpret_cast<gmem<T const>*>(gp(c.data));
  field_desc const& F = P.fields[J.f_m2];
  uint32_t const listed = aload(J.m2_counts);
  for (uint32_t i = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); i < listed; i += gridDim.x * (BLOCK / 64)) {
    uint32_t const g    = gp(J.m2_list)[i];
    int64_t const first = int64_t(gp(P.slot_rows)[g]);
    double const cnt    = double(count_of(P, F.aux_count, g));
    double const mean   = field_as_double(P.fields[F.aux_sum], g) / cnt;
    double acc          = 0.0;
    for (int64_t r0 = first; r0 < n; r0 += 64) {
      int64_t const row = r0 + lane;
      bool const in     = row < n && gid[row] == g;
      if (in && dcol_valid(c, row)) {
        double const d = double(data[row]) - mean;
        acc += d * d;
      }
      if (__builtin_amdgcn_ballot_w64(in) != ~0ull) break;  // the run ended in this chunk
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) *gp(reinterpret_cast<double*>(field_at(F, g))) = acc;
  }
}

__global__ void k_zero_u32(uint32_t* p, int n)
{
  if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}

template <int ST>
static void launch_stats_job(gb_params const* dP, sorted_stats_job const& J, int64_t n, uint32_t const* gid,
                             int num_cus, hipStream_t s)
{
  int64_t per, W;
  sorted_ranges(n, num_cus, per, W);
  // ranges of whole chunks: a range then ends at a chunk's last lane, so a
  // group at a range's end is always carried (and combined atomically)
  constexpr int64_t CH = 64 * STATS_ROWS;
  per = (per + CH - 1) / CH * CH;
  W   = (n + per - 1) / per;
  unsigned const blocks = unsigned((W + BLOCK / 64 - 1) / (BLOCK / 64));
  hipLaunchKernelGGL(k_sorted_stats<ST>, dim3(blocks), dim3(BLOCK), 0, s, dP, gid, per, W, J);
}

template <int ST>
static void launch_stats_m2(gb_params const* dP, sorted_stats_job const& J, int64_t n, uint32_t const* gid,
                            int num_cus, hipStream_t s)
{
  hipLaunchKernelGGL(k_sorted_stats_m2_list<ST>, dim3(unsigned(num_cus * 8)), dim3(BLOCK), 0, s, dP, gid, J);
  hipLaunchKernelGGL(k_sorted_stats_m2<ST>, dim3(unsigned(row_grid(n, num_cus, 8))), dim3(BLOCK), 0, s, dP, gid, J);
}

static bool stats_kind_ok(int32_t st, int32_t kind)
{
  switch (kind) {
    case FK_COUNT:
    case FK_M2: return true;
    case FK_SUM_I64: return st == ST_I32 || st == ST_I64;
    case FK_SUM_F64: return st == ST_F64;
    case FK_SUMSQ_I64: return st == ST_I32 || st == ST_I64;
    case FK_SUMSQ_F64: return st == ST_F64;
    case FK_SUM_F32:
    case FK_SUMSQ_F32:
    case FK_MIN_F32:
    case FK_MAX_F32: return st == ST_F32;
    case FK_MIN_S32:
    case FK_MAX_S32: return st == ST_I32;
    case FK_MIN_S64:
    case FK_MAX_S64: return st == ST_I64;
    case FK_MIN_F64:
    case FK_MAX_F64: return st == ST_F64;
    default: return false;
  }
}

// The column jobs of a request: every value column whose fields all have a
// specialised form.  The first job also counts the group rows when the
// request needs them.
// Device scratch of one column with M2: per-group flags, the list of groups
// and two counters.
static size_t m2_column_bytes(int64_t G)
{
  size_t const g = size_t(G > 0 ? G : 1);
  return ((g + 255) & ~size_t(255)) + ((4 * g + 255) & ~size_t(255)) + 256;
}

size_t sorted_stats_scratch_bytes(int64_t G, int m2_columns) { return m2_column_bytes(G) * size_t(m2_columns); }

static std::vector<std::pair<int32_t, sorted_stats_job>> stats_jobs(gb_params const& hP, uint8_t* m2_scratch,
                                                                   int64_t G)
{
  std::vector<std::pair<int32_t, sorted_stats_job>> jobs;
  int64_t m2_jobs = 0;
  for (int j = 0; j < hP.num_vals; j++) {
    int32_t const st = hP.vals[j].storage;
    int const fb = hP.val_field_begin[j], fe = hP.val_field_begin[j + 1];
    if (fb == fe || (st != ST_I32 && st != ST_I64 && st != ST_F64 && st != ST_F32)) continue;
    sorted_stats_job J{};
    J.vcol = j;
    J.f_sum = J.f_count = J.f_min = J.f_max = J.f_m2 = J.f_sumsq = -1;
    bool ok = true;
    // the kernels index accumulators as dense arrays
    if (hP.need_count_all && hP.count_stride != 4) continue;
    for (int f = fb; f < fe && ok; f++) {
      int32_t const k = hP.fields[f].kind;
      ok              = stats_kind_ok(st, k) && hP.fields[f].stride == hP.fields[f].width;
      int32_t* slot   = k == FK_COUNT ? &J.f_count
                        : k == FK_M2  ? &J.f_m2
                        : (k == FK_SUMSQ_I64 || k == FK_SUMSQ_F64 || k == FK_SUMSQ_F32) ? &J.f_sumsq
                        : (k == FK_SUM_I64 || k == FK_SUM_F64 || k == FK_SUM_F32)       ? &J.f_sum
                        : (k == FK_MIN_S32 || k == FK_MIN_S64 || k == FK_MIN_F64 || k == FK_MIN_F32) ? &J.f_min
                                                                                   : &J.f_max;
      if (ok) {
        ok    = *slot < 0;
        *slot = f;
      }
    }
    if (!ok || (J.f_m2 >= 0 && (J.f_sum < 0 || hP.fields[J.f_m2].aux_sum != J.f_sum))) continue;
    auto ptr = [&](int32_t f) { return f >= 0 ? hP.fields[f].ptr : nullptr; };
    J.p_rows  = hP.slot_count;
    J.p_count = static_cast<uint32_t*>(ptr(J.f_count));
    J.p_sum   = ptr(J.f_sum);
    J.p_min   = ptr(J.f_min);
    J.p_max   = ptr(J.f_max);
    J.p_sumsq = ptr(J.f_sumsq);
    J.p_m2    = static_cast<double*>(ptr(J.f_m2));
    if (st == ST_F32) {  // float32 order codes with the sign bit flipped: a signed 32-bit domain
      J.min_ident = int64_t(int32_t(uint32_t(field_identity(FK_MIN_F32)) ^ 0x80000000u));
      J.max_ident = int64_t(int32_t(uint32_t(field_identity(FK_MAX_F32)) ^ 0x80000000u));
    } else {
      int32_t const kmin = st == ST_I32 ? FK_MIN_S32 : st == ST_I64 ? FK_MIN_S64 : FK_MIN_F64;
      int32_t const kmax = st == ST_I32 ? FK_MAX_S32 : st == ST_I64 ? FK_MAX_S64 : FK_MAX_F64;
      int32_t op, xf;
      scan_domain(kmin, op, xf);
      J.min_ident = int64_t(xf_fwd(xf, field_identity(kmin)));
      scan_domain(kmax, op, xf);
      J.max_ident = int64_t(xf_fwd(xf, field_identity(kmax)));
    }
    if (J.f_m2 >= 0 && m2_scratch != nullptr) {  // scratch of its own per column
      size_t const g  = size_t(G > 0 ? G : 1);
      uint8_t* base   = m2_scratch + m2_column_bytes(G) * size_t(m2_jobs++);
      J.m2_flag       = base;
      J.m2_list       = reinterpret_cast<uint32_t*>(base + ((g + 255) & ~size_t(255)));
      J.m2_counts     = reinterpret_cast<uint32_t*>(base + ((g + 255) & ~size_t(255)) + ((4 * g + 255) & ~size_t(255)));
    }
    jobs.push_back({st, J});
  }
  if (!jobs.empty() && hP.need_count_all) jobs.front().second.count_rows = 1;
  return jobs;
}

sorted_stats_cover sorted_stats_plan(gb_params const& hP)
{
  sorted_stats_cover cover;
  cover.field.assign(size_t(hP.num_fields), 0);
  for (auto const& [st, J] : stats_jobs(hP, nullptr, 0)) {
    for (int f = hP.val_field_begin[J.vcol]; f < hP.val_field_begin[J.vcol + 1]; f++) cover.field[size_t(f)] = 1;
    cover.rows |= J.count_rows != 0;
    cover.m2_columns += J.f_m2 >= 0 ? 1 : 0;
  }
  return cover;
}

void launch_sorted_stats(gb_params const* dP, gb_params const& hP, int64_t n, int64_t G, uint32_t const* gid,
                         int num_cus, uint8_t* m2_scratch, hipStream_t s)
{
  if (n <= 0) return;
  for (auto const& [st, J] : stats_jobs(hP, m2_scratch, G)) {
    if (J.f_m2 >= 0) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, J.m2_counts, 2);
    if (st == ST_I32) launch_stats_job<ST_I32>(dP, J, n, gid, num_cus, s);
    else if (st == ST_I64) launch_stats_job<ST_I64>(dP, J, n, gid, num_cus, s);
    else if (st == ST_F32) launch_stats_job<ST_F32>(dP, J, n, gid, num_cus, s);
    else launch_stats_job<ST_F64>(dP, J, n, gid, num_cus, s);
  }
}

void launch_sorted_stats_m2(gb_params const* dP, gb_params const& hP, int64_t n, int64_t G, uint32_t const* gid,
                            int num_cus, uint8_t* m2_scratch, hipStream_t s)
{
  if (n <= 0) return;
  for (auto const& [st, J] : stats_jobs(hP, m2_scratch, G)) {
    if (J.f_m2 < 0) continue;
    if (st == ST_I32) launch_stats_m2<ST_I32>(dP, J, n, gid, num_cus, s);
    else if (st == ST_I64) launch_stats_m2<ST_I64>(dP, J, n, gid, num_cus, s);
    else if (st == ST_F32) launch_stats_m2<ST_F32>(dP, J, n, gid, num_cus, s);
    else launch_stats_m2<ST_F64>(dP, J, n, gid, num_cus, s);
  }
}

static void launch_sorted_passes(gb_params const* dP, gb_params const& hP, int64_t n, uint32_t const* gid,
                                 int num_cus, bool m2, sorted_stats_cover const* cover, hipStream_t s)
{
  std::vector<int32_t> items;
  if (!m2 && hP.need_count_all && !(cover && cover->rows)) items.push_back(-1);
  for (int f = 0; f < hP.num_fields; f++) {
    int32_t const k = hP.fields[f].kind;
    if (cover && cover->field[size_t(f)]) continue;  // launch_sorted_stats
    if (k == FK_NUNIQUE || k == FK_NUNIQUE_NULLS || k == FK_NTH) continue;  // kernels of their own
    if ((k == FK_ARGMIN || k == FK_ARGMAX) && hP.vals[hP.fields[f].vcol].storage == ST_STRING)
      continue;  // launch_sorted_string_arg
    if ((hP.fields[f].kind == FK_M2) == m2) items.push_back(f);
  }
  int64_t per, W;
  sorted_ranges(n, num_cus, per, W);
  unsigned const blocks = unsigned((W + BLOCK / 64 - 1) / (BLOCK / 64));
  size_t i = 0;
  while (i < items.size()) {
    sorted_pass pass{};
    while (i < items.size() && pass.n < SORTED_FPP) {
      int32_t const f = items[i];
      int32_t slot    = -1;
      if (f >= 0) {
        int32_t const vc = hP.fields[f].vcol;
        for (int c = 0; c < pass.ncols; c++)
          if (pass.col[c] == vc) slot = c;
        if (slot < 0) {
          if (pass.ncols == SORTED_COLS) break;  // next pass
          slot                  = pass.ncols;
          pass.col[pass.ncols++] = vc;
        }
      }
      int32_t const kind  = f < 0 ? int32_t(FK_COUNT) : hP.fields[f].kind;
      pass.field[pass.n]  = f;
      pass.col_of[pass.n] = slot;
      pass.kind[pass.n]   = kind;
      scan_domain(kind, pass.op[pass.n], pass.xf[pass.n]);
      pass.ident[pass.n]  = kind == FK_M2 ? 0ull : xf_fwd(pass.xf[pass.n], field_identity(kind));
      pass.n++;
      i++;
    }
    bool has_arg = false;
    for (int k = 0; k < pass.n; k++) has_arg |= pass.kind[k] == FK_ARGMIN || pass.kind[k] == FK_ARGMAX;
    if (has_arg) hipLaunchKernelGGL(k_sorted_reduce<true>, dim3(blocks), dim3(BLOCK), 0, s, dP, gid, per, W, pass);
    else hipLaunchKernelGGL(k_sorted_reduce<false>, dim3(blocks), dim3(BLOCK), 0, s, dP, gid, per, W, pass);
  }
}

// ARGMIN / ARGMAX of a string column on runs (and through them string MIN /
// MAX): each wave reduces its rows of every run to the run's best row with
// string comparisons (segmented over lanes of equal group numbers, which are
// contiguous), and the run's first lane folds that row into the group's
// accumulator with the compare-and-swap of the device-wide path.  Equal
// strings keep the smaller row, so the result does not depend on the order
// of the updates.
__global__ void __launch_bounds__(BLOCK)
  k_sorted_string_arg(gb_params const* __restrict__ Pp, int32_t f, int64_t n, uint32_t const* __restrict__ gid)
{
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  gb_params const& P  = *Pp;
  field_desc const& F = P.fields[f];
  dcol const& c       = P.vals[F.vcol];
  int const lane      = threadIdx.x & 63;
  for (int64_t base = int64_t(blockIdx.x) * BLOCK + (threadIdx.x & ~63); base < n;
       base += int64_t(gridDim.x) * BLOCK) {
    int64_t const row = base + lane;
    uint32_t const g  = row < n ? gid[row] : NO_GROUP;
    uint32_t best     = (g != NO_GROUP && dcol_valid(c, row)) ? uint32_t(row) : NONE;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t const other = uint32_t(__shfl_down(int(best), d));
      uint32_t const og    = uint32_t(__shfl_down(int(g), d));
      if (lane + d < 64 && og == g && other != NONE &&
          (best == NONE || arg_better(c, F.kind, int64_t(other), int64_t(best))))
        best = other;
    }
    uint32_t const gprev = uint32_t(__shfl_up(int(g), 1));
    bool const head      = lane == 0 || gprev != g;
    if (head && g != NO_GROUP && best != NONE) arg_update(F, int64_t(g), c, int64_t(best));
  }
}

void launch_sorted_string_arg(gb_params const* dP, gb_params const& hP, int64_t n, uint32_t const* gid,
                              hipStream_t s)
{
  if (n <= 0) return;
  unsigned const blocks = unsigned(std::min<int64_t>((n + BLOCK - 1) / BLOCK, int64_t(1) << 20));
  for (int f = 0; f < hP.num_fields; f++) {
    int32_t const k = hP.fields[f].kind;
    if ((k == FK_ARGMIN || k == FK_ARGMAX) && hP.vals[hP.fields[f].vcol].storage == ST_STRING)
      hipLaunchKernelGGL(k_sorted_string_arg, dim3(blocks), dim3(BLOCK), 0, s, dP, f, n, gid);
  }
}

// Group rows of the gathered (group-ordered) columns back to input rows:
// rows[g] = perm[rows[g]] where the row is set (-1 stays).
__global__ void __launch_bounds__(BLOCK)
  k_map_rows(int32_t* __restrict__ rows, int64_t G, uint32_t const* __restrict__ perm)
{
  int64_t const g = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (g < G && rows[g] >= 0) rows[g] = int32_t(perm[rows[g]]);
}

void launch_map_rows(int32_t* rows, int64_t G, uint32_t const* perm, hipStream_t s)
{
  if (G > 0) hipLaunchKernelGGL(k_map_rows, dim3(unsigned((G + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, rows, G, perm);
}

void launch_sorted_aggregate(gb_params const* dP, gb_params const& hP, int64_t n, uint32_t const* gid, int num_cus,
                             sorted_stats_cover const* cover, hipStream_t s)
{
  launch_sorted_passes(dP, hP, n, gid, num_cus, false, cover, s);
}

void launch_sorted_second_pass(gb_params const* dP, gb_params const& hP, int64_t n, uint32_t const* gid,
                               int num_cus, sorted_stats_cover const* cover, hipStream_t s)
{
  launch_sorted_passes(dP, hP, n, gid, num_cus, true, cover, s);
}

// NUNIQUE on runs.  The reference sorts each group's values and counts the
// heads of equal runs (sort/group_nunique.cu:44-64).  Here a group whose rows
// all lie in one 64-row chunk of a wave is counted inside the wave: a counted
// row is a new value when no earlier row of its group holds an equal one
// (shuffles over at most the group's length).  Every other counted row (longer
// or chunk-crossing groups, and strings) inserts (group, value) into an
// open-addressing set of row indices, and the row that claims an empty slot
// adds one distinct value to its group.  Rows
// are equal when they share the group and both are null (counted only with
// nulls included) or hold the same value, -0.0 equal to 0.0 and NaN equal to
// NaN (the row comparator's defaults).  Group g probes only its own region
// of the set, slots [2 * first, 2 * end) where [first, end) are the rows from
// its first row to the next group's: the region has at least twice as many
// slots as the group has rows, so every probe sequence reaches an empty slot
// or a match, and a group's probes stay within a few cache lines (the groups
// of sorted keys are contiguous rows).
__device__ __forceinline__ uint64_t nunique_bits(dcol const& c, int64_t row)
{
  uint64_t b = raw_bits(c, row);
  if (c.storage == ST_F32) {
    float const x = __uint_as_float(uint32_t(b));
    b             = isnan(x) ? 0x7fc00000ull : (x == 0.0f ? 0ull : b);
  } else if (c.storage == ST_F64) {
    double const x = __longlong_as_double(int64_t(b));
    b              = isnan(x) ? 0x7ff8000000000000ull : (x == 0.0 ? 0ull : b);
  } else if (c.storage == ST_BOOL) {
    b = b != 0;
  }
  return b;
}

__global__ void __launch_bounds__(BLOCK) k_fill_empty(uint32_t* __restrict__ out, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += int64_t(gridDim.x) * BLOCK)
    out[i] = 0xFFFFFFFFu;
}

__global__ void __launch_bounds__(BLOCK)
  k_sorted_nunique(gb_params const* __restrict__ Pp, int32_t f, int64_t n, int64_t G,
                   uint32_t const* __restrict__ gid, uint32_t* __restrict__ table)
{
  constexpr uint32_t EMPTY = 0xFFFFFFFFu;
  gb_params const& P   = *Pp;
  field_desc const& F  = P.fields[f];
  dcol const& c        = P.vals[F.vcol];
  bool const with_null = F.kind == FK_NUNIQUE_NULLS;
  bool const str       = c.storage == ST_STRING;  // strings: hashed, then compared byte-wise
  int const lane       = threadIdx.x & 63;
  // each wave takes 64 consecutive rows per step (the loop bound is uniform
  // across the wave, so every lane reaches the ballots below)
  for (int64_t base0 = int64_t(blockIdx.x) * BLOCK + (threadIdx.x & ~63); base0 < n;
       base0 += int64_t(gridDim.x) * BLOCK) {
    int64_t const row = base0 + lane;
    uint32_t const g  = row < n ? gid[row] : NO_GROUP;
    bool const valid  = g != NO_GROUP && dcol_valid(c, row);
    bool const counts = g != NO_GROUP && (valid || with_null);
    bool inserted     = false;
    uint64_t const b  = !valid ? 0ull : (str ? uint64_t(elem_hash(c, row, 0u)) : nunique_bits(c, row));
    uint64_t first = 0, end = 0;
    if (g != NO_GROUP) {
      first = P.slot_rows[g];
      end   = int64_t(g) + 1 < G ? P.slot_rows[g + 1] : uint64_t(n);
    }
    // A group whose rows all lie in this chunk is counted here: a counted
    // lane is a new value when no lower lane of its group holds an equal one
    // (the group's lanes are contiguous, from lane first - base0).
    bool const local    = counts && !str && first >= uint64_t(base0) && end <= uint64_t(base0 + 64);
    uint64_t const lmask = __ballot(local);
    if (lmask != 0) {
      uint64_t const cmask = __ballot(counts);
      uint64_t const vmask = __ballot(valid);
      int dist             = local ? lane - int(first - uint64_t(base0)) : 0;
      int maxd             = dist;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) maxd = max(maxd, __shfl_xor(maxd, off));
      bool dup = false;
      for (int d = 1; d <= maxd; d++) {
        uint64_t const ob = __shfl_up(b, d);
        int const o       = lane - d;  // the other lane, in the group when d <= dist
        if (d <= dist && ((cmask >> o) & 1u) && (((vmask >> o) & 1u) != 0) == valid && (!valid || ob == b))
          dup = true;
      }
      inserted = local && !dup;
    }
    if (counts && !local) {
      uint32_t const h     = mm_fmix32(uint32_t(b) ^ mm_fmix32(uint32_t(b >> 32) ^ uint32_t(valid)));
      uint64_t const hbase = 2 * first, len = 2 * (end - first);
      uint64_t off         = (uint64_t(h) * len) >> 32;  // in [0, len)
      for (;;) {
        uint64_t const slot = hbase + off;
        uint32_t cur = __hip_atomic_load(&table[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == EMPTY) {
          uint32_t const prev = atomicCAS(&table[slot], EMPTY, uint32_t(row));
          if (prev == EMPTY) {
            inserted = true;
            break;
          }
          cur = prev;
        }
        // the region holds rows of group g only
        bool const cv = dcol_valid(c, cur);
        if (cv == valid && (!valid || (str ? elem_equal(c, cur, row) : nunique_bits(c, cur) == b)))
          break;  // already counted
        off = off + 1 == len ? 0 : off + 1;
      }
    }
    // one add per run of equal groups in the wave (group numbers do not
    // decrease along the rows): the run's first lane adds the run's insertions
    uint32_t const gp_  = uint32_t(__shfl(int(g), lane == 0 ? 0 : lane - 1));
    bool const head     = lane == 0 || gp_ != g;
    uint64_t const heads = __ballot(head);
    uint64_t const ins   = __ballot(inserted);
    if (head && g != NO_GROUP) {
      uint64_t const later = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);  // heads above this lane
      uint64_t const upto  = later ? ((later & (~later + 1ull)) - 1ull) : ~0ull;     // lanes below the next head
      uint32_t const cnt   = uint32_t(__popcll(ins & upto & ~((1ull << lane) - 1ull)));
      if (cnt) atomicAdd(reinterpret_cast<unsigned int*>(field_at(F, g)), cnt);
    }
  }
}

// NTH_ELEMENT on runs (sort/group_nth_element.cu:50-140): the group's n-th
// row in row order (from the end for n < 0), counting every row under
// INCLUDE and rows with a valid value under EXCLUDE; none when the group has
// too few.  Group g holds rows first..last (slot_rows[g],
// group_last[g]) with no dropped row among them, so when every row counts
// (nulls included, or a value column without nulls) the n-th row is found by
// index arithmetic, one thread per group.
__global__ void __launch_bounds__(BLOCK)
  k_sorted_nth_direct(gb_params const* __restrict__ Pp, int32_t f, int64_t G, uint32_t const* __restrict__ group_last)
{
  int64_t const g = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (g >= G) return;
  gb_params const& P  = *Pp;
  field_desc const& F = P.fields[f];
  int64_t const nth   = F.aux_sum;
  int64_t const first = int64_t(P.slot_rows[g]);
  int64_t const last  = min(int64_t(group_last[g]), P.num_rows - 1);
  int64_t const size  = last - first + 1;
  uint32_t found      = 0xFFFFFFFFu;
  if (nth >= 0 ? nth < size : -nth <= size) found = uint32_t(nth >= 0 ? first + nth : last + 1 + nth);
  *reinterpret_cast<uint32_t*>(field_at(F, g)) = found;
}

// The r-th (0-based) set bit of `bits`, counted from the low end or, when
// `high`, from the high end; r < popcount(bits).
__device__ __forceinline__ int select_bit(uint32_t bits, int r, bool high)
{
  if (high) {
    for (int i = 0; i < r; i++) bits &= ~(0x80000000u >> __clz(bits));
    return 31 - __clz(bits);
  }
  for (int i = 0; i < r; i++) bits &= bits - 1u;
  return __ffs(bits) - 1;
}

// NTH_ELEMENT on runs with null values excluded.  The group's rows are bits
// b0..b1 of the value column's mask; the n-th valid row (from the front for
// n >= 0, from the back for n < 0) is found by counting set bits word by
// word.  Groups of at most NTH_THREAD_WORDS mask words are counted by one
// thread each; larger groups are listed (large[0] = count, large[1..] =
// groups) for k_sorted_nth_scan, where a wave counts 64 words per step.
constexpr int NTH_THREAD_WORDS = 8;

struct nth_bits {
  int64_t b0, b1, w0, w1;
};

__device__ __forceinline__ nth_bits nth_group_bits(gb_params const& P, dcol const& c, uint32_t const* group_last,
                                                   int64_t g)
{
  nth_bits r;
  r.b0 = int64_t(P.slot_rows[g]) + c.mask_offset;
  r.b1 = min(int64_t(group_last[g]), P.num_rows - 1) + c.mask_offset;
  r.w0 = r.b0 >> 5;
  r.w1 = r.b1 >> 5;
  return r;
}

// Mask word w restricted to the group's bits.
__device__ __forceinline__ uint32_t nth_word(gmem<uint32_t const>* mask, nth_bits const& q, int64_t w)
{
  uint32_t bits = mask[w];
  if (w == q.w0) bits &= ~0u << (q.b0 & 31);
  if (w == q.w1 && (q.b1 & 31) != 31) bits &= (2u << (q.b1 & 31)) - 1u;
  return bits;
}

__global__ void __launch_bounds__(BLOCK)
  k_sorted_nth_small(gb_params const* __restrict__ Pp, int32_t f, int64_t G, uint32_t const* __restrict__ group_last,
                     uint32_t* __restrict__ large)
{
  int64_t const g = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (g >= G) return;
  gb_params const& P  = *Pp;
  field_desc const& F = P.fields[f];
  dcol const& c       = P.vals[F.vcol];
  nth_bits const q    = nth_group_bits(P, c, group_last, g);
  if (q.w1 - q.w0 >= NTH_THREAD_WORDS) {
    uint32_t const i = atomicAdd(&large[0], 1u);
    large[1 + i]     = uint32_t(g);
    return;
  }
  auto const mask    = gp(c.mask);
  int64_t const nth  = F.aux_sum;
  bool const back    = nth < 0;
  int64_t const want = back ? -nth - 1 : nth;
  int64_t seen       = 0;
  uint32_t found     = 0xFFFFFFFFu;
  for (int64_t k = 0; k <= q.w1 - q.w0; k++) {
    int64_t const w     = back ? q.w1 - k : q.w0 + k;
    uint32_t const bits = nth_word(mask, q, w);
    int const cnt       = __popc(bits);
    if (seen + cnt > want) {
      found = uint32_t((w << 5) + select_bit(bits, int(want - seen), back) - c.mask_offset);
      break;
    }
    seen += cnt;
  }
  *reinterpret_cast<uint32_t*>(field_at(F, g)) = found;
}

__global__ void __launch_bounds__(BLOCK)
  k_sorted_nth_scan(gb_params const* __restrict__ Pp, int32_t f, uint32_t const* __restrict__ group_last,
                    uint32_t const* __restrict__ large)
{
  gb_params const& P  = *Pp;
  field_desc const& F = P.fields[f];
  dcol const& c       = P.vals[F.vcol];
  auto const mask     = gp(c.mask);
  int64_t const nth   = F.aux_sum;
  bool const back     = nth < 0;
  int64_t const want  = back ? -nth - 1 : nth;  // 0-based among the valid rows, from the chosen end
  int const lane      = threadIdx.x & 63;
  int64_t const waves = int64_t(gridDim.x) * (BLOCK / 64);
  int64_t const count = int64_t(large[0]);
  for (int64_t i = (int64_t(blockIdx.x) * BLOCK + threadIdx.x) >> 6; i < count; i += waves) {
    int64_t const g  = int64_t(large[1 + i]);
    nth_bits const q = nth_group_bits(P, c, group_last, g);
    int64_t seen     = 0;
    uint32_t found   = 0xFFFFFFFFu;
    for (int64_t k = 0; k <= q.w1 - q.w0; k += 64) {
      int64_t const w = back ? q.w1 - k - lane : q.w0 + k + lane;
      uint32_t bits   = 0;
      if (w >= q.w0 && w <= q.w1) bits = nth_word(mask, q, w);
      int const cnt = __popc(bits);
      int upto      = cnt;  // inclusive prefix over the lanes
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        int const t = __shfl_up(upto, d);
        if (lane >= d) upto += t;
      }
      int const total = __shfl(upto, 63);
      if (seen + total > want) {
        int64_t const r = want - seen;  // < total
        bool const mine = upto - cnt <= r && r < upto;
        uint32_t row    = 0;
        if (mine) row = uint32_t((w << 5) + select_bit(bits, int(r - (upto - cnt)), back) - c.mask_offset);
        int const src = __ffsll(static_cast<unsigned long long>(__ballot(mine))) - 1;
        found         = uint32_t(__shfl(int(row), src));
        break;
      }
      seen += total;
    }
    if (lane == 0) *reinterpret_cast<uint32_t*>(field_at(F, g)) = found;
  }
}

void launch_sorted_nth(gb_params const* dP, gb_params const& hP, int32_t f, int64_t G, uint32_t const* group_last,
                       uint32_t* large, hipStream_t s)
{
  if (G <= 0) return;
  field_desc const& F = hP.fields[f];
  unsigned const per_group = unsigned((G + BLOCK - 1) / BLOCK);
  if (F.aux_count != 0 || hP.vals[F.vcol].mask == nullptr) {
    hipLaunchKernelGGL(k_sorted_nth_direct, dim3(per_group), dim3(BLOCK), 0, s, dP, f, G, group_last);
    return;
  }
  MIG_HIP_TRY(hipMemsetAsync(large, 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_sorted_nth_small, dim3(per_group), dim3(BLOCK), 0, s, dP, f, G, group_last, large);
  int64_t const blocks = std::min<int64_t>((G + BLOCK / 64 - 1) / (BLOCK / 64), 4096);
  hipLaunchKernelGGL(k_sorted_nth_scan, dim3(unsigned(blocks)), dim3(BLOCK), 0, s, dP, f, group_last, large);
}

void launch_sorted_nunique(gb_params const* dP, int32_t f, int64_t n, int64_t G, uint32_t const* gid,
                           uint32_t* table, hipStream_t s)
{
  if (n <= 0 || G <= 0) return;
  auto const grid = [](int64_t m) { return dim3(unsigned(std::min<int64_t>((m + BLOCK - 1) / BLOCK, 1 << 20))); };
  hipLaunchKernelGGL(k_fill_empty, grid(2 * n), dim3(BLOCK), 0, s, table, 2 * n);
  hipLaunchKernelGGL(k_sorted_nunique, grid(n), dim3(BLOCK), 0, s, dP, f, n, G, gid, table);
}

// String MIN / MAX from the groups' ARGMIN / ARGMAX rows: lengths (0 for
// groups without a valid value) and validity words.
__global__ void __launch_bounds__(BLOCK)
  k_str_arg_sizes(dcol src, int32_t const* __restrict__ rows, int64_t G, int32_t* __restrict__ sizes,
                  uint32_t* __restrict__ mask)
{
  int64_t const g = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  bool valid      = false;
  if (g < G) {
    int32_t const r = rows[g];
    valid           = r != -1 && dcol_valid(src, r);  // NTH_ELEMENT's row may hold a null
    int64_t b = 0, e = 0;
    if (valid) str_span(src, r, b, e);
    sizes[g] = int32_t(e - b);
  }
  if (mask) store_valid_bits(mask, g, G, valid);
}

void launch_str_arg_sizes(dcol const& src, int32_t const* rows, int64_t G, int32_t* sizes, uint32_t* mask,
                          hipStream_t s)
{
  if (G > 0)
    hipLaunchKernelGGL(k_str_arg_sizes, dim3((G + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, src, rows, G, sizes, mask);
}

void launch_iota_u32(uint32_t* out, int64_t n, hipStream_t s)
{
  if (n > 0) hipLaunchKernelGGL(k_iota_u32, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, out, n);
}

}  // namespace mig
