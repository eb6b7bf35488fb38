// Device-side helpers shared by the group-by kernels: agent-scope atomics,
// order-preserving float encodings, element loads, MurmurHash3 row hashing
// (cudf/table/experimental/row_operators.cuh:1858-1883) and row equality
// (row_operators.cuh:1352-1364, 1429-1448: nulls equal, NaN equal NaN,
// -0.0 equal 0.0).
#pragma once

#include "kernels.hpp"
#include "murmur.hpp"

namespace mig {

// ---------------------------------------------------------------------------
// atomics (agent scope: the table is shared by all XCDs of the device)
template <typename T>
__device__ __forceinline__ T aload(T const* p)
{
  return __hip_atomic_load(gp(const_cast<T*>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void astore(T* p, T v)
{
  __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t acas64(uint64_t* p, uint64_t expected, uint64_t desired)
{
  __hip_atomic_compare_exchange_strong(gp(p), &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;
}
template <typename T>
__device__ __forceinline__ void aadd(T* p, T v)
{
  __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void amin(T* p, T v)
{
  __hip_atomic_fetch_min(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void amax(T* p, T v)
{
  __hip_atomic_fetch_max(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void gadd_f32(float* p, float v)
{
  __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gadd_f64(double* p, double v)
{
  __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Order-preserving unsigned encodings of float bit patterns.
__device__ __forceinline__ uint32_t ord32(uint32_t b) { return (b & 0x80000000u) ? ~b : (b | 0x80000000u); }
__device__ __forceinline__ uint32_t unord32(uint32_t o) { return (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o; }
__device__ __forceinline__ uint64_t ord64(uint64_t b)
{
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ uint64_t unord64(uint64_t o)
{
  return (o & 0x8000000000000000ull) ? (o & 0x7fffffffffffffffull) : ~o;
}

// ---------------------------------------------------------------------------
// element access
__device__ __forceinline__ uint64_t load_fixed_bits(dcol const& c, int64_t row)
{
  auto const d = gp(c.data);
  switch (c.storage) {
    case ST_I8:
    case ST_U8: return d[row];
    case ST_BOOL: return d[row] != 0;
    case ST_I16:
    case ST_U16: return reinterpret_cast<gmem<uint16_t const>*>(d)[row];
    case ST_I32:
    case ST_U32: return reinterpret_cast<gmem<uint32_t const>*>(d)[row];
    case ST_F32: return normalize_f32_bits(reinterpret_cast<gmem<uint32_t const>*>(d)[row]);
    case ST_I64:
    case ST_U64: return reinterpret_cast<gmem<uint64_t const>*>(d)[row];
    case ST_F64: return normalize_f64_bits(reinterpret_cast<gmem<uint64_t const>*>(d)[row]);
    default: return 0;
  }
}

__device__ __forceinline__ uint32_t hash_fixed_bits(int32_t st, uint64_t bits, uint32_t seed)
{
  switch (storage_width(st)) {
    case 1: return mm_hash_small(static_cast<uint32_t>(bits), 1, seed);
    case 2: return mm_hash_small(static_cast<uint32_t>(bits), 2, seed);
    case 4: return mm_hash4(static_cast<uint32_t>(bits), seed);
    default: return mm_hash8(bits, seed);
  }
}

__device__ __forceinline__ void str_span(dcol const& c, int64_t row, int64_t& b, int64_t& e)
{
  if (c.offsets64) {
    auto o = gp(reinterpret_cast<int64_t const*>(c.offsets));
    b = o[row];
    e = o[row + 1];
  } else {
    auto o = gp(reinterpret_cast<int32_t const*>(c.offsets));
    b = o[row];
    e = o[row + 1];
  }
}

__device__ __forceinline__ uint32_t elem_hash(dcol const& c, int64_t row, uint32_t seed)
{
  if (!dcol_valid(c, row)) return NULL_HASH;
  if (c.storage == ST_STRING) {
    int64_t b, e;
    str_span(c, row, b, e);
    return mm_hash_bytes(c.data + b, e - b, seed);
  }
  return hash_fixed_bits(c.storage, load_fixed_bits(c, row), seed);
}

__device__ __forceinline__ uint32_t row_hash(dcol const* cols, int ncols, int64_t row, uint32_t seed)
{
  uint32_t h = seed;
  for (int c = 0; c < ncols; c++) h = hash_combine(h, elem_hash(cols[c], row, seed));
  return h;
}

// row_hash(cols, ncols, row, 0) and row_hash(cols, ncols, row, s2) together,
// each element read once.
__device__ __forceinline__ void row_hash2(dcol const* cols, int ncols, int64_t row, uint32_t s2, uint32_t& h1,
                                          uint32_t& h2)
{
  h1 = 0u;
  h2 = s2;
  for (int c = 0; c < ncols; c++) {
    dcol const& col = cols[c];
    uint32_t e1 = NULL_HASH, e2 = NULL_HASH;
    if (dcol_valid(col, row)) {
      if (col.storage == ST_STRING) {
        int64_t b, e;
        str_span(col, row, b, e);
        mm_hash_bytes2(col.data + b, e - b, 0u, s2, e1, e2);
      } else {
        uint64_t const bits = load_fixed_bits(col, row);
        e1 = hash_fixed_bits(col.storage, bits, 0u);
        e2 = hash_fixed_bits(col.storage, bits, s2);
      }
    }
    h1 = hash_combine(h1, e1);
    h2 = hash_combine(h2, e2);
  }
}

__device__ __forceinline__ bool elem_equal(dcol const& c, int64_t a, int64_t b)
{
  bool const va = dcol_valid(c, a), vb = dcol_valid(c, b);
  if (!va || !vb) return va == vb;
  if (c.storage == ST_STRING) {
    int64_t ab, ae, bb, be;
    str_span(c, a, ab, ae);
    str_span(c, b, bb, be);
    int64_t const len = ae - ab;
    if (len != be - bb) return false;
    // 8 bytes of each string per step, their loads in flight together (a
    // byte-by-byte loop with an exit per byte waits for every load in turn)
    auto const d = gp(c.data);
    for (int64_t i = 0; i < len; i += 8) {
      uint32_t diff = 0;
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (i + j < len) diff |= uint32_t(d[ab + i + j] ^ d[bb + i + j]);
      if (diff) return false;
    }
    return true;
  }
  return load_fixed_bits(c, a) == load_fixed_bits(c, b);
}

__device__ __forceinline__ bool rows_equal(gb_params const& P, int64_t a, int64_t b)
{
  for (int c = 0; c < P.num_keys; c++)
    if (!elem_equal(P.keys[c], a, b)) return false;
  return true;
}

__device__ __forceinline__ bool any_key_null(gb_params const& P, int64_t row)
{
  for (int c = 0; c < P.num_keys; c++)
    if (!dcol_valid(P.keys[c], row)) return true;
  return false;
}


__host__ __device__ inline uint64_t field_identity(int32_t kind)
{
  switch (kind) {
    case FK_MIN_S32: return 0x7fffffffull;
    case FK_MAX_S32: return 0x80000000ull;
    case FK_MIN_S64: return 0x7fffffffffffffffull;
    case FK_MAX_S64: return 0x8000000000000000ull;
    case FK_MIN_U32: return 0xffffffffull;
    case FK_MIN_U64: return 0xffffffffffffffffull;
    case FK_MIN_F32: return 0xff800000ull;              // ord32(+inf)
    case FK_MAX_F32: return 0x007fffffull;              // ord32(-inf)
    case FK_MIN_F64: return 0xfff0000000000000ull;      // ord64(+inf)
    case FK_MAX_F64: return 0x000fffffffffffffull;      // ord64(-inf)
    case FK_ARGMIN:
    case FK_ARGMAX: return 0xffffffffull;               // no row yet
    case FK_PROD_I64: return 1ull;
    case FK_PROD_F32: return 0x3f800000ull;             // 1.0f
    case FK_PROD_F64: return 0x3ff0000000000000ull;     // 1.0
    default: return 0;                                  // sums, counts, max of unsigned
  }
}

__device__ __forceinline__ bool is_unsigned_storage_dev(int32_t st)
{
  return st == ST_U8 || st == ST_U16 || st == ST_U32 || st == ST_U64 || st == ST_BOOL;
}

struct loaded_value {
  int64_t i;   // integer sources: sign- or zero-extended
  double d;    // value as double (all sources)
  float f;     // float32 sources
  uint64_t b;  // float sources: raw bits
};

__device__ __forceinline__ loaded_value load_value(dcol const& c, int64_t row)
{
  loaded_value v{0, 0.0, 0.0f, 0};
  auto const d = gp(c.data);
  switch (c.storage) {
    case ST_I8: v.i = reinterpret_cast<gmem<int8_t const>*>(d)[row]; v.d = double(v.i); break;
    case ST_I16: v.i = reinterpret_cast<gmem<int16_t const>*>(d)[row]; v.d = double(v.i); break;
    case ST_I32: v.i = reinterpret_cast<gmem<int32_t const>*>(d)[row]; v.d = double(v.i); break;
    case ST_I64: v.i = reinterpret_cast<gmem<int64_t const>*>(d)[row]; v.d = double(v.i); break;
    case ST_U8: v.i = d[row]; v.d = double(v.i); break;
    case ST_BOOL: v.i = d[row] != 0; v.d = double(v.i); break;
    case ST_U16: v.i = reinterpret_cast<gmem<uint16_t const>*>(d)[row]; v.d = double(v.i); break;
    case ST_U32: v.i = reinterpret_cast<gmem<uint32_t const>*>(d)[row]; v.d = double(v.i); break;
    case ST_U64: {
      uint64_t u = reinterpret_cast<gmem<uint64_t const>*>(d)[row];
      v.i = int64_t(u);
      v.d = double(u);
      break;
    }
    case ST_F32: {
      uint32_t u = reinterpret_cast<gmem<uint32_t const>*>(d)[row];
      v.b = u;
      v.f = __uint_as_float(u);
      v.d = double(v.f);
      break;
    }
    case ST_F64: {
      uint64_t u = reinterpret_cast<gmem<uint64_t const>*>(d)[row];
      v.b = u;
      v.d = __longlong_as_double(int64_t(u));
      break;
    }
    default: break;
  }
  return v;
}

// Accumulate one valid value into the accumulator at `a` (an element of the
// kind's width; agent-scope atomics; FK_M2 is handled by a second pass).
__device__ __forceinline__ void acc_apply(int32_t kind, void* a, loaded_value const& v)
{
  switch (kind) {
    case FK_SUM_I64: aadd(static_cast<uint64_t*>(a), uint64_t(v.i)); break;
    case FK_SUM_F32: gadd_f32(static_cast<float*>(a), v.f); break;
    case FK_SUM_F64: gadd_f64(static_cast<double*>(a), v.d); break;
    case FK_SUMSQ_I64: aadd(static_cast<uint64_t*>(a), uint64_t(v.i) * uint64_t(v.i)); break;
    case FK_SUMSQ_F32: gadd_f32(static_cast<float*>(a), v.f * v.f); break;
    case FK_SUMSQ_F64: gadd_f64(static_cast<double*>(a), v.d * v.d); break;
    case FK_MIN_S32: amin(static_cast<int32_t*>(a), int32_t(v.i)); break;
    case FK_MAX_S32: amax(static_cast<int32_t*>(a), int32_t(v.i)); break;
    case FK_MIN_S64: amin(static_cast<int64_t*>(a), v.i); break;
    case FK_MAX_S64: amax(static_cast<int64_t*>(a), v.i); break;
    case FK_MIN_U32: amin(static_cast<uint32_t*>(a), uint32_t(v.i)); break;
    case FK_MAX_U32: amax(static_cast<uint32_t*>(a), uint32_t(v.i)); break;
    case FK_MIN_U64: amin(static_cast<uint64_t*>(a), uint64_t(v.i)); break;
    case FK_MAX_U64: amax(static_cast<uint64_t*>(a), uint64_t(v.i)); break;
    case FK_MIN_F32:
      if (!isnan(v.f)) amin(static_cast<uint32_t*>(a), ord32(uint32_t(v.b)));
      break;
    case FK_MAX_F32:
      if (!isnan(v.f)) amax(static_cast<uint32_t*>(a), ord32(uint32_t(v.b)));
      break;
    case FK_MIN_F64:
      if (!isnan(v.d)) amin(static_cast<uint64_t*>(a), ord64(v.b));
      break;
    case FK_MAX_F64:
      if (!isnan(v.d)) amax(static_cast<uint64_t*>(a), ord64(v.b));
      break;
    case FK_COUNT: aadd(static_cast<uint32_t*>(a), 1u); break;
    case FK_PROD_I64:
    case FK_PROD_F64: {
      // no atomic multiply: compare-and-swap loop (DeviceProduct through
      // atomicCAS in the reference's device_atomics)
      auto* p      = static_cast<uint64_t*>(a);
      uint64_t cur = __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (;;) {
        uint64_t const next =
          kind == FK_PROD_I64 ? cur * uint64_t(v.i)
                              : uint64_t(__double_as_longlong(__longlong_as_double(int64_t(cur)) * v.d));
        uint64_t const seen = acas64(p, cur, next);
        if (seen == cur) break;
        cur = seen;
      }
      break;
    }
    case FK_PROD_F32: {
      auto* p      = static_cast<unsigned int*>(a);
      uint32_t cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (;;) {
        uint32_t const next = __float_as_uint(__uint_as_float(cur) * v.f);
        uint32_t const seen = atomicCAS(p, cur, next);
        if (seen == cur) break;
        cur = seen;
      }
      break;
    }
    default: break;
  }
}

// Global accumulator of field F at `slot`, and a slot's row count.
__device__ __forceinline__ uint8_t* field_at(field_desc const& F, int64_t slot)
{
  return static_cast<uint8_t*>(F.ptr) + slot * F.stride;
}
__device__ __forceinline__ uint32_t* count_at(gb_params const& P, int64_t slot)
{
  return reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(P.slot_count) + slot * P.count_stride);
}

// Accumulate into slot `slot` of a field's global accumulators.
__device__ __forceinline__ void apply_field(field_desc const& F, int64_t slot, loaded_value const& v)
{
  acc_apply(F.kind, field_at(F, slot), v);
}

// Combine a pre-aggregated partial (same encoding as the table) into `a`.
__device__ __forceinline__ void acc_combine(int32_t kind, void* a, uint64_t part)
{
  switch (kind) {
    case FK_SUM_I64:
    case FK_SUMSQ_I64: aadd(static_cast<uint64_t*>(a), part); break;
    case FK_SUM_F32:
    case FK_SUMSQ_F32: gadd_f32(static_cast<float*>(a), __uint_as_float(uint32_t(part))); break;
    case FK_SUM_F64:
    case FK_SUMSQ_F64: gadd_f64(static_cast<double*>(a), __longlong_as_double(int64_t(part))); break;
    case FK_MIN_S32: amin(static_cast<int32_t*>(a), int32_t(part)); break;
    case FK_MAX_S32: amax(static_cast<int32_t*>(a), int32_t(part)); break;
    case FK_MIN_S64: amin(static_cast<int64_t*>(a), int64_t(part)); break;
    case FK_MAX_S64: amax(static_cast<int64_t*>(a), int64_t(part)); break;
    case FK_MIN_U32:
    case FK_MIN_F32: amin(static_cast<uint32_t*>(a), uint32_t(part)); break;
    case FK_MAX_U32:
    case FK_MAX_F32: amax(static_cast<uint32_t*>(a), uint32_t(part)); break;
    case FK_MIN_U64:
    case FK_MIN_F64: amin(static_cast<uint64_t*>(a), part); break;
    case FK_MAX_U64:
    case FK_MAX_F64: amax(static_cast<uint64_t*>(a), part); break;
    case FK_COUNT: aadd(static_cast<uint32_t*>(a), uint32_t(part)); break;
    case FK_PROD_I64:
    case FK_PROD_F32:
    case FK_PROD_F64: {
      // multiply by the partial with the row update's compare-and-swap loop
      loaded_value v{int64_t(part), __longlong_as_double(int64_t(part)), __uint_as_float(uint32_t(part)), part};
      acc_apply(kind, a, v);
      break;
    }
    default: break;
  }
}

// The per-row update into LDS partials (workgroup-scope atomics).
__device__ __forceinline__ void lds_apply(int32_t kind, void* a, loaded_value const& v)
{
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
  switch (kind) {
    case FK_SUM_I64: __hip_atomic_fetch_add(static_cast<uint64_t*>(a), uint64_t(v.i), __ATOMIC_RELAXED, WG); break;
    case FK_SUM_F32: __hip_atomic_fetch_add(static_cast<float*>(a), v.f, __ATOMIC_RELAXED, WG); break;
    case FK_SUM_F64: __hip_atomic_fetch_add(static_cast<double*>(a), v.d, __ATOMIC_RELAXED, WG); break;
    case FK_SUMSQ_I64:
      __hip_atomic_fetch_add(static_cast<uint64_t*>(a), uint64_t(v.i) * uint64_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_SUMSQ_F32: __hip_atomic_fetch_add(static_cast<float*>(a), v.f * v.f, __ATOMIC_RELAXED, WG); break;
    case FK_SUMSQ_F64: __hip_atomic_fetch_add(static_cast<double*>(a), v.d * v.d, __ATOMIC_RELAXED, WG); break;
    case FK_MIN_S32: __hip_atomic_fetch_min(static_cast<int32_t*>(a), int32_t(v.i), __ATOMIC_RELAXED, WG); break;
    case FK_MAX_S32: __hip_atomic_fetch_max(static_cast<int32_t*>(a), int32_t(v.i), __ATOMIC_RELAXED, WG); break;
    case FK_MIN_S64: __hip_atomic_fetch_min(static_cast<int64_t*>(a), v.i, __ATOMIC_RELAXED, WG); break;
    case FK_MAX_S64: __hip_atomic_fetch_max(static_cast<int64_t*>(a), v.i, __ATOMIC_RELAXED, WG); break;
    case FK_MIN_U32: __hip_atomic_fetch_min(static_cast<uint32_t*>(a), uint32_t(v.i), __ATOMIC_RELAXED, WG); break;
    case FK_MAX_U32: __hip_atomic_fetch_max(static_cast<uint32_t*>(a), uint32_t(v.i), __ATOMIC_RELAXED, WG); break;
    case FK_MIN_U64: __hip_atomic_fetch_min(static_cast<uint64_t*>(a), uint64_t(v.i), __ATOMIC_RELAXED, WG); break;
    case FK_MAX_U64: __hip_atomic_fetch_max(static_cast<uint64_t*>(a), uint64_t(v.i), __ATOMIC_RELAXED, WG); break;
    case FK_MIN_F32:
      if (!isnan(v.f)) __hip_atomic_fetch_min(static_cast<uint32_t*>(a), ord32(uint32_t(v.b)), __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_F32:
      if (!isnan(v.f)) __hip_atomic_fetch_max(static_cast<uint32_t*>(a), ord32(uint32_t(v.b)), __ATOMIC_RELAXED, WG);
      break;
    case FK_MIN_F64:
      if (!isnan(v.d)) __hip_atomic_fetch_min(static_cast<uint64_t*>(a), ord64(v.b), __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_F64:
      if (!isnan(v.d)) __hip_atomic_fetch_max(static_cast<uint64_t*>(a), ord64(v.b), __ATOMIC_RELAXED, WG);
      break;
    case FK_COUNT: __hip_atomic_fetch_add(static_cast<uint32_t*>(a), 1u, __ATOMIC_RELAXED, WG); break;
    default: break;
  }
}

// Insert `key` into an LDS open-addressing table of S (a power of two) u64
// slots at `tkeys` whose free slots hold `empty`, with at most `limit` keys
// (`fill` counts them); returns its slot, or -1 when the table is full.
__device__ __forceinline__ int lds_insert(uint64_t* tkeys, uint32_t* fill, uint64_t key, uint32_t h,
                                          uint64_t empty, int S, uint32_t limit)
{
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
  uint32_t s       = h & uint32_t(S - 1);
  for (int p = 0; p < S; p++) {
    uint64_t const cur = __hip_atomic_load(tkeys + s, __ATOMIC_RELAXED, WG);
    if (cur == key) return int(s);
    if (cur == empty) {
      if (__hip_atomic_load(fill, __ATOMIC_RELAXED, WG) >= limit) return -1;
      uint64_t expected = empty;
      if (__hip_atomic_compare_exchange_strong(tkeys + s, &expected, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               WG)) {
        __hip_atomic_fetch_add(fill, 1u, __ATOMIC_RELAXED, WG);
        return int(s);
      }
      if (expected == key) return int(s);
    }
    s = (s + 1) & uint32_t(S - 1);
  }
  return -1;
}

__device__ __forceinline__ void store_width(void* data, int64_t g, int w, uint64_t bits)
{
  auto const d = gp(static_cast<uint8_t*>(data));
  switch (w) {
    case 1: d[g] = uint8_t(bits); break;
    case 2: reinterpret_cast<gmem<uint16_t>*>(d)[g] = uint16_t(bits); break;
    case 4: reinterpret_cast<gmem<uint32_t>*>(d)[g] = uint32_t(bits); break;
    default: reinterpret_cast<gmem<uint64_t>*>(d)[g] = bits; break;
  }
}

// Sum over a block of NT threads (wave64 shuffles, then across waves through
// `lds`, which holds NT / 64 words); the total is returned to thread 0.
template <int NT = 256>
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t* lds)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  uint32_t t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NT / 64; w++) t += lds[w];
  __syncthreads();
  return t;
}

// Address of element `i` of width `w` in an LDS byte array.
template <typename T>
__device__ __forceinline__ T* lds_at(uint8_t* lds, int32_t off, int64_t i)
{
  return reinterpret_cast<T*>(lds + off) + i;
}

}  // namespace mig
