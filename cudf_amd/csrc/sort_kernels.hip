// Kernels of the sort-based aggregations (MEDIAN, QUANTILE) on gfx950.
//
// The reference sorts the keys, then sorts each group's values and reads the
// quantiles off the sorted segments (sort/sort_helper.cu, sort/
// group_quantiles.cu:30-80, quantiles/quantiles_util.hpp).  Here the rows are
// labelled with their group by the hash table, and one stable LSD counting
// sort orders them by (group, null, value) using the order-preserving bits
// written by k_quantile_keys; k_group_quantiles then reads every group's
// segment.
#include "device_ops.hpp"
#include "kernels.hpp"

namespace mig {

namespace {

constexpr int SBLOCK = 256;

__device__ __forceinline__ bool is_signed_st(int32_t st)
{
  return st == ST_I8 || st == ST_I16 || st == ST_I32 || st == ST_I64;
}

// Order-preserving bits of element `row` (32 bits for elements of up to 4
// bytes, 64 bits otherwise): signed integers with the sign bit flipped,
// unsigned integers unchanged, floats with NaN above +inf and -0.0 as 0.0.
__device__ __forceinline__ uint64_t order_bits(dcol const& c, int64_t row)
{
  auto const d = gp(c.data);
  switch (c.storage) {
    case ST_I8: return uint32_t(int32_t(reinterpret_cast<gmem<int8_t const>*>(d)[row])) ^ 0x80000000u;
    case ST_I16: return uint32_t(int32_t(reinterpret_cast<gmem<int16_t const>*>(d)[row])) ^ 0x80000000u;
    case ST_I32: return uint32_t(reinterpret_cast<gmem<int32_t const>*>(d)[row]) ^ 0x80000000u;
    case ST_I64: return uint64_t(reinterpret_cast<gmem<int64_t const>*>(d)[row]) ^ (1ull << 63);
    case ST_U8:
    case ST_BOOL: return d[row];
    case ST_U16: return reinterpret_cast<gmem<uint16_t const>*>(d)[row];
    case ST_U32: return reinterpret_cast<gmem<uint32_t const>*>(d)[row];
    case ST_U64: return reinterpret_cast<gmem<uint64_t const>*>(d)[row];
    case ST_F32: {
      float const x = reinterpret_cast<gmem<float const>*>(d)[row];
      if (isnan(x)) return 0xFFFFFFFFu;
      uint32_t const b = __float_as_uint(x == 0.0f ? 0.0f : x);
      return (b >> 31) ? ~b : (b | 0x80000000u);
    }
    default: {  // ST_F64
      double const x = reinterpret_cast<gmem<double const>*>(d)[row];
      if (isnan(x)) return ~0ull;
      uint64_t const b = uint64_t(__double_as_longlong(x == 0.0 ? 0.0 : x));
      return (b >> 63) ? ~b : (b | (1ull << 63));
    }
  }
}

__global__ void __launch_bounds__(SBLOCK)
  k_quantile_keys(dcol v, int64_t n, uint32_t* __restrict__ lo, uint32_t* __restrict__ hi,
                  uint32_t* __restrict__ null_flag)
{
  int64_t const r = int64_t(blockIdx.x) * SBLOCK + threadIdx.x;
  if (r >= n) return;
  bool const valid = dcol_valid(v, r);
  uint64_t const o = valid ? order_bits(v, r) : 0ull;
  gp(lo)[r]        = uint32_t(o);
  if (hi) gp(hi)[r] = uint32_t(o >> 32);
  if (null_flag) gp(null_flag)[r] = valid ? 0u : 1u;
}

// The element of storage `st` whose order bits are (lo, hi), as double.
__device__ __forceinline__ double order_value(int32_t st, uint32_t lo, uint32_t hi)
{
  switch (st) {
    case ST_I8:
    case ST_I16:
    case ST_I32: return double(int32_t(lo ^ 0x80000000u));
    case ST_I64: return double(int64_t(((uint64_t(hi) << 32) | lo) ^ (1ull << 63)));
    case ST_U64: return double((uint64_t(hi) << 32) | lo);
    case ST_F32: {
      uint32_t const b = (lo >> 31) ? (lo & 0x7FFFFFFFu) : ~lo;
      return double(__uint_as_float(b));
    }
    case ST_F64: {
      uint64_t const o = (uint64_t(hi) << 32) | lo;
      uint64_t const b = (o >> 63) ? (o & ~(1ull << 63)) : ~o;
      return __longlong_as_double(int64_t(b));
    }
    default: return double(lo);  // unsigned up to 32 bits, bool
  }
}

__device__ __forceinline__ int64_t order_int64(uint32_t lo, uint32_t hi)
{
  return int64_t(((uint64_t(hi) << 32) | lo) ^ (1ull << 63));
}

// quantile_index and select_quantile_data of quantiles/quantiles_util.hpp:
// v = q * (n - 1) with q clamped to [0, 1]; LINEAR and MIDPOINT in double,
// MIDPOINT of int64 values by halves (no overflow); the arithmetic is not
// contracted into fused multiply-adds.
__global__ void __launch_bounds__(SBLOCK)
  k_group_quantiles(uint32_t const* __restrict__ ord_lo, uint32_t const* __restrict__ ord_hi,
                    uint32_t const* __restrict__ null_flag, int32_t const* __restrict__ off, int64_t G,
                    int32_t storage, double const* __restrict__ q, int32_t nq, int32_t interp,
                    double* __restrict__ out, uint32_t* __restrict__ out_mask)
{
#pragma clang fp contract(off)
  int64_t const g = int64_t(blockIdx.x) * SBLOCK + threadIdx.x;
  if (g >= G) return;
  int64_t const b = gp(off)[g], e = gp(off)[g + 1];
  // valid values come first in the segment (the null flag sorts above them)
  int64_t m = e - b;
  if (null_flag) {
    int64_t lo = b, hi = e;
    while (lo < hi) {
      int64_t const mid = (lo + hi) >> 1;
      if (gp(null_flag)[mid]) hi = mid;
      else lo = mid + 1;
    }
    m = lo - b;
  }
  auto elem = [&](int64_t i) {
    return order_value(storage, gp(ord_lo)[b + i], ord_hi ? gp(ord_hi)[b + i] : 0u);
  };
  for (int32_t j = 0; j < nq; j++) {
    int64_t const o = g * nq + j;
    double r        = 0.0;
    if (m > 0) {
      double qq = gp(q)[j];
      qq        = qq < 0.0 ? 0.0 : (qq > 1.0 ? 1.0 : qq);
      double const v   = qq * double(m - 1);
      int64_t const lo = int64_t(floor(v)), hi = int64_t(ceil(v)), ne = int64_t(rint(v));
      double const f   = v - double(lo);
      switch (interp) {
        case MIG_INTERP_LINEAR: {
          double const a = elem(lo), c = elem(hi);
          double const x = (1.0 - f) * a;
          double const y = f * c;
          r              = x + y;
          break;
        }
        case MIG_INTERP_LOWER: r = elem(lo); break;
        case MIG_INTERP_HIGHER: r = elem(hi); break;
        case MIG_INTERP_NEAREST: r = elem(ne); break;
        default: {  // MIDPOINT
          if (storage == ST_I64) {
            int64_t const a = order_int64(gp(ord_lo)[b + lo], gp(ord_hi)[b + lo]);
            int64_t const c = order_int64(gp(ord_lo)[b + hi], gp(ord_hi)[b + hi]);
            int64_t const half = a / 2 + c / 2, rest = a % 2 + c % 2;
            r = double(half) + double(rest) * 0.5;
          } else {
            r = elem(lo) / 2 + elem(hi) / 2;
          }
          break;
        }
      }
      __hip_atomic_fetch_or(out_mask + (o >> 5), 1u << (o & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    gp(out)[o] = r;
  }
}

__global__ void __launch_bounds__(SBLOCK)
  k_first_rows(uint32_t const* __restrict__ map, int32_t const* __restrict__ off, int64_t G, uint32_t* __restrict__ rep)
{
  int64_t const g = int64_t(blockIdx.x) * SBLOCK + threadIdx.x;
  if (g < G) gp(rep)[g] = gp(map)[gp(off)[g]];
}

// Rows the null policy dropped (NO_GROUP) take group number G.
__global__ void __launch_bounds__(SBLOCK) k_clamp_groups(uint32_t* __restrict__ gid, int64_t n, uint32_t G)
{
  int64_t const r = int64_t(blockIdx.x) * SBLOCK + threadIdx.x;
  if (r < n && gp(gid)[r] > G) gp(gid)[r] = G;
}

int64_t sgrid(int64_t n) { return (n + SBLOCK - 1) / SBLOCK; }

}  // namespace

void launch_quantile_keys(dcol const& v, int64_t n, uint32_t* lo, uint32_t* hi, uint32_t* null_flag, hipStream_t s)
{
  if (n > 0) hipLaunchKernelGGL(k_quantile_keys, dim3(sgrid(n)), dim3(SBLOCK), 0, s, v, n, lo, hi, null_flag);
}

void launch_group_quantiles(uint32_t const* ord_lo, uint32_t const* ord_hi, uint32_t const* null_flag,
                            int32_t const* off, int64_t G, int32_t storage, double const* q, int32_t nq,
                            int32_t interp, double* out, uint32_t* out_mask, hipStream_t s)
{
  if (G > 0)
    hipLaunchKernelGGL(k_group_quantiles, dim3(sgrid(G)), dim3(SBLOCK), 0, s, ord_lo, ord_hi, null_flag, off, G,
                       storage, q, nq, interp, out, out_mask);
}

void launch_clamp_groups(uint32_t* gid, int64_t n, uint32_t G, hipStream_t s)
{
  if (n > 0) hipLaunchKernelGGL(k_clamp_groups, dim3(sgrid(n)), dim3(SBLOCK), 0, s, gid, n, G);
}

void launch_first_rows(uint32_t const* map, int32_t const* off, int64_t G, uint32_t* rep, hipStream_t s)
{
  if (G > 0) hipLaunchKernelGGL(k_first_rows, dim3(sgrid(G)), dim3(SBLOCK), 0, s, map, off, G, rep);
}

}  // namespace mig
