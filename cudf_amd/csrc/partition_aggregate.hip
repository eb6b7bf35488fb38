// k_rpa_aggregate: one workgroup aggregates one final bin of the radix
// partition entirely in LDS.
//
//   pass A  every row of the bin: find-or-insert its key in an LDS open-
//           addressing table of S slots (CAS on 64-bit LDS words), then apply
//           every accumulator of the request at the slot with LDS atomics.
//           Accumulators are indexed by slot, so a row never waits for
//           another thread's insert to publish anything.
//   pass B  VARIANCE / STD: squared deviations from the group mean, with the
//           slots of the first rows of every thread kept in registers and the
//           values re-read from L2.
//   output  occupied slots are compacted by a block scan; consecutive lanes
//           write consecutive output rows (coalesced) and validity words are
//           assembled by wave ballots.
//
// Replaces the per-row device-wide atomics of the reference
// (groupby/hash/groupby_kernels.cuh:98-105, detail/aggregation/
// aggregation.cuh:131-490) and its second VARIANCE pass over the whole input
// (multi_pass_kernels.cuh:74-111), with the same results per group.
#include "partition_common.hpp"

namespace mig {

constexpr int WG_SCOPE = __HIP_MEMORY_SCOPE_WORKGROUP;

constexpr int RPA_CACHED = 4;     // rows per thread per chunk; their slots stay in registers
constexpr int RPA_OUT_BATCH = 4;  // output columns whose LDS reads are issued together
// Probes before a row's key counts as not fitting: the table is filled to at
// most ~85%, where a double-hashing probe longer than this has probability
// below 1e-18, so reaching it means the bin has more groups than the table
// holds (the call is retried with more bins).
constexpr int RPA_PROBE_LIMIT = 256;

// Apply `f(k)` to the rows k of a chunk whose bit is set in `ok` (the loop is
// unrolled; the caller's switch on the accumulator kind stays outside it).
template <typename F>
__device__ __forceinline__ void for_rows(uint32_t ok, F&& f)
{
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++)
    if ((ok >> k) & 1u) f(k);
}

// One accumulator over the rows of a chunk: `c[k]` holds row k's value in
// canonical form (integers sign- or zero-extended to 64 bits, float bits
// unchanged), `slot[k]` its group slot, bit k of `ok` whether it counts.
__device__ __forceinline__ double canon_double(int32_t st, uint64_t c);

__device__ __forceinline__ void acc_rows_lds(int32_t kind, int32_t lds_off, int32_t st, uint8_t* lds,
                                             int const (&slot)[RPA_CACHED], uint32_t ok,
                                             uint64_t const (&c)[RPA_CACHED], bool wide_f32)
{
  uint8_t* const base = lds + lds_off;
  auto u64 = [&](int k) { return reinterpret_cast<uint64_t*>(base) + slot[k]; };
  auto u32 = [&](int k) { return reinterpret_cast<uint32_t*>(base) + slot[k]; };
  auto fl  = [&](int k) { return __uint_as_float(uint32_t(c[k])); };
  auto db  = [&](int k) { return __longlong_as_double(int64_t(c[k])); };
  constexpr int R = __ATOMIC_RELAXED;
  switch (kind) {
    case FK_SUM_I64: for_rows(ok, [&](int k) { __hip_atomic_fetch_add(u64(k), c[k], R, WG_SCOPE); }); break;
    case FK_SUM_F32:  // in float64 when wide (rounded to float32 at output)
      if (wide_f32)
        for_rows(ok, [&](int k) {
          __hip_atomic_fetch_add(reinterpret_cast<double*>(u64(k)), double(fl(k)), R, WG_SCOPE);
        });
      else
        for_rows(ok, [&](int k) { __hip_atomic_fetch_add(reinterpret_cast<float*>(u32(k)), fl(k), R, WG_SCOPE); });
      break;
    case FK_SUM_F64:
      for_rows(ok, [&](int k) { __hip_atomic_fetch_add(reinterpret_cast<double*>(u64(k)), db(k), R, WG_SCOPE); });
      break;
    case FK_SUMSQ_I64:
      for_rows(ok, [&](int k) { __hip_atomic_fetch_add(u64(k), c[k] * c[k], R, WG_SCOPE); });
      break;
    case FK_SUMSQ_F32:
      if (wide_f32)
        for_rows(ok, [&](int k) {
          double const x = double(fl(k));
          __hip_atomic_fetch_add(reinterpret_cast<double*>(u64(k)), x * x, R, WG_SCOPE);
        });
      else
        for_rows(ok, [&](int k) {
          __hip_atomic_fetch_add(reinterpret_cast<float*>(u32(k)), fl(k) * fl(k), R, WG_SCOPE);
        });
      break;
    case FK_SUMSQ_F64:
      for_rows(ok, [&](int k) {
        __hip_atomic_fetch_add(reinterpret_cast<double*>(u64(k)), db(k) * db(k), R, WG_SCOPE);
      });
      break;
    case FK_MIN_S32:
      for_rows(ok, [&](int k) {
        __hip_atomic_fetch_min(reinterpret_cast<int32_t*>(u32(k)), int32_t(c[k]), R, WG_SCOPE);
      });
      break;
    case FK_MAX_S32:
      for_rows(ok, [&](int k) {
        __hip_atomic_fetch_max(reinterpret_cast<int32_t*>(u32(k)), int32_t(c[k]), R, WG_SCOPE);
      });
      break;
    case FK_MIN_S64:
      for_rows(ok, [&](int k) {
        __hip_atomic_fetch_min(reinterpret_cast<int64_t*>(u64(k)), int64_t(c[k]), R, WG_SCOPE);
      });
      break;
    case FK_MAX_S64:
      for_rows(ok, [&](int k) {
        __hip_atomic_fetch_max(reinterpret_cast<int64_t*>(u64(k)), int64_t(c[k]), R, WG_SCOPE);
      });
      break;
    case FK_MIN_U32: for_rows(ok, [&](int k) { __hip_atomic_fetch_min(u32(k), uint32_t(c[k]), R, WG_SCOPE); }); break;
    case FK_MAX_U32: for_rows(ok, [&](int k) { __hip_atomic_fetch_max(u32(k), uint32_t(c[k]), R, WG_SCOPE); }); break;
    case FK_MIN_U64: for_rows(ok, [&](int k) { __hip_atomic_fetch_min(u64(k), c[k], R, WG_SCOPE); }); break;
    case FK_MAX_U64: for_rows(ok, [&](int k) { __hip_atomic_fetch_max(u64(k), c[k], R, WG_SCOPE); }); break;
    case FK_MIN_F32:
      for_rows(ok, [&](int k) {
        if (!isnan(fl(k))) __hip_atomic_fetch_min(u32(k), ord32(uint32_t(c[k])), R, WG_SCOPE);
      });
      break;
    case FK_MAX_F32:
      for_rows(ok, [&](int k) {
        if (!isnan(fl(k))) __hip_atomic_fetch_max(u32(k), ord32(uint32_t(c[k])), R, WG_SCOPE);
      });
      break;
    case FK_MIN_F64:
      for_rows(ok, [&](int k) {
        if (!isnan(db(k))) __hip_atomic_fetch_min(u64(k), ord64(c[k]), R, WG_SCOPE);
      });
      break;
    case FK_MAX_F64:
      for_rows(ok, [&](int k) {
        if (!isnan(db(k))) __hip_atomic_fetch_max(u64(k), ord64(c[k]), R, WG_SCOPE);
      });
      break;
    case FK_COUNT: for_rows(ok, [&](int k) { __hip_atomic_fetch_add(u32(k), 1u, R, WG_SCOPE); }); break;
    case FK_M2:  // pass A: sum of squares in double (finalised into M2 after the pass)
      for_rows(ok, [&](int k) {
        double const x = canon_double(st, c[k]);
        __hip_atomic_fetch_add(reinterpret_cast<double*>(u64(k)), x * x, R, WG_SCOPE);
      });
      break;
    default: break;
  }
}

__device__ __forceinline__ double lds_field_double(field_desc const& F, uint8_t* lds, int id, bool wide_f32)
{
  switch (F.kind) {
    case FK_SUM_I64: return double(*lds_at<int64_t>(lds, F.lds_off, id));
    case FK_SUM_F32: return wide_f32 ? *lds_at<double>(lds, F.lds_off, id) : double(*lds_at<float>(lds, F.lds_off, id));
    case FK_SUM_F64: return *lds_at<double>(lds, F.lds_off, id);
    case FK_COUNT: return double(*lds_at<uint32_t>(lds, F.lds_off, id));
    default: return 0.0;
  }
}

// A sum or count partial, given as its raw bits, as double.
__device__ __forceinline__ double field_bits_double(int32_t kind, uint64_t bits, bool wide_f32)
{
  switch (kind) {
    case FK_SUM_I64: return double(int64_t(bits));
    case FK_SUM_F32: return wide_f32 ? __longlong_as_double(int64_t(bits)) : double(__uint_as_float(uint32_t(bits)));
    case FK_SUM_F64: return __longlong_as_double(int64_t(bits));
    case FK_COUNT: return double(uint32_t(bits));
    default: return 0.0;
  }
}

__device__ __forceinline__ uint32_t lds_count(gb_params const& P, rpa_params const& R, uint8_t* lds,
                                              int32_t f, int id)
{
  return f == -2 ? *lds_at<uint32_t>(lds, R.count_all_off, id)
                 : *lds_at<uint32_t>(lds, P.fields[f].lds_off, id);
}

// Validity bits of the 64 output rows [c0, c0 + 64) (c0 a multiple of 64)
// of a wave whose workgroup owns the rows [base, end): one 64-bit word,
// stored when the workgroup owns all of it, else OR-ed in.  Words at or past
// `cap` (the rows allocated) are left alone.
__device__ __forceinline__ void store_valid_word(uint32_t* mask, uint64_t c0, uint64_t base, uint64_t end,
                                                 uint64_t cap, bool valid)
{
  uint64_t const bal = __ballot(valid);
  if ((threadIdx.x & 63) != 0 || c0 >= cap) return;
  auto const w = gp(reinterpret_cast<uint64_t*>(mask)) + (c0 >> 6);
  if (c0 >= base && c0 + 64 <= end) *w = bal;
  else if (bal) __hip_atomic_fetch_or(w, bal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct rpa_misc {
  uint32_t fill;      // keys inserted into the slot table
  uint32_t present0;  // a null key (INCLUDE) was seen
  uint32_t present1;  // a key equal to EMPTY_KEY was seen
  int32_t overflow;
  uint32_t g_out;
  uint32_t m2_pending;  // some group's M2 needs the exact second pass
  uint32_t rep1;        // hashed general keys: row representing the EMPTY_KEY word group
  unsigned long long base;
  uint32_t wave_sum[RPA_AGG_BLOCK / 64];
};

// Raw elements of a partition value column for the rows [r0 + k * block)
// of a chunk, zero-extended to 64 bits (canon_rows gives the canonical form).
// Rows past the bin read its last row; their lanes take no part.
template <typename T>
__device__ __forceinline__ void load_rows(gmem<T const>* p, int64_t r0, int64_t e, uint64_t (&c)[RPA_CACHED])
{
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) {
    int64_t const r = r0 + int64_t(k) * RPA_AGG_BLOCK;
    c[k]            = uint64_t(p[r < e ? r : e - 1]);
  }
}

__device__ __forceinline__ void load_chunk(gmem<uint8_t const>* src, int w, int64_t r0, int64_t e,
                                           uint64_t (&c)[RPA_CACHED])
{
  switch (w) {
    case 1: load_rows(src, r0, e, c); break;
    case 2: load_rows(reinterpret_cast<gmem<uint16_t const>*>(src), r0, e, c); break;
    case 4: load_rows(reinterpret_cast<gmem<uint32_t const>*>(src), r0, e, c); break;
    default: load_rows(reinterpret_cast<gmem<uint64_t const>*>(src), r0, e, c); break;
  }
}

// Canonical form of a chunk's raw elements of storage `st`: integers sign-
// or zero-extended to 64 bits (bool as 0/1), float bits unchanged.
__device__ __forceinline__ void canon_rows(int32_t st, uint64_t (&c)[RPA_CACHED])
{
  switch (st) {
    case ST_I8:
#pragma unroll
      for (int k = 0; k < RPA_CACHED; k++) c[k] = uint64_t(int64_t(int8_t(c[k])));
      break;
    case ST_I16:
#pragma unroll
      for (int k = 0; k < RPA_CACHED; k++) c[k] = uint64_t(int64_t(int16_t(c[k])));
      break;
    case ST_I32:
#pragma unroll
      for (int k = 0; k < RPA_CACHED; k++) c[k] = uint64_t(int64_t(int32_t(c[k])));
      break;
    case ST_BOOL:
#pragma unroll
      for (int k = 0; k < RPA_CACHED; k++) c[k] = c[k] != 0;
      break;
    default: break;
  }
}

// A canonical element as double.
__device__ __forceinline__ double canon_double(int32_t st, uint64_t c)
{
  if (st == ST_F64) return __longlong_as_double(int64_t(c));
  if (st == ST_F32) return double(__uint_as_float(uint32_t(c)));
  if (st == ST_U64) return double(c);
  return double(int64_t(c));
}

// What pass A needs of the request, held in registers for the whole bin.
// Every load from the kernel parameters is a scalar memory read, and waiting
// for one also waits for all LDS operations in flight (one counter covers
// both), so reading descriptors inside the row loop would drain the LDS
// atomics of every field before the next one starts.  The per-column and
// per-field descriptors sit one per lane of a VGPR and are fetched with
// v_readlane (a uniform index into a register array would go to memory).
struct acc_plan {
  int32_t ncols;
  int32_t nfields;
  bool wide_f32;
  int32_t key_storage;
  int32_t count_all_off;
  uint32_t cd_lane;     // lane i < ncols: storage | (flag bit + 1) << 8 | first field << 16 | end << 24
  uint32_t cp_lo_lane;  // lane i < ncols: partitioned values of column i (low, high word)
  uint32_t cp_hi_lane;
  uint32_t fd_lane;     // lane f: kind | lds_off << 8 of field f
  gmem<uint64_t const>* keys;
  gmem<uint8_t const>* flags;
  gmem<uint32_t const>* rows;
};

__device__ __forceinline__ acc_plan load_plan(gb_params const& P, rpa_params const& R)
{
  acc_plan A;
  A.ncols         = R.plan_ncols;
  A.nfields       = P.num_fields;
  A.wide_f32      = R.wide_f32 != 0;
  A.key_storage   = P.key_storage;
  A.count_all_off = R.count_all_off;
  A.keys          = gp(static_cast<uint64_t const*>(R.keysB));
  A.flags         = gp(static_cast<uint8_t const*>(R.flagsB));
  A.rows          = gp(static_cast<uint32_t const*>(R.rowsB));
  int const lane  = int(threadIdx.x & 63);
  int const i     = lane < A.ncols ? lane : 0;
  int const v     = R.plan_col[i];
  A.cd_lane       = uint32_t(P.vals[v].storage) | (uint32_t(R.val_flag_bit[v] + 1) << 8) |
              (uint32_t(R.plan_fb[i]) << 16) | (uint32_t(R.plan_fb[i + 1]) << 24);
  uint64_t const cp = reinterpret_cast<uint64_t>(R.valsB[v]);
  A.cp_lo_lane      = uint32_t(cp);
  A.cp_hi_lane      = uint32_t(cp >> 32);
  A.fd_lane         = R.plan_fd[lane < RPA_PLAN_FIELDS ? lane : 0];
  return A;
}

__device__ __forceinline__ uint32_t lane_value(uint32_t v, int i)
{
  return uint32_t(__builtin_amdgcn_readlane(int(v), i));
}

// Raw values of the first RPA_AGG_PREFETCH value columns of a chunk, loaded
// together with the chunk's keys and flags so that the chunk waits for memory
// once, not once per column.
constexpr int RPA_AGG_PREFETCH = 1;
struct chunk_prefetch {
  uint64_t c[RPA_AGG_PREFETCH][RPA_CACHED];
};

// A chunk's records as loaded: keys, flags (byte k = row k) and the
// prefetched value columns.
struct chunk_in {
  uint64_t key[RPA_CACHED];
  uint64_t flp;
  uint32_t row[RPA_CACHED];  // input rows (hashed general keys)
  chunk_prefetch pf;
};

// Key forms of the aggregating kernel.
constexpr int KF_INLINE = 0;   // one fixed-width key, normalised bits in the table
constexpr int KF_GENERAL = 1;  // general keys: (row hash, row) words, key rows compared while probing
constexpr int KF_HASHED = 2;   // general keys: (row hash, second hash) words, pairs verified afterwards

__device__ __forceinline__ gmem<uint8_t const>* plan_column(acc_plan const& A, int i)
{
  return reinterpret_cast<gmem<uint8_t const>*>((uint64_t(lane_value(A.cp_hi_lane, i)) << 32) |
                                                lane_value(A.cp_lo_lane, i));
}

__device__ __forceinline__ void prefetch_values(acc_plan const& A, int64_t r0, int64_t e, chunk_prefetch& pf)
{
#pragma unroll
  for (int i = 0; i < RPA_AGG_PREFETCH; i++) {
    if (i >= A.ncols) break;
    uint32_t const cd = lane_value(A.cd_lane, i);
    load_chunk(plan_column(A, i), storage_width(int(cd & 0xffu)), r0, e, pf.c[i]);
  }
}


// Slots of the rows [r0 + tid + k * block) of one chunk: S = null key,
// S + 1 = key equal to EMPTY_KEY, else the row's LDS table slot (inserted
// when INSERT); -1 when the row is past the bin or its key is not found.
// All keys and flags are loaded first (RPA_CACHED loads in flight), then the
// rows are probed one after the other.  The table is open-addressed with
// double hashing: the first slot takes the low bits of the row hash and the
// step (odd, so every slot is reached) the next ones, which keeps probe
// sequences short at the table's ~60% fill where linear probing clusters.
// Inserted keys are counted once per chunk; a bin whose count passes
// max_fill is retried with more bins.
// Whether two general-key table words (row hash, row) are the same group:
// equal hashes and equal key rows (nulls equal, NaN equal NaN, -0.0 equal
// 0.0, strings byte-wise).
__device__ __forceinline__ bool same_general(gb_params const& P, uint64_t a, uint64_t b)
{
  return (a >> 32) == (b >> 32) && rows_equal(P, int64_t(uint32_t(a)), int64_t(uint32_t(b)));
}

// Issues every load of a chunk (keys, flags and, with PREFETCH, the first
// value columns) so that the chunk waits for memory once.
template <bool PREFETCH, int KM>
__device__ __forceinline__ void chunk_load(acc_plan const& A, int64_t r0, int64_t e, chunk_in& in)
{
  uint8_t fl[RPA_CACHED];
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) {  // rows past the bin read its last row (e > r0 - tid)
    int64_t const r  = r0 + int64_t(k) * RPA_AGG_BLOCK;
    int64_t const rc = r < e ? r : e - 1;
    fl[k]            = A.flags[rc];
    in.key[k]        = A.keys[rc];
    if (KM == KF_HASHED) in.row[k] = A.rows[rc];
  }
  if (PREFETCH) prefetch_values(A, r0, e, in.pf);
  in.flp = 0;
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) in.flp |= uint64_t(fl[k]) << (8 * k);
}

template <bool INSERT, int KM>
__device__ __forceinline__ void chunk_slots(gb_params const& P, acc_plan const& A, rpa_params const& R,
                                            uint64_t* tkeys, rpa_misc* misc, int S, int64_t r0, int64_t e,
                                            chunk_in const& in, int (&slot)[RPA_CACHED], int64_t rbase)
{
  constexpr bool GENERAL = KM != KF_INLINE;
  uint64_t key[RPA_CACHED];
  uint8_t fl[RPA_CACHED];
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) {
    key[k] = in.key[k];
    fl[k]  = uint8_t(in.flp >> (8 * k));
  }
  uint32_t pending = 0;
  uint32_t reps    = 0;  // records that became their group's representative
  uint32_t pos[RPA_CACHED], step[RPA_CACHED];
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) {
    int64_t const r = r0 + int64_t(k) * RPA_AGG_BLOCK;
    slot[k]         = -1;
    pos[k]          = 0;
    step[k]         = 1;
    if (r >= e) continue;
    if (fl[k] & RPA_KEY_NULL) {
      if (INSERT) misc->present0 = 1;
      slot[k] = S;
    } else if (key[k] == EMPTY_KEY) {
      if (INSERT) misc->present1 = 1;
      slot[k] = S + 1;
      if (KM == KF_HASHED && INSERT) {  // the first such row represents the group
        uint32_t expected = 0xffffffffu;
        if (__hip_atomic_compare_exchange_strong(&misc->rep1, &expected, in.row[k], __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, WG_SCOPE)) {
          __hip_atomic_store(gp(R.rep_rows) + rbase + S + 1, in.row[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          reps |= 1u << k;
        }
      }
    } else {
      uint32_t const h = GENERAL ? uint32_t(key[k] >> 32) : inline_key_hash(A.key_storage, key[k]);
      pos[k]           = h & uint32_t(S - 1);
      step[k]          = ((h >> 9) | 1u) & uint32_t(S - 1);
      pending |= 1u << k;
    }
  }
  uint32_t inserted = 0;
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) {
    if (!((pending >> k) & 1u)) continue;
    uint32_t p      = pos[k];
    int const limit = S < RPA_PROBE_LIMIT ? S : RPA_PROBE_LIMIT;
    int i           = 0;
    for (; i < limit; i++) {
      uint64_t const cur = __hip_atomic_load(tkeys + p, __ATOMIC_RELAXED, WG_SCOPE);
      if (cur == key[k]) break;
      if (KM == KF_GENERAL && cur != EMPTY_KEY && same_general(P, cur, key[k])) break;
      if (cur == EMPTY_KEY) {
        if (!INSERT) {
          i = limit;
          break;
        }
        uint64_t expected = EMPTY_KEY;
        if (__hip_atomic_compare_exchange_strong(tkeys + p, &expected, key[k], __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, WG_SCOPE)) {
          inserted++;
          if (KM == KF_HASHED) {  // this row represents the new group
            __hip_atomic_store(gp(R.rep_rows) + rbase + p, in.row[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            reps |= 1u << k;
          }
          break;
        }
        if (expected == key[k]) break;
        if (KM == KF_GENERAL && same_general(P, expected, key[k])) break;
      }
      p = (p + step[k]) & uint32_t(S - 1);
    }
    if (i < limit) slot[k] = int(p);
    else if (INSERT) __hip_atomic_store(&misc->overflow, 1, __ATOMIC_RELAXED, WG_SCOPE);  // table full
  }
  if (INSERT && inserted) __hip_atomic_fetch_add(&misc->fill, inserted, __ATOMIC_RELAXED, WG_SCOPE);
  if (KM == KF_HASHED && INSERT) {
    // each joined row's group for the verification, indexed by input row so
    // that the verification reads the rows' own keys in order; a group's
    // representative keeps 0xFFFFFFFF, having nothing to compare
#pragma unroll
    for (int k = 0; k < RPA_CACHED; k++) {
      int64_t const r = r0 + int64_t(k) * RPA_AGG_BLOCK;
      if (r < e && slot[k] >= 0 && !((reps >> k) & 1u)) gp(R.rec_slot)[in.row[k]] = uint32_t(rbase + slot[k]);
    }
  }
}

// Pass A of one chunk: every accumulator, column by column, each accumulator
// applied to all rows of the chunk at once (the first columns' values come
// prefetched, the others are loaded here).
__device__ __forceinline__ void chunk_accumulate(acc_plan const& A, uint8_t* lds, int64_t r0, int64_t e,
                                                 int const (&slot)[RPA_CACHED], uint64_t flp,
                                                 chunk_prefetch const& pf)
{
  uint32_t live = 0;
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) live |= uint32_t(slot[k] >= 0) << k;
  if (A.count_all_off >= 0)
    for_rows(live, [&](int k) {
      __hip_atomic_fetch_add(lds_at<uint32_t>(lds, A.count_all_off, slot[k]), 1u, __ATOMIC_RELAXED, WG_SCOPE);
    });
  auto column = [&](int i, uint64_t (&c)[RPA_CACHED]) {
    uint32_t const cd = lane_value(A.cd_lane, i);
    int const st      = int(cd & 0xffu);
    int const bit     = int((cd >> 8) & 0xffu) - 1;
    int const fb      = int((cd >> 16) & 0xffu);
    int const fe      = int(cd >> 24);
    canon_rows(st, c);
    uint32_t ok = live;
    if (bit >= 0) {
#pragma unroll
      for (int k = 0; k < RPA_CACHED; k++) ok &= ~(uint32_t(((flp >> (8 * k + bit)) & 1) == 0) << k);
    }
    for (int f = fb; f < fe; f++) {
      uint32_t const d = lane_value(A.fd_lane, f);
      acc_rows_lds(int32_t(d & 0xffu), int32_t(d >> 8), st, lds, slot, ok, c, A.wide_f32);
    }
  };
#pragma unroll
  for (int i = 0; i < RPA_AGG_PREFETCH; i++) {
    if (i >= A.ncols) break;
    uint64_t c[RPA_CACHED];
#pragma unroll
    for (int k = 0; k < RPA_CACHED; k++) c[k] = pf.c[i][k];
    column(i, c);
  }
  for (int i = RPA_AGG_PREFETCH; i < A.ncols; i++) {
    uint64_t c[RPA_CACHED];
    load_chunk(plan_column(A, i), storage_width(int(lane_value(A.cd_lane, i) & 0xffu)), r0, e, c);
    column(i, c);
  }
}

// Pass B of one chunk: squared deviations from the group mean.
__device__ __forceinline__ void chunk_deviations(gb_params const& P, rpa_params const& R, uint8_t* lds,
                                                 int64_t r0, int64_t e, int const (&slot)[RPA_CACHED],
                                                 uint64_t flp)
{
  uint8_t const* m2_flag = lds + R.slot_of_j_off;  // groups whose M2 is recomputed
  uint32_t live = 0, flagged = 0;
#pragma unroll
  for (int k = 0; k < RPA_CACHED; k++) {
    live |= uint32_t(slot[k] >= 0) << k;
    flagged |= uint32_t(slot[k] >= 0 && m2_flag[slot[k]] != 0) << k;
  }
  if (flagged == 0) return;
  for (int v = 0; v < P.num_vals; v++) {
    int const fb = P.val_field_begin[v], fe = P.val_field_begin[v + 1];
    int m2 = -1;
    for (int f = fb; f < fe; f++)
      if (P.fields[f].kind == FK_M2) m2 = f;
    if (m2 < 0) continue;
    int const bit       = R.val_flag_bit[v];
    int const st        = P.vals[v].storage;
    int const w         = storage_width(st);
    field_desc const& F = P.fields[m2];
    uint64_t c[RPA_CACHED];
    load_chunk(gp(static_cast<uint8_t const*>(R.valsB[v])), w, r0, e, c);
    canon_rows(st, c);
    uint32_t ok = live;
    if (bit >= 0) {
#pragma unroll
      for (int k = 0; k < RPA_CACHED; k++) ok &= ~(uint32_t(((flp >> (8 * k + bit)) & 1) == 0) << k);
    }
    for_rows(ok & flagged, [&](int k) {
      double const x    = canon_double(st, c[k]);
      double const cnt  = double(lds_count(P, R, lds, F.aux_count, slot[k]));
      double const mean = lds_field_double(P.fields[F.aux_sum], lds, slot[k], R.wide_f32 != 0) / cnt;
      double const d    = x - mean;
      __hip_atomic_fetch_add(lds_at<double>(lds, F.lds_off, slot[k]), d * d, __ATOMIC_RELAXED, WG_SCOPE);
    });
  }
}

// M2 of a group from its pass-A sums: `*m2` holds Σx² (double) on entry and
// Σ(x - mean)² on success.  Integer sources whose Σx² stays below 2^53 (then
// every partial sum and square is exact in double) use the exact integer
// identity n·M2 = n·Σx² − (Σx)² in 128 bits.  Float64 sources use
// Σx² − (Σx)²/n when its rounding error, at most ~3·n·ε·Σx², is below 1e-11
// of the result.  Anything else (float32 sums, large or ill-conditioned
// groups) returns false: the exact second pass computes that group.
__device__ __forceinline__ bool m2_from_sums(int32_t sum_kind, int32_t sum_off, int32_t st, uint8_t* lds, int i,
                                             uint32_t n, double* m2)
{
  double const s2 = *m2;
  if (n <= 1 && isfinite(s2)) {  // no deviation from the mean of one finite value
    *m2 = 0.0;
    return true;
  }  // one NaN or infinite value: the exact pass gives (x - x)² = NaN, as the reference
  if (st == ST_F32) return false;
  if (st == ST_F64) {
    if (n > 256) return false;
    double const s1 = *lds_at<double>(lds, sum_off, i);
    double const v  = s2 - s1 * (s1 / double(n));
    if (!(v > 0.0) || 3.0 * double(n) * 1.1102230246251565e-16 * s2 > 1e-11 * v) return v == 0.0 && s2 == 0.0;
    *m2 = v;
    return true;
  }
  if (sum_kind != FK_SUM_I64 || !(s2 < 9007199254740992.0)) return false;
  int64_t const s1        = *lds_at<int64_t>(lds, sum_off, i);
  __int128 const num      = __int128(n) * __int128(uint64_t(s2)) - __int128(s1) * __int128(s1);
  *m2 = double(num) / double(n);
  return true;
}

template <int KM>
__global__ void __launch_bounds__(RPA_AGG_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
  k_rpa_aggregate(gb_params const* __restrict__ Pp, rpa_params const* __restrict__ Rp, int64_t bin_begin)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  gb_params const& P = *Pp;
  rpa_params const& R = *Rp;
  int const S       = R.lds_slots;
  int const NS      = S + 2;
  auto* tkeys       = reinterpret_cast<uint64_t*>(lds);
  auto* slot_of_j   = reinterpret_cast<uint16_t*>(lds + R.slot_of_j_off);
  auto* misc        = reinterpret_cast<rpa_misc*>(lds + R.misc_off);
  int const tid     = threadIdx.x;
  acc_plan const A  = load_plan(P, R);
  constexpr int64_t CHUNK = int64_t(RPA_CACHED) * RPA_AGG_BLOCK;
  // One workgroup per bin.  The phases share only LDS, so the barriers wait
  // for LDS operations alone, not for the output stores.
  int64_t const b = bin_begin + blockIdx.x;
  {
  // the bin's records: from its start to where the scatter's cursor stopped
  // (within its capacity; a bin that overflowed a speculative layout is
  // redone with an exact one)
  int64_t const s     = R.bin_start[b];
  int64_t e           = int64_t(gp(R.two_level ? R.cursor2 : R.cursor1)[b * RPA_CURSOR_STRIDE]);
  if (e - s > R.cap2) e = s + R.cap2;
  if (s == e) return;
  // the first chunk's loads are in flight while the table is initialised
  chunk_in in;
  chunk_load<true, KM>(A, s + tid, e, in);
  int64_t const rbase = b * R.rep_stride;  // this bin's representative rows (hashed general keys)

  // ---- init
  for (int i = tid; i < S; i += RPA_AGG_BLOCK) tkeys[i] = EMPTY_KEY;
  for (int f = 0; f < A.nfields; f++) {  // identities, from the register plan
    uint32_t const d    = lane_value(A.fd_lane, f);
    int32_t const kind  = int32_t(d & 0xffu);
    int32_t const off   = int32_t(d >> 8);
    uint64_t const idv  = field_identity(kind);
    if (rpa_lds_width(kind, A.wide_f32) == 8)
      for (int i = tid; i < NS; i += RPA_AGG_BLOCK) *lds_at<uint64_t>(lds, off, i) = idv;
    else
      for (int i = tid; i < NS; i += RPA_AGG_BLOCK) *lds_at<uint32_t>(lds, off, i) = uint32_t(idv);
  }
  if (R.count_all_off >= 0)
    for (int i = tid; i < NS; i += RPA_AGG_BLOCK) *lds_at<uint32_t>(lds, R.count_all_off, i) = 0;
  if (tid == 0) {
    misc->fill     = 0;
    misc->present0 = 0;
    misc->present1 = 0;
    misc->overflow = 0;
    misc->rep1     = 0xffffffffu;
  }
  lds_barrier();

  // ---- pass A: find-or-insert and accumulate, chunk by chunk
  for (int64_t c = s; c < e; c += CHUNK) {
    if (__hip_atomic_load(&misc->overflow, __ATOMIC_RELAXED, WG_SCOPE) ||
        __hip_atomic_load(&misc->fill, __ATOMIC_RELAXED, WG_SCOPE) > uint32_t(R.max_fill))
      break;
    if (c != s) chunk_load<true, KM>(A, c + tid, e, in);
    int slot[RPA_CACHED];
    chunk_slots<true, KM>(P, A, R, tkeys, misc, S, c + tid, e, in, slot, rbase);
    chunk_accumulate(A, lds, c + tid, e, slot, in.flp, in.pf);
  }
  lds_barrier();
  if (misc->overflow || misc->fill > uint32_t(R.max_fill)) {  // the call is redone with more bins
    if (tid == 0) astore(R.overflow, 1);
    return;
  }

  // ---- M2 from the pass-A sums where that is exact or well conditioned;
  // the other groups are flagged for the exact second pass
  if (P.need_second_pass) {
    uint8_t* m2_flag = lds + R.slot_of_j_off;  // the compaction map's bytes, free until then
    int const nm2    = R.m2_n;
    int const lane   = tid & 63;
    int const jl     = lane < nm2 ? lane : 0;  // the plan, one M2 field per lane
    int32_t const m2_off_l = R.m2_off[jl], m2_cnt_l = R.m2_cnt_off[jl], m2_st_l = R.m2_storage[jl];
    uint32_t const m2_sum_l = R.m2_sum[jl];
    if (tid == 0) misc->m2_pending = 0;
    lds_barrier();
    for (int i = tid; i < NS; i += RPA_AGG_BLOCK) {
      bool redo = false;
      for (int j = 0; j < nm2; j++) {
        int32_t const off   = int32_t(lane_value(uint32_t(m2_off_l), j));
        uint32_t const sum  = lane_value(m2_sum_l, j);
        double* m2          = lds_at<double>(lds, off, i);
        uint32_t const n    = *lds_at<uint32_t>(lds, int32_t(lane_value(uint32_t(m2_cnt_l), j)), i);
        redo |= !m2_from_sums(int32_t(sum & 0xffu), int32_t(sum >> 8), int32_t(lane_value(uint32_t(m2_st_l), j)),
                              lds, i, n, m2);
      }
      m2_flag[i] = redo ? 1 : 0;
      if (redo) {
        for (int j = 0; j < nm2; j++) *lds_at<double>(lds, int32_t(lane_value(uint32_t(m2_off_l), j)), i) = 0.0;
        __hip_atomic_store(&misc->m2_pending, 1u, __ATOMIC_RELAXED, WG_SCOPE);
      }
    }
    lds_barrier();
  }

  // ---- pass B: squared deviations from the group mean, for flagged groups
  if (P.need_second_pass && misc->m2_pending) {
    for (int64_t c = s; c < e; c += CHUNK) {
      int slot[RPA_CACHED];
      chunk_load<false, KM>(A, c + tid, e, in);
      chunk_slots<false, KM>(P, A, R, tkeys, misc, S, c + tid, e, in, slot, rbase);
      chunk_deviations(P, R, lds, c + tid, e, slot, in.flp);
    }
    lds_barrier();
  }

  // the output plan, one result column per lane (loads in flight during the
  // compaction)
  int const nouts = R.out_n;
  uint32_t out_op_l, out_dlo_l, out_dhi_l, out_mlo_l, out_mhi_l;
  int32_t out_foff_l, out_coff_l, out_ddof_l;
  {
    int const ol = (tid & 63) < nouts ? (tid & 63) : 0;
    out_op_l     = R.out_op[ol];
    out_foff_l   = R.out_foff[ol];
    out_coff_l   = R.out_coff[ol];
    out_ddof_l   = R.out_ddof[ol];
    uint64_t const d = reinterpret_cast<uint64_t>(P.outs[ol].data);
    uint64_t const m = reinterpret_cast<uint64_t>(P.outs[ol].mask);
    out_dlo_l = uint32_t(d);
    out_dhi_l = uint32_t(d >> 32);
    out_mlo_l = uint32_t(m);
    out_mhi_l = uint32_t(m >> 32);
  }

  // ---- compact occupied slots (null-key group first, EMPTY_KEY group next)
  uint32_t const p0 = misc->present0, p1 = misc->present1;
  int const per     = (NS + RPA_AGG_BLOCK - 1) / RPA_AGG_BLOCK;
  int const i0      = tid * per;
  int const i1      = (i0 + per < NS) ? i0 + per : NS;
  // order: S (null), S+1 (EMPTY_KEY), then table slots 0..S-1
  auto occupied = [&](int k) -> bool {
    int const slot = k < 2 ? S + k : k - 2;
    if (slot == S) return p0 != 0;
    if (slot == S + 1) return p1 != 0;
    return tkeys[slot] != EMPTY_KEY;
  };
  uint32_t c = 0;
  for (int k = i0; k < i1; k++) c += occupied(k) ? 1u : 0u;
  // block exclusive scan of c
  int const lane = tid & 63, wave = tid >> 6;
  uint32_t incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) misc->wave_sum[wave] = incl;
  lds_barrier();
  uint32_t before = 0, total = 0;
  for (int w = 0; w < RPA_AGG_BLOCK / 64; w++) {
    if (w < wave) before += misc->wave_sum[w];
    total += misc->wave_sum[w];
  }
  uint32_t j = before + incl - c;
  for (int k = i0; k < i1; k++)
    if (occupied(k)) slot_of_j[j++] = uint16_t(k < 2 ? S + k : k - 2);
  if (tid == 0) {
    misc->g_out = total;
    misc->base  = __hip_atomic_fetch_add(gp(R.group_counter), (unsigned long long)total, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
  }
  lds_barrier();

  // ---- outputs: group j of the workgroup goes to output row base + j.  The
  // waves take 64-row chunks aligned to 64 output rows, so a data store
  // covers whole lines and the validity bits of a chunk are one 64-bit word
  // (stored, or OR-ed in where a neighbouring workgroup shares it; the masks
  // start zeroed).
  uint32_t const g_out          = misc->g_out;
  unsigned long long const base = misc->base;
  uint64_t const end            = base + g_out;
  uint64_t const cap            = R.out_capacity;
  int const kw                  = storage_width(P.key_storage);
  for (uint64_t c0 = (base & ~63ull) + 64 * uint64_t(wave); c0 < end; c0 += RPA_AGG_BLOCK) {
    uint64_t const pos  = c0 + lane;
    bool const live     = pos >= base && pos < end;
    int const slot      = live ? int(slot_of_j[pos - base]) : 0;
    bool const in_range = live && pos < cap;
    bool const kvalid   = slot != S;
    if (live && !in_range) astore(R.overflow, 1);
    if (KM == KF_HASHED) {  // keys are gathered from the representative rows afterwards
      if (in_range)
        gp(R.out_rep)[pos] = __hip_atomic_load(gp(R.rep_rows) + rbase + slot, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    } else if (KM == KF_GENERAL) {
      if (in_range) gp(R.out_rep)[pos] = uint32_t(tkeys[slot]);
    } else {
      if (in_range) {
        uint64_t const key = slot < S ? tkeys[slot] : (slot == S ? 0ull : EMPTY_KEY);
        store_width(P.out_keys[0], int64_t(pos), kw, key);
      }
      if (P.out_key_mask[0]) store_valid_word(P.out_key_mask[0], c0, base, end, cap, in_range && kvalid);
    }
    // outputs in batches: the LDS reads of a batch are issued together, then
    // the batch's values are converted and stored; descriptors come from the
    // register plan
    for (int o0 = 0; o0 < nouts; o0 += RPA_OUT_BATCH) {
      uint64_t raw[RPA_OUT_BATCH];
      uint32_t cnt[RPA_OUT_BATCH];
#pragma unroll
      for (int u = 0; u < RPA_OUT_BATCH; u++) {
        raw[u]      = 0;
        cnt[u]      = 0;
        int const o = o0 + u;
        if (o >= nouts || !in_range) continue;
        uint32_t const op  = lane_value(out_op_l, o);
        int32_t const foff = int32_t(lane_value(uint32_t(out_foff_l), o));
        int32_t const coff = int32_t(lane_value(uint32_t(out_coff_l), o));
        bool const w8      = (op & 0xffu) != OK_COUNT_ALL && rpa_lds_width(int32_t((op >> 8) & 0xffu), A.wide_f32) == 8;
        raw[u] = w8 ? *lds_at<uint64_t>(lds, foff, slot) : uint64_t(*lds_at<uint32_t>(lds, foff, slot));
        if (coff >= 0) cnt[u] = *lds_at<uint32_t>(lds, coff, slot);
      }
#pragma unroll
      for (int u = 0; u < RPA_OUT_BATCH; u++) {
        int const o = o0 + u;
        if (o >= nouts) break;
        uint32_t const op   = lane_value(out_op_l, o);
        int32_t const okind = int32_t(op & 0xffu);
        int32_t const fk    = int32_t((op >> 8) & 0xffu);
        int32_t const ow    = int32_t((op >> 16) & 0xffu);
        bool const counted  = int32_t(lane_value(uint32_t(out_coff_l), o)) >= 0;
        void* const data    = reinterpret_cast<void*>((uint64_t(lane_value(out_dhi_l, o)) << 32) | lane_value(out_dlo_l, o));
        auto* const mask    = reinterpret_cast<uint32_t*>((uint64_t(lane_value(out_mhi_l, o)) << 32) |
                                                       lane_value(out_mlo_l, o));
        bool valid          = !counted || cnt[u] > 0;
        if (in_range) {
          switch (okind) {
            case OK_COPY_I64:
            case OK_COPY_F64: gp(reinterpret_cast<uint64_t*>(data))[pos] = raw[u]; break;
            case OK_COPY_F32: {  // float32 sums come from float64 accumulators
              float const f = rpa_wide_sum(fk, A.wide_f32) ? float(__longlong_as_double(int64_t(raw[u])))
                                                           : __uint_as_float(uint32_t(raw[u]));
              gp(reinterpret_cast<float*>(data))[pos] = f;
              break;
            }
            case OK_COUNT:
            case OK_COUNT_ALL:
              gp(reinterpret_cast<int32_t*>(data))[pos] = int32_t(uint32_t(raw[u]));
              valid = true;
              break;
            case OK_MINMAX: {
              uint64_t bits = raw[u];
              if (fk == FK_MIN_F32 || fk == FK_MAX_F32) bits = unord32(uint32_t(bits));
              else if (fk == FK_MIN_F64 || fk == FK_MAX_F64) bits = unord64(bits);
              store_width(data, int64_t(pos), ow, bits);
              break;
            }
            case OK_MEAN: {
              double const sum = field_bits_double(fk, raw[u], A.wide_f32);
              gp(reinterpret_cast<double*>(data))[pos] = valid ? sum / double(cnt[u]) : 0.0;
              break;
            }
            case OK_SUM_I32: gp(reinterpret_cast<int32_t*>(data))[pos] = int32_t(uint32_t(raw[u])); break;
            case OK_MEAN_INT: {  // fixed-point mean: the representation's sum / count, truncated
              int64_t sum = int64_t(raw[u]);
              if (ow == 4) sum = int64_t(int32_t(uint32_t(sum)));
              store_width(data, int64_t(pos), ow, uint64_t(valid ? sum / int64_t(cnt[u]) : 0));
              break;
            }
            case OK_VAR:
            case OK_STD: {
              int64_t const dn = int64_t(cnt[u]) - int32_t(lane_value(uint32_t(out_ddof_l), o));
              valid            = cnt[u] > 0 && dn > 0;
              double var       = valid ? __longlong_as_double(int64_t(raw[u])) / double(dn) : 0.0;
              gp(reinterpret_cast<double*>(data))[pos] = (okind == OK_STD) ? sqrt(var) : var;
              break;
            }
            default: break;
          }
        }
        if (mask) store_valid_word(mask, c0, base, end, cap, in_range && valid);
      }
    }
  }
  }
}

template <int KM>
static void launch_aggregate_form(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int64_t bin_begin,
                                  int64_t bin_end, hipStream_t s)
{
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_rpa_aggregate<KM>), hR.lds_bytes);
  hipLaunchKernelGGL(k_rpa_aggregate<KM>, dim3(bin_end - bin_begin), dim3(RPA_AGG_BLOCK), hR.lds_bytes, s, dP, dR,
                     bin_begin);
}

void launch_rpa_aggregate(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int key_form,
                          int64_t bin_begin, int64_t bin_end, hipStream_t s)
{
  if (bin_end <= bin_begin) return;
  if (key_form == KF_HASHED) launch_aggregate_form<KF_HASHED>(dP, dR, hR, bin_begin, bin_end, s);
  else if (key_form == KF_GENERAL) launch_aggregate_form<KF_GENERAL>(dP, dR, hR, bin_begin, bin_end, s);
  else launch_aggregate_form<KF_INLINE>(dP, dR, hR, bin_begin, bin_end, s);
}

// Exact comparison of every record of the hashed form with its group's
// representative: rows whose group was found by equal 64-bit words must also
// have equal keys (nulls equal, NaN equal NaN, -0.0 equal 0.0, strings
// byte-wise), otherwise verify_failed is set and the call is redone comparing
// key rows.  Rows are visited in input order, so a row's own keys are read
// sequentially and only the representatives' are random reads.
__global__ void __launch_bounds__(256) k_rpa_verify(gb_params const* __restrict__ Pp, rpa_params const* __restrict__ Rp,
                                                    int64_t records)
{
  __shared__ uint32_t block_count;
  gb_params const& P  = *Pp;
  rpa_params const& R = *Rp;
  if (threadIdx.x == 0) block_count = 0;
  __syncthreads();
  bool bad       = false;
  uint32_t count = 0;
  for (int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x; r < records; r += int64_t(gridDim.x) * 256) {
    uint32_t const gs = gp(R.rec_slot)[r];
    if (gs == 0xffffffffu) continue;  // a dropped row, or a representative
    uint32_t const row = uint32_t(r);
    uint32_t const rep = __hip_atomic_load(gp(R.rep_rows) + gs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (rep == row) continue;
    count++;
    bad |= !rows_equal(P, int64_t(row), int64_t(rep));
  }
  if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) astore(R.verify_failed, 1);
  if (count) __hip_atomic_fetch_add(&block_count, count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __syncthreads();
  if (threadIdx.x == 0 && block_count)
    __hip_atomic_fetch_add(gp(R.verified), (unsigned long long)block_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

void launch_rpa_verify(gb_params const* dP, rpa_params const* dR, int64_t records, hipStream_t s)
{
  if (records <= 0) return;
  int64_t blocks = (records + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_rpa_verify, dim3(blocks), dim3(256), 0, s, dP, dR, records);
}

}  // namespace mig
