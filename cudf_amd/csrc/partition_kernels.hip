// Radix-partitioned hash aggregation for high-cardinality keys on gfx950.
//
// The reference aggregates every row with device-wide atomics into sparse
// result columns (groupby/hash/groupby_kernels.cuh:98-105) and runs a second
// full pass for VARIANCE (multi_pass_kernels.cuh:74-111).  With ~10 rows per
// group and millions of groups that state does not fit in any on-chip cache,
// so every row costs several random HBM lines.  Here rows are first moved, by
// the top bits of their MurmurHash3 row hash, into bins whose groups fit in
// one workgroup's LDS; each bin is then aggregated entirely with LDS atomics
// (including the variance pass, whose re-read of the bin hits L2) and its
// groups are written straight to the dense output columns.
//
//   k_rpa_hist      rows per final bin (LDS histogram per block)
//   k_rpa_scatter1  input columns -> first-level bins (buffer A, or B when
//                   one level suffices); ranks from LDS atomics, one global
//                   range reservation per (tile, bin)
//   k_rpa_scatter2  first-level bins -> final bins (buffer B)
//   k_rpa_aggregate one workgroup per final bin: LDS hash table of the bin's
//                   keys, dense ids, LDS accumulators, outputs
#include "device_ops.hpp"

namespace mig {

__device__ __forceinline__ uint32_t inline_key_hash(int32_t st, uint64_t bits)
{
  return hash_combine(0u, hash_fixed_bits(st, bits, 0u));
}
__device__ __forceinline__ uint32_t null_key_hash() { return hash_combine(0u, NULL_HASH); }

__device__ __forceinline__ uint32_t top_bits(uint32_t h, int nbits)
{
  return nbits == 0 ? 0u : (h >> (32 - nbits));
}

// Hash of input row `row`; returns false when the row is dropped (null key
// under EXCLUDE).  `bits` receives the normalised key, `key_null` whether the
// key is null (INCLUDE).
__device__ __forceinline__ bool input_key(gb_params const& P, int64_t row, uint64_t& bits,
                                          bool& key_null, uint32_t& h)
{
  dcol const& kc = P.keys[0];
  if (!dcol_valid(kc, row)) {
    if (P.skip_null_keys) return false;
    key_null = true;
    bits     = 0;
    h        = null_key_hash();
    return true;
  }
  key_null = false;
  bits     = load_fixed_bits(kc, row);
  h        = inline_key_hash(kc.storage, bits);
  return true;
}

// ---------------------------------------------------------------------------
// Rows per final bin.  Each block counts its slice of the rows in LDS and
// writes its own histogram row; k_rpa_hist_reduce sums the rows.
__global__ void __launch_bounds__(RPA_HIST_BLOCK)
  k_rpa_hist(gb_params const* __restrict__ Pp, rpa_params const* __restrict__ Rp, int64_t n)
{
  extern __shared__ uint32_t hist[];
  gb_params const& P = *Pp;
  rpa_params const& R = *Rp;
  int const B = 1 << R.log_bins;
  for (int i = threadIdx.x; i < B; i += RPA_HIST_BLOCK) hist[i] = 0;
  __syncthreads();
  int64_t const per = (n + gridDim.x - 1) / gridDim.x;
  int64_t const r0  = int64_t(blockIdx.x) * per;
  int64_t const r1  = (r0 + per < n) ? r0 + per : n;
  constexpr int U   = 4;  // independent rows in flight per thread
  for (int64_t row0 = r0 + threadIdx.x; row0 < r1; row0 += int64_t(U) * RPA_HIST_BLOCK) {
    uint32_t bin[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int64_t const row = row0 + int64_t(u) * RPA_HIST_BLOCK;
      bin[u]            = 0xffffffffu;
      uint64_t bits;
      bool key_null;
      uint32_t h;
      if (row < r1 && input_key(P, row, bits, key_null, h)) bin[u] = top_bits(h, R.log_bins);
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (bin[u] != 0xffffffffu)
        __hip_atomic_fetch_add(&hist[bin[u]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  uint32_t* out = R.hist_partial + int64_t(blockIdx.x) * B;
  for (int i = threadIdx.x; i < B; i += RPA_HIST_BLOCK) out[i] = hist[i];
}

__global__ void __launch_bounds__(RPA_SMALL_BLOCK) k_rpa_hist_reduce(rpa_params const* __restrict__ Rp)
{
  rpa_params const& R = *Rp;
  int const B         = 1 << R.log_bins;
  int const b         = blockIdx.x * RPA_SMALL_BLOCK + threadIdx.x;
  if (b >= B) return;
  uint32_t s = 0;
  for (int k = 0; k < R.hist_blocks; k++) s += R.hist_partial[int64_t(k) * B + b];
  R.hist[b] = s;
}

void launch_rpa_hist(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int64_t n,
                     int num_cus, hipStream_t s)
{
  size_t lds = sizeof(uint32_t) << hR.log_bins;
  if (lds > 64 * 1024)
    MIG_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<void const*>(&k_rpa_hist),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  hipLaunchKernelGGL(k_rpa_hist, dim3(hR.hist_blocks), dim3(RPA_HIST_BLOCK), lds, s, dP, dR, n);
  int64_t const B = int64_t(1) << hR.log_bins;
  hipLaunchKernelGGL(k_rpa_hist_reduce, dim3((B + RPA_SMALL_BLOCK - 1) / RPA_SMALL_BLOCK),
                     dim3(RPA_SMALL_BLOCK), 0, s, dR);
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ void copy_width(uint8_t* dst, int64_t di, uint8_t const* src, int64_t si, int w)
{
  switch (w) {
    case 1: dst[di] = src[si]; break;
    case 2: reinterpret_cast<uint16_t*>(dst)[di] = reinterpret_cast<uint16_t const*>(src)[si]; break;
    case 4: reinterpret_cast<uint32_t*>(dst)[di] = reinterpret_cast<uint32_t const*>(src)[si]; break;
    default: reinterpret_cast<uint64_t*>(dst)[di] = reinterpret_cast<uint64_t const*>(src)[si]; break;
  }
}

constexpr int RPA_MAX_RPT    = RPA_TILE / RPA_SCATTER_BLOCK;  // rows per thread per tile (at most)
constexpr int RPA_MAX_FANOUT = 256;

// Exclusive scan of cnt[0, fan) into off[0, fan] by wave 0 (fan <= 256).
__device__ __forceinline__ void scan_counts(uint32_t const* cnt, uint32_t* off, int fan)
{
  if (threadIdx.x >= 64) return;
  int const lane = threadIdx.x;
  uint32_t v[4];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int const i = lane * 4 + k;
    v[k]        = i < fan ? cnt[i] : 0u;
    sum += v[k];
  }
  uint32_t incl = sum;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  uint32_t run = incl - sum;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int const i = lane * 4 + k;
    if (i < fan) off[i] = run;
    run += v[k];
  }
  if (lane == 63) off[fan] = incl;
}

// One tile of the scatter: rows are ranked per bin with LDS atomics, staged
// in LDS in bin order, and written out so that consecutive threads store
// consecutive rows of a bin (coalesced runs of tile_rows / fanout rows).
//   SECOND = false: input columns -> first-level bins
//   SECOND = true : first-level bin in buffer A -> final bins in buffer B
template <bool SECOND>
__global__ void __launch_bounds__(RPA_SCATTER_BLOCK)
  k_rpa_scatter(gb_params const* __restrict__ Pp, rpa_params const* __restrict__ Rp, int64_t n)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t stage[];
  __shared__ uint32_t cnt[RPA_MAX_FANOUT];
  __shared__ uint32_t toff[RPA_MAX_FANOUT + 1];
  __shared__ unsigned long long base[RPA_MAX_FANOUT];
  __shared__ int64_t tstart[SECOND ? RPA_MAX_FANOUT + 1 : 1];
  gb_params const& P = *Pp;
  rpa_params const& R = *Rp;
  int const T     = R.tile_rows;
  int const rpt   = T / RPA_SCATTER_BLOCK;
  int const lf2   = R.log_bins - R.log_f1;
  int const F1    = 1 << R.log_f1;
  int const fan   = SECOND ? (1 << lf2) : F1;
  auto* skeys     = reinterpret_cast<uint64_t*>(stage);
  auto* sflags    = stage + R.stage_flags_off;
  auto* sbins     = reinterpret_cast<uint16_t*>(stage + R.stage_bin_off);
  uint64_t* const okeys = (SECOND || !R.two_level) ? R.keysB : R.keysA;
  uint8_t* const oflags = (SECOND || !R.two_level) ? R.flagsB : R.flagsA;
  int64_t ntiles;
  if constexpr (SECOND) {
    for (int i = threadIdx.x; i <= F1; i += RPA_SCATTER_BLOCK) tstart[i] = R.tile_start1[i];
    __syncthreads();
    ntiles = R.num_tiles1;
  } else {
    ntiles = (n + T - 1) / T;
  }
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // rows [r0, r1) of this tile; for the second pass `b1` is the first-level bin
    int64_t r0, r1;
    int b1 = 0;
    if constexpr (SECOND) {
      int lo = 0, hi = F1 - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tstart[mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      b1                 = lo;
      int64_t const bs   = R.bin_start[int64_t(b1) << lf2];
      int64_t const be   = R.bin_start[int64_t(b1 + 1) << lf2];
      r0                 = bs + (t - tstart[b1]) * T;
      r1                 = (r0 + T < be) ? r0 + T : be;
    } else {
      r0 = t * T;
      r1 = (r0 + T < n) ? r0 + T : n;
    }
    for (int i = threadIdx.x; i < fan; i += RPA_SCATTER_BLOCK) cnt[i] = 0;
    __syncthreads();
    // ---- rank rows per bin
    int bin[RPA_MAX_RPT];
    uint32_t rank[RPA_MAX_RPT];
    uint64_t kb[RPA_MAX_RPT];
    uint8_t kf[RPA_MAX_RPT];
#pragma unroll
    for (int k = 0; k < RPA_MAX_RPT; k++) {
      bin[k] = -1;
      if (k >= rpt) continue;
      int64_t const row = r0 + int64_t(k) * RPA_SCATTER_BLOCK + threadIdx.x;
      if (row >= r1) continue;
      uint32_t h;
      if constexpr (SECOND) {
        kf[k] = R.flagsA[row];
        kb[k] = R.keysA[row];
        h     = (kf[k] & RPA_KEY_NULL) ? null_key_hash() : inline_key_hash(P.key_storage, kb[k]);
        bin[k] = int(top_bits(h, R.log_bins) & uint32_t(fan - 1));
      } else {
        bool kn;
        if (!input_key(P, row, kb[k], kn, h)) continue;
        kf[k]  = kn ? RPA_KEY_NULL : 0;
        bin[k] = int(top_bits(h, R.log_f1));
      }
      rank[k] = __hip_atomic_fetch_add(&cnt[bin[k]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    scan_counts(cnt, toff, fan);
    for (int i = threadIdx.x; i < fan; i += RPA_SCATTER_BLOCK) {
      if (cnt[i]) {
        unsigned long long* cur = SECOND ? R.cursor2 + ((int64_t(b1) << lf2) + i) : R.cursor1 + i;
        base[i] = __hip_atomic_fetch_add(cur, (unsigned long long)cnt[i], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
    // ---- stage the tile in bin order
#pragma unroll
    for (int k = 0; k < RPA_MAX_RPT; k++) {
      if (bin[k] < 0) continue;
      int64_t const row = r0 + int64_t(k) * RPA_SCATTER_BLOCK + threadIdx.x;
      int const lp      = int(toff[bin[k]] + rank[k]);
      skeys[lp]         = kb[k];
      sbins[lp]         = uint16_t(bin[k]);
      uint8_t flags     = kf[k];
      for (int v = 0; v < P.num_vals; v++) {
        if (P.val_field_begin[v] == P.val_field_begin[v + 1]) continue;
        int const w = storage_width(P.vals[v].storage);
        if constexpr (SECOND) {
          copy_width(stage + R.stage_val_off[v], lp, R.valsA[v], row, w);
        } else {
          dcol const& c = P.vals[v];
          copy_width(stage + R.stage_val_off[v], lp, c.data, row, w);
          int const fb = R.val_flag_bit[v];
          if (fb >= 0 && dcol_valid(c, row)) flags |= uint8_t(1u << fb);
        }
      }
      sflags[lp] = flags;
    }
    __syncthreads();
    // ---- write bin runs
    int const total = int(toff[fan]);
    for (int lp = threadIdx.x; lp < total; lp += RPA_SCATTER_BLOCK) {
      int const b       = sbins[lp];
      int64_t const dst = int64_t(base[b]) + (lp - int(toff[b]));
      okeys[dst]        = skeys[lp];
      oflags[dst]       = sflags[lp];
      for (int v = 0; v < P.num_vals; v++) {
        if (P.val_field_begin[v] == P.val_field_begin[v + 1]) continue;
        uint8_t* const ov = (SECOND || !R.two_level) ? R.valsB[v] : R.valsA[v];
        copy_width(ov, dst, stage + R.stage_val_off[v], lp, storage_width(P.vals[v].storage));
      }
    }
    __syncthreads();
  }
}

// Dynamic LDS above 64 KB must be opted into per kernel (gfx950 has 160 KB).
static void allow_large_lds(void const* fn, int bytes)
{
  if (bytes > 64 * 1024)
    MIG_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

void launch_rpa_scatter1(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int64_t n,
                         int num_cus, hipStream_t s)
{
  int64_t tiles  = (n + hR.tile_rows - 1) / hR.tile_rows;
  int64_t blocks = int64_t(num_cus) * 4;
  if (blocks > tiles) blocks = tiles;
  if (blocks < 1) blocks = 1;
  allow_large_lds(reinterpret_cast<void const*>(&k_rpa_scatter<false>), hR.stage_bytes);
  hipLaunchKernelGGL(k_rpa_scatter<false>, dim3(blocks), dim3(RPA_SCATTER_BLOCK), hR.stage_bytes, s, dP, dR, n);
}

void launch_rpa_scatter2(gb_params const* dP, rpa_params const* dR, rpa_params const& hR,
                         int num_cus, hipStream_t s)
{
  int64_t blocks = int64_t(num_cus) * 4;
  if (blocks > hR.num_tiles1) blocks = hR.num_tiles1;
  if (blocks < 1) blocks = 1;
  allow_large_lds(reinterpret_cast<void const*>(&k_rpa_scatter<true>), hR.stage_bytes);
  hipLaunchKernelGGL(k_rpa_scatter<true>, dim3(blocks), dim3(RPA_SCATTER_BLOCK), hR.stage_bytes, s, dP, dR,
                     int64_t(0));
}

// ---------------------------------------------------------------------------
// LDS accumulators
__device__ __forceinline__ void apply_field_lds(field_desc const& F, uint8_t* lds, int id,
                                                loaded_value const& v)
{
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
  switch (F.kind) {
    case FK_SUM_I64:
      __hip_atomic_fetch_add(lds_at<uint64_t>(lds, F.lds_off, id), uint64_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_SUM_F32:
      __hip_atomic_fetch_add(lds_at<float>(lds, F.lds_off, id), v.f, __ATOMIC_RELAXED, WG);
      break;
    case FK_SUM_F64:
      __hip_atomic_fetch_add(lds_at<double>(lds, F.lds_off, id), v.d, __ATOMIC_RELAXED, WG);
      break;
    case FK_SUMSQ_I64:
      __hip_atomic_fetch_add(lds_at<uint64_t>(lds, F.lds_off, id), uint64_t(v.i) * uint64_t(v.i),
                             __ATOMIC_RELAXED, WG);
      break;
    case FK_SUMSQ_F32:
      __hip_atomic_fetch_add(lds_at<float>(lds, F.lds_off, id), v.f * v.f, __ATOMIC_RELAXED, WG);
      break;
    case FK_SUMSQ_F64:
      __hip_atomic_fetch_add(lds_at<double>(lds, F.lds_off, id), v.d * v.d, __ATOMIC_RELAXED, WG);
      break;
    case FK_MIN_S32:
      __hip_atomic_fetch_min(lds_at<int32_t>(lds, F.lds_off, id), int32_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_S32:
      __hip_atomic_fetch_max(lds_at<int32_t>(lds, F.lds_off, id), int32_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_MIN_S64:
      __hip_atomic_fetch_min(lds_at<int64_t>(lds, F.lds_off, id), v.i, __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_S64:
      __hip_atomic_fetch_max(lds_at<int64_t>(lds, F.lds_off, id), v.i, __ATOMIC_RELAXED, WG);
      break;
    case FK_MIN_U32:
      __hip_atomic_fetch_min(lds_at<uint32_t>(lds, F.lds_off, id), uint32_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_U32:
      __hip_atomic_fetch_max(lds_at<uint32_t>(lds, F.lds_off, id), uint32_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_MIN_U64:
      __hip_atomic_fetch_min(lds_at<uint64_t>(lds, F.lds_off, id), uint64_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_U64:
      __hip_atomic_fetch_max(lds_at<uint64_t>(lds, F.lds_off, id), uint64_t(v.i), __ATOMIC_RELAXED, WG);
      break;
    case FK_MIN_F32:
      if (!isnan(v.f))
        __hip_atomic_fetch_min(lds_at<uint32_t>(lds, F.lds_off, id), ord32(uint32_t(v.b)), __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_F32:
      if (!isnan(v.f))
        __hip_atomic_fetch_max(lds_at<uint32_t>(lds, F.lds_off, id), ord32(uint32_t(v.b)), __ATOMIC_RELAXED, WG);
      break;
    case FK_MIN_F64:
      if (!isnan(v.d))
        __hip_atomic_fetch_min(lds_at<uint64_t>(lds, F.lds_off, id), ord64(v.b), __ATOMIC_RELAXED, WG);
      break;
    case FK_MAX_F64:
      if (!isnan(v.d))
        __hip_atomic_fetch_max(lds_at<uint64_t>(lds, F.lds_off, id), ord64(v.b), __ATOMIC_RELAXED, WG);
      break;
    case FK_COUNT:
      __hip_atomic_fetch_add(lds_at<uint32_t>(lds, F.lds_off, id), 1u, __ATOMIC_RELAXED, WG);
      break;
    default: break;
  }
}

__device__ __forceinline__ uint64_t lds_field_bits(field_desc const& F, uint8_t* lds, int id)
{
  return F.width == 8 ? *lds_at<uint64_t>(lds, F.lds_off, id) : *lds_at<uint32_t>(lds, F.lds_off, id);
}

__device__ __forceinline__ double lds_field_double(field_desc const& F, uint8_t* lds, int id)
{
  switch (F.kind) {
    case FK_SUM_I64: return double(*lds_at<int64_t>(lds, F.lds_off, id));
    case FK_SUM_F32: return double(*lds_at<float>(lds, F.lds_off, id));
    case FK_SUM_F64: return *lds_at<double>(lds, F.lds_off, id);
    case FK_COUNT: return double(*lds_at<uint32_t>(lds, F.lds_off, id));
    default: return 0.0;
  }
}

__device__ __forceinline__ uint32_t lds_count(gb_params const& P, rpa_params const& R, uint8_t* lds,
                                              int32_t f, int id)
{
  return f == -2 ? *lds_at<uint32_t>(lds, R.count_all_off, id) : *lds_at<uint32_t>(lds, P.fields[f].lds_off, id);
}

// Dense id of partitioned row r: 0 = null key, 1 = key equal to EMPTY_KEY,
// otherwise the id stored with the key's LDS slot.
__device__ __forceinline__ int find_id(rpa_params const& R, int32_t key_storage, uint64_t const* tkeys,
                                       uint16_t const* tids, int S, uint8_t fl, uint64_t key)
{
  if (fl & RPA_KEY_NULL) return 0;
  if (key == EMPTY_KEY) return 1;
  uint32_t slot = inline_key_hash(key_storage, key) & uint32_t(S - 1);
  for (int i = 0; i < S; i++) {
    if (tkeys[slot] == key) return tids[slot];
    slot = (slot + 1) & uint32_t(S - 1);
  }
  return -1;
}

// Validity bits of the 64 output rows [p0, p0 + 64) of one wave, clipped to
// the workgroup's rows [.., end): words wholly inside are stored, words shared
// with a neighbouring workgroup are OR-ed in (the mask starts zeroed).
__device__ __forceinline__ void write_valid_bits(uint32_t* mask, uint64_t p0, uint64_t end, bool valid)
{
  uint64_t const bal = __ballot(valid);
  if ((threadIdx.x & 63) != 0) return;
  uint64_t const last = (p0 + 64 < end) ? p0 + 64 : end;
  if (last <= p0) return;
  for (uint64_t w = p0 >> 5; w <= (last - 1) >> 5; w++) {
    int64_t const shift = int64_t(w * 32) - int64_t(p0);
    uint32_t bits       = shift >= 0 ? uint32_t(bal >> shift) : uint32_t(bal << (-shift));
    bool const whole    = (w * 32 >= p0) && (w * 32 + 32 <= last);
    if (whole) mask[w] = bits;
    else if (bits)
      __hip_atomic_fetch_or(mask + w, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

struct rpa_misc {
  uint32_t ngroups;
  uint32_t present0;
  uint32_t present1;
  int32_t overflow;
  unsigned long long base;
};

__global__ void __launch_bounds__(RPA_AGG_BLOCK)
  k_rpa_aggregate(gb_params const* __restrict__ Pp, rpa_params const* __restrict__ Rp)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  gb_params const& P = *Pp;
  rpa_params const& R = *Rp;
  int const b       = blockIdx.x;
  int64_t const s   = R.bin_start[b];
  int64_t const e   = R.bin_start[b + 1];
  if (s == e) return;
  int const S = R.lds_slots, D = R.max_ids;
  auto* tkeys = reinterpret_cast<uint64_t*>(lds);
  auto* tids  = reinterpret_cast<uint16_t*>(lds + R.ids_off);
  auto* misc  = reinterpret_cast<rpa_misc*>(lds + R.misc_off);
  auto* slot_of_id = reinterpret_cast<uint16_t*>(lds + R.slot_of_id_off);
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;

  // ---- init
  for (int i = threadIdx.x; i < S; i += RPA_AGG_BLOCK) tkeys[i] = EMPTY_KEY;
  for (int f = 0; f < P.num_fields; f++) {
    field_desc const& F = P.fields[f];
    uint64_t const id   = field_identity(F.kind);
    if (F.width == 8)
      for (int i = threadIdx.x; i < D; i += RPA_AGG_BLOCK) *lds_at<uint64_t>(lds, F.lds_off, i) = id;
    else
      for (int i = threadIdx.x; i < D; i += RPA_AGG_BLOCK) *lds_at<uint32_t>(lds, F.lds_off, i) = uint32_t(id);
  }
  if (R.count_all_off >= 0)
    for (int i = threadIdx.x; i < D; i += RPA_AGG_BLOCK) *lds_at<uint32_t>(lds, R.count_all_off, i) = 0;
  if (threadIdx.x == 0) {
    misc->ngroups  = 0;
    misc->present0 = 0;
    misc->present1 = 0;
    misc->overflow = 0;
  }
  __syncthreads();

  // ---- phase 1: keys of the bin -> LDS table, dense ids from 2
  for (int64_t r = s + threadIdx.x; r < e; r += RPA_AGG_BLOCK) {
    // once the bin is known not to fit, stop inserting: the call is retried
    // with more bins, and a full LDS table would make every probe long
    if (__hip_atomic_load(&misc->overflow, __ATOMIC_RELAXED, WG)) break;
    uint8_t const fl = R.flagsB[r];
    if (fl & RPA_KEY_NULL) {
      misc->present0 = 1;
      continue;
    }
    uint64_t const key = R.keysB[r];
    if (key == EMPTY_KEY) {
      misc->present1 = 1;
      continue;
    }
    uint32_t slot = inline_key_hash(P.key_storage, key) & uint32_t(S - 1);
    int i = 0;
    for (; i < S; i++) {
      uint64_t cur = __hip_atomic_load(tkeys + slot, __ATOMIC_RELAXED, WG);
      if (cur == key) break;
      if (cur == EMPTY_KEY) {
        uint64_t expected = EMPTY_KEY;
        if (__hip_atomic_compare_exchange_strong(tkeys + slot, &expected, key, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, WG)) {
          uint32_t id = __hip_atomic_fetch_add(&misc->ngroups, 1u, __ATOMIC_RELAXED, WG) + 2u;
          if (id < uint32_t(D)) {
            tids[slot]       = uint16_t(id);
            slot_of_id[id]   = uint16_t(slot);
          }
          else __hip_atomic_store(&misc->overflow, 1, __ATOMIC_RELAXED, WG);
          break;
        }
        if (expected == key) break;
      }
      slot = (slot + 1) & uint32_t(S - 1);
    }
    if (i == S) __hip_atomic_store(&misc->overflow, 1, __ATOMIC_RELAXED, WG);
  }
  __syncthreads();
  if (misc->overflow) {
    if (threadIdx.x == 0) astore(R.overflow, 1);
    return;
  }

  // ---- phase 2: accumulate
  for (int64_t r = s + threadIdx.x; r < e; r += RPA_AGG_BLOCK) {
    uint8_t const fl = R.flagsB[r];
    int const id     = find_id(R, P.key_storage, tkeys, tids, S, fl, R.keysB[r]);
    if (id < 0) continue;
    if (R.count_all_off >= 0)
      __hip_atomic_fetch_add(lds_at<uint32_t>(lds, R.count_all_off, id), 1u, __ATOMIC_RELAXED, WG);
    for (int v = 0; v < P.num_vals; v++) {
      int const fb = P.val_field_begin[v], fe = P.val_field_begin[v + 1];
      if (fb == fe) continue;
      int const bit = R.val_flag_bit[v];
      if (bit >= 0 && !((fl >> bit) & 1)) continue;
      dcol c{};
      c.data    = R.valsB[v];
      c.storage = P.vals[v].storage;
      loaded_value const val = load_value(c, r);
      for (int f = fb; f < fe; f++) apply_field_lds(P.fields[f], lds, id, val);
    }
  }
  __syncthreads();

  // ---- phase 3: squared deviations from the group mean (VARIANCE / STD)
  if (P.need_second_pass) {
    for (int64_t r = s + threadIdx.x; r < e; r += RPA_AGG_BLOCK) {
      uint8_t const fl = R.flagsB[r];
      int id           = -2;
      for (int v = 0; v < P.num_vals; v++) {
        int const fb = P.val_field_begin[v], fe = P.val_field_begin[v + 1];
        bool has_m2  = false;
        for (int f = fb; f < fe; f++) has_m2 |= (P.fields[f].kind == FK_M2);
        if (!has_m2) continue;
        int const bit = R.val_flag_bit[v];
        if (bit >= 0 && !((fl >> bit) & 1)) continue;
        if (id == -2) id = find_id(R, P.key_storage, tkeys, tids, S, fl, R.keysB[r]);
        if (id < 0) break;
        dcol c{};
        c.data    = R.valsB[v];
        c.storage = P.vals[v].storage;
        loaded_value const val = load_value(c, r);
        for (int f = fb; f < fe; f++) {
          field_desc const& F = P.fields[f];
          if (F.kind != FK_M2) continue;
          double const cnt  = double(lds_count(P, R, lds, F.aux_count, id));
          double const mean = lds_field_double(P.fields[F.aux_sum], lds, id) / cnt;
          double const d    = val.d - mean;
          __hip_atomic_fetch_add(lds_at<double>(lds, F.lds_off, id), d * d, __ATOMIC_RELAXED, WG);
        }
      }
    }
    __syncthreads();
  }

  // ---- phase 4: outputs
  uint32_t const p0 = misc->present0, p1 = misc->present1;
  uint32_t const ng = misc->ngroups;
  if (threadIdx.x == 0)
    misc->base = __hip_atomic_fetch_add(R.group_counter, (unsigned long long)(ng + p0 + p1),
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  unsigned long long const base = misc->base;
  int const kw            = storage_width(P.key_storage);
  uint32_t const g_out    = ng + p0 + p1;
  uint64_t const end      = base + g_out;
  int const lane          = threadIdx.x & 63;
  // output index j = consecutive rows base + j; one wave covers 64 of them,
  // so values are stored coalesced and validity words are built by ballot
  for (uint32_t j0 = threadIdx.x & ~63u; j0 < g_out; j0 += RPA_AGG_BLOCK) {
    uint32_t const j  = j0 + lane;
    bool const live   = j < g_out;
    uint64_t const pos = base + j;
    int id            = 0;
    uint64_t key      = 0;
    bool kvalid       = true;
    if (live) {
      if (p0 && j == 0) {
        id = 0;
        kvalid = false;
      } else if (p1 && j == p0) {
        id  = 1;
        key = EMPTY_KEY;
      } else {
        id  = int(j - p0 - p1) + 2;
        key = tkeys[slot_of_id[id]];
      }
    }
    bool const in_range = live && pos < R.out_capacity;
    if (live && !in_range) astore(R.overflow, 1);
    if (in_range) store_width(P.out_keys[0], int64_t(pos), kw, kvalid ? key : 0);
    if (P.out_key_mask[0]) write_valid_bits(P.out_key_mask[0], base + j0, end, in_range && kvalid);
    for (int o = 0; o < P.num_outs; o++) {
      out_desc const& O = P.outs[o];
      bool valid        = true;
      if (in_range) {
        uint32_t cnt = 0;
        if (O.count_field != -1) {
          cnt   = lds_count(P, R, lds, O.count_field, id);
          valid = cnt > 0;
        }
        switch (O.kind) {
          case OK_COPY_I64:
          case OK_COPY_F64:
            reinterpret_cast<uint64_t*>(O.data)[pos] = lds_field_bits(P.fields[O.field], lds, id);
            break;
          case OK_COPY_F32:
            reinterpret_cast<uint32_t*>(O.data)[pos] = uint32_t(lds_field_bits(P.fields[O.field], lds, id));
            break;
          case OK_COUNT:
            reinterpret_cast<int32_t*>(O.data)[pos] = int32_t(*lds_at<uint32_t>(lds, P.fields[O.field].lds_off, id));
            valid = true;
            break;
          case OK_COUNT_ALL:
            reinterpret_cast<int32_t*>(O.data)[pos] = int32_t(*lds_at<uint32_t>(lds, R.count_all_off, id));
            valid = true;
            break;
          case OK_MINMAX: {
            field_desc const& F = P.fields[O.field];
            uint64_t bits       = lds_field_bits(F, lds, id);
            if (F.kind == FK_MIN_F32 || F.kind == FK_MAX_F32) bits = unord32(uint32_t(bits));
            else if (F.kind == FK_MIN_F64 || F.kind == FK_MAX_F64) bits = unord64(bits);
            store_width(O.data, int64_t(pos), O.out_width, bits);
            break;
          }
          case OK_MEAN: {
            double const sum = lds_field_double(P.fields[O.field], lds, id);
            reinterpret_cast<double*>(O.data)[pos] = valid ? sum / double(cnt) : 0.0;
            break;
          }
          case OK_VAR:
          case OK_STD: {
            int64_t const dn = int64_t(cnt) - O.ddof;
            valid            = cnt > 0 && dn > 0;
            double var = valid ? *lds_at<double>(lds, P.fields[O.field].lds_off, id) / double(dn) : 0.0;
            reinterpret_cast<double*>(O.data)[pos] = (O.kind == OK_STD) ? sqrt(var) : var;
            break;
          }
          default: break;
        }
      }
      if (O.mask) write_valid_bits(O.mask, base + j0, end, in_range && valid);
    }
  }
}

void launch_rpa_aggregate(gb_params const* dP, rpa_params const* dR, rpa_params const& hR,
                          hipStream_t s)
{
  int64_t const B = int64_t(1) << hR.log_bins;
  allow_large_lds(reinterpret_cast<void const*>(&k_rpa_aggregate), hR.lds_bytes);
  hipLaunchKernelGGL(k_rpa_aggregate, dim3(B), dim3(RPA_AGG_BLOCK), hR.lds_bytes, s, dP, dR);
}

}  // namespace mig
