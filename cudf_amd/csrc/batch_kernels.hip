// Batched hash group-by kernels: many independent group-bys (one fixed-width
// key each, per-instance key and value types) in one persistent launch.
//
//   k_batch_sample     per instance: distinct row hashes of a strided sample
//   k_batch_init_*     tables, accumulators (each instance's identities),
//                      flags and output masks reset by fill kernels (the
//                      sequence is captured in a hipGraph, where memset nodes
//                      are avoided)
//   k_batch_aggregate  persistent workgroups pull (instance, row tile) items
//                      from a device queue; instances estimated at few groups
//                      are pre-aggregated per tile in LDS and flushed into
//                      their table, the others update their table directly
//   k_batch_second     variance: squared deviations from the group means,
//                      over the same work items
//   k_batch_count / k_batch_scan / k_batch_outputs / k_batch_valid_counts
//                      per-instance compaction into dense output rows
//
// Per row the semantics are those of the single group-by
// (groupby_kernels.hip; reference groupby/hash/groupby_kernels.cuh:98-105).
#include "batch.hpp"
#include "device_ops.hpp"

namespace mig {

constexpr int BBLOCK = 256;

__device__ __forceinline__ uint32_t b_key_hash(int32_t st, uint64_t bits)
{
  return hash_combine(0u, hash_fixed_bits(st, bits, 0u));
}

__device__ __forceinline__ int fk_width(int32_t k)
{
  switch (k) {
    case FK_SUM_F32:
    case FK_SUMSQ_F32:
    case FK_MIN_S32:
    case FK_MAX_S32:
    case FK_MIN_U32:
    case FK_MAX_U32:
    case FK_MIN_F32:
    case FK_MAX_F32:
    case FK_COUNT: return 4;
    default: return 8;
  }
}

// Accumulator element `slot` of field `f` (8-byte element slots).
__device__ __forceinline__ uint8_t* fa(field_desc const& F, int64_t slot)
{
  return static_cast<uint8_t*>(F.ptr) + slot * BATCH_ELEM;
}

__device__ __forceinline__ uint64_t read_kind(int32_t k, void const* a)
{
  return fk_width(k) == 4 ? uint64_t(*static_cast<uint32_t const*>(a)) : *static_cast<uint64_t const*>(a);
}

__device__ __forceinline__ double sum_as_double(int32_t k, void const* a)
{
  if (k == FK_SUM_I64) return double(*static_cast<int64_t const*>(a));
  if (k == FK_SUM_F32) return double(*static_cast<float const*>(a));
  return *static_cast<double const*>(a);
}

// Hash part of a row key for instances split over `parts` LDS passes
// (multiplicative range map of a remixed hash, independent of the low bits
// that place the key in the LDS and global tables).
__device__ __forceinline__ uint32_t part_of(uint32_t h, uint32_t parts)
{
  return uint32_t((uint64_t(mm_fmix32(h ^ 0x9e3779b9u)) * parts) >> 32);
}

// Key hash of a row (null keys hash as the reference hashes a null).
__device__ __forceinline__ uint32_t b_row_hash(batch_inst const& I, int64_t row)
{
  if (!dcol_valid(I.key, row)) return hash_combine(0u, NULL_HASH);
  return b_key_hash(I.key.storage, load_fixed_bits(I.key, row));
}

// Work item: (instance << 40) | (hash part << 32) | first row.
__device__ __forceinline__ void decode_item(uint64_t item, int& inst, uint32_t& part, int64_t& r0)
{
  inst = int(item >> 40);
  part = uint32_t(item >> 32) & 0xffu;
  r0   = int64_t(item & 0xffffffffull);
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BBLOCK)
  k_batch_sample(batch_params const* __restrict__ Bp, int32_t per_inst, uint64_t* tables, uint32_t table_slots,
                 uint32_t* counters)
{
  batch_params const& B = *Bp;
  int const i           = blockIdx.x;
  batch_inst const I    = B.inst[i];
  int64_t const ns      = I.rows < per_inst ? I.rows : per_inst;
  uint64_t* table       = tables + uint64_t(i) * table_slots;
  for (int64_t k = threadIdx.x; k < ns; k += BBLOCK) {
    int64_t const row = (k * I.rows) / ns;
    uint32_t h;
    if (!dcol_valid(I.key, row)) {
      if (B.skip_null_keys) continue;
      h = hash_combine(0u, NULL_HASH);
    } else {
      h = b_key_hash(I.key.storage, load_fixed_bits(I.key, row));
    }
    aadd(counters + 2 * i, 1u);
    uint64_t const tag = (1ull << 32) | h;
    uint32_t s         = mm_fmix32(h) & (table_slots - 1);
    for (uint32_t p = 0; p < table_slots; p++) {
      uint64_t cur = aload(table + s);
      if (cur == tag) break;
      if (cur == 0) {
        uint64_t old = acas64(table + s, 0ull, tag);
        if (old == 0) {
          aadd(counters + 2 * i + 1, 1u);
          break;
        }
        if (old == tag) break;
      }
      s = (s + 1) & (table_slots - 1);
    }
  }
}

void launch_batch_sample(batch_params const* dB, int32_t n_inst, int32_t per_inst, uint64_t* tables,
                         uint32_t table_slots, uint32_t* counters, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_sample, dim3(n_inst), dim3(BBLOCK), 0, s, dB, per_inst, tables, table_slots,
                     counters);
}

// ---------------------------------------------------------------------------
// One block per slot tile: keys, row counts and every field's identity of
// the tile's instance.
__global__ void __launch_bounds__(BBLOCK) k_batch_init_slots(batch_params const* __restrict__ Bp)
{
  batch_params const& B = *Bp;
  uint64_t const t      = B.slot_tiles[blockIdx.x];
  int const inst        = int(t >> 40);
  uint64_t const l0     = t & ((1ull << 40) - 1);
  batch_inst const I    = B.inst[inst];
  int32_t const* kinds  = B.inst_kinds + int64_t(inst) * B.num_fields;
  for (int k = threadIdx.x; k < SCAN_TILE; k += BBLOCK) {
    uint64_t const l = l0 + k;
    if (l >= I.cap + 2) break;
    int64_t const slot = int64_t(I.slot_base + l);
    B.slot_keys[slot]  = EMPTY_KEY;
    if (B.slot_count) B.slot_count[slot] = 0;
    for (int f = 0; f < B.num_fields; f++)
      *reinterpret_cast<uint64_t*>(fa(B.fields[f], slot)) = kinds[f] == FK_M2 ? 0ull : field_identity(kinds[f]);
  }
}

// Small per-instance state, the work queues and the output masks.
__global__ void __launch_bounds__(BBLOCK) k_batch_init_misc(batch_params const* __restrict__ Bp)
{
  batch_params const& B = *Bp;
  int64_t const stride  = int64_t(gridDim.x) * BBLOCK;
  int64_t const i0      = int64_t(blockIdx.x) * BBLOCK + threadIdx.x;
  for (int64_t j = i0; j < int64_t(B.n_inst); j += stride) {
    B.special[2 * j]     = 0;
    B.special[2 * j + 1] = 0;
    B.groups[j]          = 0;
    B.overflow[j]        = 0;
  }
  int64_t const nv = int64_t(B.n_inst) * (B.num_outs + 1);
  for (int64_t j = i0; j < nv; j += stride) B.valid_counts[j] = 0;
  if (i0 == 0) {
    B.queue[0] = 0;
    B.queue[1] = 0;
  }
  int64_t const words = int64_t((B.out_rows + 31) / 32);
  for (int o = 0; o < B.num_outs; o++)
    if (B.outs[o].mask)
      for (int64_t j = i0; j < words; j += stride) B.outs[o].mask[j] = 0;
  if (B.out_key_mask)
    for (int64_t j = i0; j < words; j += stride) B.out_key_mask[j] = 0;
}

void launch_batch_init(batch_params const* dB, batch_params const& hB, int num_cus, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_init_slots, dim3(hB.n_slot_tiles), dim3(BBLOCK), 0, s, dB);
  int64_t blocks = int64_t((hB.out_rows / 32 + BBLOCK - 1) / BBLOCK);
  if (blocks > int64_t(num_cus) * 4) blocks = int64_t(num_cus) * 4;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_batch_init_misc, dim3(blocks), dim3(BBLOCK), 0, s, dB);
}

// ---------------------------------------------------------------------------
// table of one instance
__device__ __forceinline__ int64_t b_locate(batch_params const& B, batch_inst const& I, int inst,
                                            uint64_t key, uint32_t h)
{
  if (key == EMPTY_KEY) {
    uint32_t* f = B.special + 2 * inst;
    if (aload(f) == 0u) astore(f, 1u);
    return int64_t(I.slot_base + I.cap);
  }
  uint64_t* keys   = B.slot_keys + I.slot_base;
  uint64_t const m = I.cap - 1;
  uint64_t s       = h & m;
  for (uint64_t p = 0; p < I.cap; p++) {
    uint64_t cur = aload(keys + s);
    if (cur == key) return int64_t(I.slot_base + s);
    if (cur == EMPTY_KEY) {
      uint64_t old = acas64(keys + s, EMPTY_KEY, key);
      if (old == EMPTY_KEY || old == key) return int64_t(I.slot_base + s);
    }
    s = (s + 1) & m;
  }
  astore(B.overflow + inst, 1);
  return -1;
}

// Slot of a key inserted by the first pass (-1 if absent: only after an
// overflow, whose results are discarded).
__device__ __forceinline__ int64_t b_find(batch_params const& B, batch_inst const& I, uint64_t key, uint32_t h)
{
  if (key == EMPTY_KEY) return int64_t(I.slot_base + I.cap);
  uint64_t const* keys = B.slot_keys + I.slot_base;
  uint64_t const m     = I.cap - 1;
  uint64_t s           = h & m;
  for (uint64_t p = 0; p < I.cap; p++) {
    uint64_t const cur = keys[s];
    if (cur == key) return int64_t(I.slot_base + s);
    if (cur == EMPTY_KEY) return -1;
    s = (s + 1) & m;
  }
  return -1;
}

__device__ __forceinline__ int64_t b_null_slot(batch_params const& B, batch_inst const& I, int inst)
{
  uint32_t* f = B.special + 2 * inst + 1;
  if (aload(f) == 0u) astore(f, 1u);
  return int64_t(I.slot_base + I.cap + 1);
}

__device__ __forceinline__ void b_update_global(batch_params const& B, int32_t const* kinds, int inst,
                                                int64_t slot, int64_t row)
{
  if (B.need_count_all) aadd(B.slot_count + slot, 1u);
  for (int v = 0; v < B.num_vals; v++) {
    int const fb = B.val_field_begin[v], fe = B.val_field_begin[v + 1];
    if (fb == fe) continue;
    dcol const& c = B.vals[int64_t(inst) * B.num_vals + v];
    if (!dcol_valid(c, row)) continue;
    loaded_value const val = load_value(c, row);
    for (int f = fb; f < fe; f++) acc_apply(kinds[f], fa(B.fields[f], slot), val);
  }
}

// Pull the next work item of queue `q` into `item_sh` (all threads), and
// the instance's accumulator kinds into `kinds_sh`.  Returns false when the
// queue is drained.
__device__ __forceinline__ bool next_item(batch_params const& B, int q, uint64_t& item_sh, int32_t* kinds_sh,
                                          int nt)
{
  if (threadIdx.x == 0) {
    uint32_t it = __hip_atomic_fetch_add(B.queue + q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    item_sh     = it < uint64_t(B.n_items) ? B.items[it] : ~0ull;
  }
  __syncthreads();
  uint64_t const item = item_sh;
  if (item == ~0ull) return false;
  int const inst = int(item >> 40);
  for (int f = threadIdx.x; f < B.num_fields; f += nt) kinds_sh[f] = B.inst_kinds[int64_t(inst) * B.num_fields + f];
  __syncthreads();
  return true;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_batch_aggregate(batch_params const* __restrict__ Bp)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ uint64_t item_sh;
  __shared__ uint32_t fill_sh, present_sh[2];
  __shared__ int32_t kinds_sh[MAX_FIELDS];
  batch_params const& B = *Bp;
  constexpr int S       = BATCH_LDS_SLOTS;
  constexpr int WG      = __HIP_MEMORY_SCOPE_WORKGROUP;
  auto* tkeys           = reinterpret_cast<uint64_t*>(lds);
  while (next_item(B, 0, item_sh, kinds_sh, NT)) {
    int inst;
    uint32_t part;
    int64_t r0;
    decode_item(item_sh, inst, part, r0);
    batch_inst const I  = B.inst[inst];
    int64_t const r1    = (r0 + B.tile_rows < I.rows) ? r0 + B.tile_rows : I.rows;

    if (!I.lds) {
      // direct updates of the instance table, BATCH_CHUNK rows per thread at
      // a time (their key loads in flight together, then their slots, then
      // their updates)
      for (int64_t c0 = r0 + threadIdx.x; c0 < r1; c0 += int64_t(BATCH_CHUNK) * NT) {
        uint64_t bits[BATCH_CHUNK];
        bool kv[BATCH_CHUNK];
#pragma unroll
        for (int k = 0; k < BATCH_CHUNK; k++) {  // rows past the tile read its last row
          int64_t const row = c0 + int64_t(k) * NT;
          int64_t const rc  = row < r1 ? row : r1 - 1;
          kv[k]             = dcol_valid(I.key, rc);
          bits[k]           = load_fixed_bits(I.key, rc);
        }
        int64_t slot[BATCH_CHUNK];
#pragma unroll
        for (int k = 0; k < BATCH_CHUNK; k++) {
          slot[k] = -1;
          if (c0 + int64_t(k) * NT >= r1) continue;
          if (!kv[k]) {
            if (!B.skip_null_keys) slot[k] = b_null_slot(B, I, inst);
          } else {
            slot[k] = b_locate(B, I, inst, bits[k], b_key_hash(I.key.storage, bits[k]));
          }
        }
        if (B.need_count_all) {
#pragma unroll
          for (int k = 0; k < BATCH_CHUNK; k++)
            if (slot[k] >= 0) aadd(B.slot_count + slot[k], 1u);
        }
        for (int v = 0; v < B.num_vals; v++) {
          int const fb = B.val_field_begin[v], fe = B.val_field_begin[v + 1];
          if (fb == fe) continue;
          dcol const cv = B.vals[int64_t(inst) * B.num_vals + v];
          loaded_value val[BATCH_CHUNK];
          uint32_t ok = 0;
#pragma unroll
          for (int k = 0; k < BATCH_CHUNK; k++) {
            val[k]            = loaded_value{0, 0.0, 0.0f, 0};
            int64_t const row = c0 + int64_t(k) * NT;
            if (slot[k] >= 0 && dcol_valid(cv, row)) {
              val[k] = load_value(cv, row);
              ok |= 1u << k;
            }
          }
          for (int f = fb; f < fe; f++) {
            int32_t const kind = kinds_sh[f];
#pragma unroll
            for (int k = 0; k < BATCH_CHUNK; k++)
              if ((ok >> k) & 1u) acc_apply(kind, fa(B.fields[f], slot[k]), val[k]);
          }
        }
      }
      __syncthreads();
      continue;
    }

    // ---- LDS pre-aggregation of the tile (rows S, S+1: EMPTY_KEY, null key)
    for (int i = threadIdx.x; i < S; i += NT) tkeys[i] = EMPTY_KEY;
    for (int f = 0; f < B.num_fields; f++) {
      int32_t const off = B.lds_field_off[f];
      if (off < 0) continue;
      uint64_t const idv = field_identity(kinds_sh[f]);
      for (int i = threadIdx.x; i < S + 2; i += NT) *lds_at<uint64_t>(lds, off, i) = idv;
    }
    if (B.lds_count_off >= 0)
      for (int i = threadIdx.x; i < S + 2; i += NT) *lds_at<uint32_t>(lds, B.lds_count_off, i) = 0;
    if (threadIdx.x == 0) {
      fill_sh       = 0;
      present_sh[0] = 0;
      present_sh[1] = 0;
    }
    __syncthreads();
    uint32_t const parts = uint32_t(I.lds);
    // The item's field plan, one field per lane of a VGPR (kind | (LDS offset
    // + 1) << 8), fetched with v_readlane: reading descriptors from memory
    // inside the row loop would wait for the LDS atomics in flight each time.
    int const lane         = int(threadIdx.x & 63);
    uint32_t const fd_lane = lane < B.num_fields
                               ? uint32_t(kinds_sh[lane]) | (uint32_t(B.lds_field_off[lane] + 1) << 8)
                               : 0u;
    // BATCH_CHUNK rows per thread at a time: their key loads, then their
    // LDS slots, then each value column's loads, in flight together
    for (int64_t c0 = r0 + threadIdx.x; c0 < r1; c0 += int64_t(BATCH_CHUNK) * NT) {
      uint64_t bits[BATCH_CHUNK];
      bool kv[BATCH_CHUNK];
      int id[BATCH_CHUNK];
      uint32_t hh[BATCH_CHUNK];
#pragma unroll
      for (int k = 0; k < BATCH_CHUNK; k++) {  // rows past the tile read its last row
        int64_t const row = c0 + int64_t(k) * NT;
        int64_t const rc  = row < r1 ? row : r1 - 1;
        kv[k]             = dcol_valid(I.key, rc);
        bits[k]           = load_fixed_bits(I.key, rc);
      }
#pragma unroll
      for (int k = 0; k < BATCH_CHUNK; k++) {
        int64_t const row = c0 + int64_t(k) * NT;
        id[k]             = -2;  // no part in this item
        hh[k]             = 0;
        if (row >= r1) continue;
        if (!kv[k]) {
          if (B.skip_null_keys) continue;
          if (parts > 1 && part_of(hash_combine(0u, NULL_HASH), parts) != part) continue;
          id[k]         = S + 1;
          present_sh[1] = 1;
        } else {
          uint32_t const h = b_key_hash(I.key.storage, bits[k]);
          if (parts > 1 && part_of(h, parts) != part) continue;  // another item's part
          hh[k] = h;
          if (bits[k] == EMPTY_KEY) {
            id[k]         = S;
            present_sh[0] = 1;
          } else {
            id[k] = lds_insert(tkeys, &fill_sh, bits[k], h, EMPTY_KEY, S, BATCH_LDS_FILL);
          }
        }
      }
      uint32_t live = 0;
#pragma unroll
      for (int k = 0; k < BATCH_CHUNK; k++) {
        live |= uint32_t(id[k] >= 0) << k;
        if (id[k] == -1) {  // the LDS table is full: this row goes straight to the instance table
          int64_t const slot = b_locate(B, I, inst, bits[k], hh[k]);
          if (slot >= 0) b_update_global(B, kinds_sh, inst, slot, c0 + int64_t(k) * NT);
        }
      }
      if (live == 0) continue;
      if (B.lds_count_off >= 0) {
#pragma unroll
        for (int k = 0; k < BATCH_CHUNK; k++)
          if ((live >> k) & 1u)
            __hip_atomic_fetch_add(lds_at<uint32_t>(lds, B.lds_count_off, id[k]), 1u, __ATOMIC_RELAXED, WG);
      }
      for (int v = 0; v < B.num_vals; v++) {
        int const fb = B.val_field_begin[v], fe = B.val_field_begin[v + 1];
        if (fb == fe) continue;
        dcol const cv = B.vals[int64_t(inst) * B.num_vals + v];
        loaded_value val[BATCH_CHUNK];
        uint32_t ok = live;
#pragma unroll
        for (int k = 0; k < BATCH_CHUNK; k++) {
          val[k] = loaded_value{0, 0.0, 0.0f, 0};
          if (!((live >> k) & 1u)) continue;
          int64_t const row = c0 + int64_t(k) * NT;
          if (dcol_valid(cv, row)) val[k] = load_value(cv, row);
          else ok &= ~(1u << k);
        }
        for (int f = fb; f < fe; f++) {
          uint32_t const d = uint32_t(__builtin_amdgcn_readlane(int(fd_lane), f));
          int const off    = int(d >> 8) - 1;
          if (off < 0) continue;
          int32_t const kind = int32_t(d & 0xffu);
#pragma unroll
          for (int k = 0; k < BATCH_CHUNK; k++)
            if ((ok >> k) & 1u) lds_apply(kind, lds_at<uint64_t>(lds, off, id[k]), val[k]);
        }
      }
    }
    __syncthreads();
    // ---- flush the tile's groups into the instance table
    for (int i = threadIdx.x; i < S + 2; i += NT) {
      int64_t slot;
      if (i < S) {
        uint64_t const k = tkeys[i];
        if (k == EMPTY_KEY) continue;
        slot = b_locate(B, I, inst, k, b_key_hash(I.key.storage, k));
      } else if (i == S) {
        if (!present_sh[0]) continue;
        slot = b_locate(B, I, inst, EMPTY_KEY, 0);
      } else {
        if (!present_sh[1]) continue;
        slot = b_null_slot(B, I, inst);
      }
      if (slot < 0) continue;
      if (B.lds_count_off >= 0) aadd(B.slot_count + slot, *lds_at<uint32_t>(lds, B.lds_count_off, i));
      for (int f = 0; f < B.num_fields; f++) {
        if (B.lds_field_off[f] < 0) continue;
        int32_t const k     = kinds_sh[f];
        uint64_t const part = read_kind(k, lds_at<uint64_t>(lds, B.lds_field_off[f], i));
        // an identity partial leaves the accumulator as it is (the float
        // sums start at +0.0, which adding +0.0 keeps)
        if (part != field_identity(k)) acc_combine(k, fa(B.fields[f], slot), part);
      }
    }
    __syncthreads();
  }
}

void prepare_batch_aggregate(int32_t lds_bytes);

void launch_batch_aggregate(batch_params const* dB, batch_params const& hB, int num_cus, int block,
                            hipStream_t s)
{
  int64_t blocks = int64_t(num_cus) * (block == 512 ? 2 : 4);
  if (blocks > hB.n_items) blocks = hB.n_items;
  if (blocks < 1) blocks = 1;
  if (block == 512)
    hipLaunchKernelGGL(k_batch_aggregate<512>, dim3(blocks), dim3(512), hB.lds_bytes, s, dB);
  else
    hipLaunchKernelGGL(k_batch_aggregate<256>, dim3(blocks), dim3(256), hB.lds_bytes, s, dB);
}

// ---------------------------------------------------------------------------
// second pass: squared deviations from the group mean (FK_M2 fields)

__device__ __forceinline__ uint32_t b_count_of(batch_params const& B, int32_t count_field, int64_t slot)
{
  if (count_field == -2) return B.slot_count[slot];
  return *reinterpret_cast<uint32_t const*>(fa(B.fields[count_field], slot));
}

// The group slot of `row` for the second pass (-1: row dropped or absent).
__device__ __forceinline__ int64_t b_row_slot(batch_params const& B, batch_inst const& I, int64_t row)
{
  if (!dcol_valid(I.key, row)) return B.skip_null_keys ? -1 : int64_t(I.slot_base + I.cap + 1);
  uint64_t const bits = load_fixed_bits(I.key, row);
  return b_find(B, I, bits, b_key_hash(I.key.storage, bits));
}

// Deviation term of row `row` for M2 field `f` at `slot`.
__device__ __forceinline__ double b_dev2(batch_params const& B, int32_t const* kinds, int f, int64_t slot,
                                         loaded_value const& val)
{
  field_desc const& F = B.fields[f];
  double const cnt    = double(b_count_of(B, F.aux_count, slot));
  double const mean   = sum_as_double(kinds[F.aux_sum], fa(B.fields[F.aux_sum], slot)) / cnt;
  double const d      = val.d - mean;
  return d * d;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_batch_second(batch_params const* __restrict__ Bp)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ uint64_t item_sh;
  __shared__ uint32_t fill_sh;
  __shared__ int32_t kinds_sh[MAX_FIELDS];
  batch_params const& B = *Bp;
  constexpr int S       = BATCH_LDS_SLOTS;
  constexpr int WG      = __HIP_MEMORY_SCOPE_WORKGROUP;
  constexpr uint64_t NO = ~0ull;  // empty marker of the LDS slot table
  auto* tslots          = reinterpret_cast<uint64_t*>(lds);
  while (next_item(B, 1, item_sh, kinds_sh, NT)) {
    int inst;
    uint32_t part;
    int64_t r0;
    decode_item(item_sh, inst, part, r0);
    batch_inst const I  = B.inst[inst];
    int64_t const r1    = (r0 + B.tile_rows < I.rows) ? r0 + B.tile_rows : I.rows;
    bool const use_lds  = I.lds != 0;
    if (use_lds) {
      for (int i = threadIdx.x; i < S; i += NT) tslots[i] = NO;
      for (int f = 0; f < B.num_fields; f++) {
        if (B.fields[f].kind != FK_M2) continue;
        for (int i = threadIdx.x; i < S; i += NT) *lds_at<double>(lds, B.lds_m2_off[f], i) = 0.0;
      }
      if (threadIdx.x == 0) fill_sh = 0;
      __syncthreads();
    }
    uint32_t const parts = uint32_t(I.lds);
    for (int64_t row = r0 + threadIdx.x; row < r1; row += NT) {
      if (parts > 1 && part_of(b_row_hash(I, row), parts) != part) continue;  // another item's part
      int64_t slot = -2;
      int id       = -2;
      for (int v = 0; v < B.num_vals; v++) {
        int const fb = B.val_field_begin[v], fe = B.val_field_begin[v + 1];
        bool has_m2  = false;
        for (int f = fb; f < fe; f++) has_m2 |= (B.fields[f].kind == FK_M2);
        if (!has_m2) continue;
        dcol const& c = B.vals[int64_t(inst) * B.num_vals + v];
        if (!dcol_valid(c, row)) continue;
        if (slot == -2) {
          slot = b_row_slot(B, I, row);
          if (slot >= 0 && use_lds) id = lds_insert(tslots, &fill_sh, uint64_t(slot), mm_fmix32(uint32_t(slot)), NO, S, BATCH_LDS_FILL);
        }
        if (slot < 0) break;
        loaded_value const val = load_value(c, row);
        for (int f = fb; f < fe; f++) {
          if (B.fields[f].kind != FK_M2) continue;
          double const t = b_dev2(B, kinds_sh, f, slot, val);
          if (id >= 0) __hip_atomic_fetch_add(lds_at<double>(lds, B.lds_m2_off[f], id), t, __ATOMIC_RELAXED, WG);
          else gadd_f64(reinterpret_cast<double*>(fa(B.fields[f], slot)), t);
        }
      }
    }
    __syncthreads();
    if (use_lds) {
      for (int i = threadIdx.x; i < S; i += NT) {
        uint64_t const slot = tslots[i];
        if (slot == NO) continue;
        for (int f = 0; f < B.num_fields; f++) {
          if (B.fields[f].kind != FK_M2) continue;
          double const part = *lds_at<double>(lds, B.lds_m2_off[f], i);
          if (part != 0.0) gadd_f64(reinterpret_cast<double*>(fa(B.fields[f], int64_t(slot))), part);
        }
      }
      __syncthreads();
    }
  }
}

void launch_batch_second_pass(batch_params const* dB, batch_params const& hB, int num_cus, int block,
                              hipStream_t s)
{
  if (!hB.has_m2) return;
  int64_t blocks = int64_t(num_cus) * (block == 512 ? 2 : 4);
  if (blocks > hB.n_items) blocks = hB.n_items;
  if (blocks < 1) blocks = 1;
  if (block == 512)
    hipLaunchKernelGGL(k_batch_second<512>, dim3(blocks), dim3(512), hB.lds_bytes, s, dB);
  else
    hipLaunchKernelGGL(k_batch_second<256>, dim3(blocks), dim3(256), hB.lds_bytes, s, dB);
}

// ---------------------------------------------------------------------------
// compaction: occupied slots of each instance, in slot order
__device__ __forceinline__ bool b_occupied(batch_params const& B, batch_inst const& I, int inst, uint64_t local)
{
  if (local < I.cap) return B.slot_keys[I.slot_base + local] != EMPTY_KEY;
  return B.special[2 * inst + (local - I.cap)] != 0u;
}

__global__ void __launch_bounds__(BBLOCK) k_batch_count(batch_params const* __restrict__ Bp)
{
  __shared__ uint32_t red[BBLOCK / 64];
  batch_params const& B = *Bp;
  uint64_t const t      = B.slot_tiles[blockIdx.x];
  int const inst        = int(t >> 40);
  uint64_t const l0     = t & ((1ull << 40) - 1);
  batch_inst const I    = B.inst[inst];
  uint32_t c            = 0;
  for (int k = threadIdx.x; k < SCAN_TILE; k += BBLOCK) {
    uint64_t const l = l0 + k;
    if (l < I.cap + 2 && b_occupied(B, I, inst, l)) c++;
  }
  c = block_sum_u32<BBLOCK>(c, red);
  if (threadIdx.x == 0) B.tile_offsets[blockIdx.x] = c;
}

// One block per instance: exclusive scan of its tiles' counts (BBLOCK tiles
// per step: wave64 prefix sums, then across the block's waves), total groups.
__global__ void __launch_bounds__(BBLOCK) k_batch_scan(batch_params const* __restrict__ Bp)
{
  __shared__ uint32_t wave_tot[BBLOCK / 64];
  batch_params const& B = *Bp;
  int const inst        = blockIdx.x;
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int64_t const t0 = B.inst_tile_begin[inst], t1 = B.inst_tile_begin[inst + 1];
  uint32_t run = 0;
  for (int64_t base = t0; base < t1; base += BBLOCK) {
    int64_t const t  = base + threadIdx.x;
    uint32_t const c = t < t1 ? B.tile_offsets[t] : 0u;
    uint32_t incl    = c;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t const up = __shfl_up(incl, o, 64);
      if (lane >= o) incl += up;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < BBLOCK / 64; w++) {
      if (w < wave) before += wave_tot[w];
      all += wave_tot[w];
    }
    if (t < t1) B.tile_offsets[t] = run + before + incl - c;
    run += all;
    __syncthreads();
  }
  if (threadIdx.x == 0) B.groups[inst] = run;
}

__global__ void __launch_bounds__(BBLOCK) k_batch_outputs(batch_params const* __restrict__ Bp)
{
  __shared__ uint32_t wave_tot[BBLOCK / 64];
  batch_params const& B = *Bp;
  uint64_t const t      = B.slot_tiles[blockIdx.x];
  int const inst        = int(t >> 40);
  uint64_t const l0     = t & ((1ull << 40) - 1);
  batch_inst const I    = B.inst[inst];
  uint32_t running      = B.tile_offsets[blockIdx.x];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int32_t const* kinds  = B.inst_kinds + int64_t(inst) * B.num_fields;
  int32_t const* widths = B.inst_widths + int64_t(inst) * (B.num_outs + 1);
  uint64_t const rbase  = I.out_base * BATCH_ELEM;  // byte offset of the instance's regions
  for (int k = 0; k < SCAN_TILE; k += BBLOCK) {
    uint64_t const l     = l0 + k + threadIdx.x;
    bool const occ       = l < I.cap + 2 && b_occupied(B, I, inst, l);
    uint64_t const bal   = __ballot(occ);
    uint32_t const below = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(bal);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < BBLOCK / 64; w++) {
      if (w < wave) before += wave_tot[w];
      all += wave_tot[w];
    }
    if (occ) {
      uint64_t const lp  = running + before + below;  // row within the instance
      uint64_t const pos = I.out_base + lp;           // row of the concatenated masks
      int64_t const slot = int64_t(I.slot_base + l);
      bool const kvalid  = l != I.cap + 1;
      uint64_t const key = l < I.cap ? B.slot_keys[slot] : (l == I.cap ? EMPTY_KEY : 0ull);
      store_width(static_cast<uint8_t*>(B.out_key) + rbase, int64_t(lp), widths[B.num_outs], kvalid ? key : 0ull);
      if (B.out_key_mask && kvalid)
        __hip_atomic_fetch_or(B.out_key_mask + (pos >> 5), 1u << (pos & 31), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
      for (int o = 0; o < B.num_outs; o++) {
        out_desc const& O = B.outs[o];
        uint8_t* const dst = static_cast<uint8_t*>(O.data) + rbase;
        int const w        = widths[o];
        bool valid         = true;
        uint32_t cnt       = 0;
        if (O.count_field != -1) {
          cnt   = b_count_of(B, O.count_field, slot);
          valid = cnt > 0;
        }
        switch (O.kind) {
          case OK_COPY_I64:
          case OK_COPY_F32:
          case OK_COPY_F64:
            store_width(dst, int64_t(lp), w, *reinterpret_cast<uint64_t const*>(fa(B.fields[O.field], slot)));
            break;
          case OK_COUNT:
            store_width(dst, int64_t(lp), 4, *reinterpret_cast<uint32_t const*>(fa(B.fields[O.field], slot)));
            valid = true;
            break;
          case OK_COUNT_ALL:
            store_width(dst, int64_t(lp), 4, B.slot_count[slot]);
            valid = true;
            break;
          case OK_MINMAX: {
            int32_t const fk = kinds[O.field];
            uint64_t bits    = read_kind(fk, fa(B.fields[O.field], slot));
            if (fk == FK_MIN_F32 || fk == FK_MAX_F32) bits = unord32(uint32_t(bits));
            else if (fk == FK_MIN_F64 || fk == FK_MAX_F64) bits = unord64(bits);
            store_width(dst, int64_t(lp), w, bits);
            break;
          }
          case OK_MEAN: {
            double const sum = sum_as_double(kinds[O.field], fa(B.fields[O.field], slot));
            reinterpret_cast<double*>(dst)[lp] = valid ? sum / double(cnt) : 0.0;
            break;
          }
          case OK_VAR:
          case OK_STD: {
            int64_t const dn = int64_t(cnt) - O.ddof;
            valid            = cnt > 0 && dn > 0;
            double var       = 0.0;
            if (valid) var = *reinterpret_cast<double const*>(fa(B.fields[O.field], slot)) / double(dn);
            reinterpret_cast<double*>(dst)[lp] = O.kind == OK_STD ? sqrt(var) : var;
            break;
          }
          default: break;
        }
        if (O.mask && valid)
          __hip_atomic_fetch_or(O.mask + (pos >> 5), 1u << (pos & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    running += all;
    __syncthreads();
  }
}

void launch_batch_count(batch_params const* dB, batch_params const& hB, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_count, dim3(hB.n_slot_tiles), dim3(BBLOCK), 0, s, dB);
}
void launch_batch_scan(batch_params const* dB, batch_params const& hB, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_scan, dim3(hB.n_inst), dim3(BBLOCK), 0, s, dB);
}
void launch_batch_outputs(batch_params const* dB, batch_params const& hB, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_outputs, dim3(hB.n_slot_tiles), dim3(BBLOCK), 0, s, dB);
}

// Set bits of each instance's region of each output mask (grid: instances x
// (outputs + key)); the host turns them into null counts.
__global__ void __launch_bounds__(BBLOCK) k_batch_valid_counts(batch_params const* __restrict__ Bp)
{
  __shared__ uint32_t red[BBLOCK / 64];
  batch_params const& B = *Bp;
  int const inst        = blockIdx.x;
  int const o           = blockIdx.y;
  uint32_t const* mask  = o < B.num_outs ? B.outs[o].mask : B.out_key_mask;
  if (mask == nullptr) return;
  uint64_t const base = B.inst[inst].out_base;  // multiple of 64
  uint32_t const G    = B.groups[inst];
  uint32_t c          = 0;
  for (uint32_t w = threadIdx.x; w < (G + 31) / 32; w += BBLOCK) {
    uint32_t bits = mask[base / 32 + w];
    if (w == G / 32) bits &= (1u << (G & 31)) - 1u;
    c += __popc(bits);
  }
  c = block_sum_u32<BBLOCK>(c, red);
  if (threadIdx.x == 0) B.valid_counts[int64_t(inst) * (B.num_outs + 1) + o] = c;
}

void launch_batch_valid_counts(batch_params const* dB, batch_params const& hB, hipStream_t s)
{
  hipLaunchKernelGGL(k_batch_valid_counts, dim3(hB.n_inst, hB.num_outs + 1), dim3(BBLOCK), 0, s, dB);
}

void prepare_batch_aggregate(int32_t lds_bytes)
{
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_batch_aggregate<256>), lds_bytes);
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_batch_aggregate<512>), lds_bytes);
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_batch_second<256>), lds_bytes);
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_batch_second<512>), lds_bytes);
}

}  // namespace mig
