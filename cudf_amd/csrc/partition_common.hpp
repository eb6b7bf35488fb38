// Device helpers shared by the radix-partitioned aggregation kernels.
#pragma once

#include "device_ops.hpp"

namespace mig {

// Workgroup barrier for phases that communicate through LDS only: each wave
// waits for its LDS operations, not for its global stores (a __syncthreads
// would also wait until every store of the workgroup is acknowledged).
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t inline_key_hash(int32_t st, uint64_t bits)
{
  return hash_combine(0u, hash_fixed_bits(st, bits, 0u));
}
__device__ __forceinline__ uint32_t null_key_hash() { return hash_combine(0u, NULL_HASH); }

// Row hash of a partition record (key word and flags byte).
__device__ __forceinline__ uint32_t record_hash(gb_params const& P, uint64_t kb, uint8_t kf)
{
  if (P.key_mode == KM_GENERAL) return uint32_t(kb >> 32);
  return (kf & RPA_KEY_NULL) ? null_key_hash() : inline_key_hash(P.key_storage, kb);
}

__device__ __forceinline__ uint32_t top_bits(uint32_t h, int nbits)
{
  return nbits == 0 ? 0u : (h >> (32 - nbits));
}

// Seed of the second, independent row hash of the hashed general-key form.
constexpr uint32_t SECOND_ROW_HASH_SEED = 0x2f3b9a57u;

// Partition word of a row under general keys (several columns, strings):
// its row hash and its row index.  Equal keys are found by comparing rows.
__device__ __forceinline__ uint64_t general_word(uint32_t h, int64_t row)
{
  return (uint64_t(h) << 32) | uint64_t(uint32_t(row));
}

// Hash of input row `row`; returns false when the row is dropped (a null key
// under EXCLUDE).  `bits` receives the partition word (the normalised key,
// or general_word), `key_null` whether a single key is null (INCLUDE).
__device__ __forceinline__ bool input_key(gb_params const& P, int64_t row, uint64_t& bits,
                                          bool& key_null, uint32_t& h)
{
  if (P.key_mode == KM_GENERAL) {
    if (P.skip_null_keys && any_key_null(P, row)) return false;
    h        = row_hash(P.keys, P.num_keys, row, 0u);
    bits     = general_word(h, row);
    key_null = false;
    return true;
  }
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

}  // namespace mig
