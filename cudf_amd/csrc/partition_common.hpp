// Device helpers shared by the radix-partitioned aggregation kernels.
#pragma once

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

}  // namespace mig
