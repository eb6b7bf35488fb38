// MurmurHash3_x86_32 (Austin Appleby's public-domain algorithm) and the boost
// hash_combine used by libcudf's row hasher.
//   element hash : cudf/hashing/detail/murmurhash3_x86_32.cuh:40-130
//   normalisation: cudf/hashing/detail/hash_functions.cuh:28-47
//   combine      : cudf/hashing/detail/hashing.hpp:85-88
//   row hash     : cudf/table/experimental/row_operators.cuh:1869-1883
#pragma once

#include "common.hpp"

namespace mig {

constexpr uint32_t MM_C1 = 0xcc9e2d51u;
constexpr uint32_t MM_C2 = 0x1b873593u;
constexpr uint32_t MM_C3 = 0xe6546b64u;
constexpr uint32_t NULL_HASH = 0xFFFFFFFFu;

__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, int r)
{
  return (x << r) | (x >> (32 - r));
}

__host__ __device__ __forceinline__ uint32_t mm_fmix32(uint32_t h)
{
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__host__ __device__ __forceinline__ uint32_t mm_mix_block(uint32_t h, uint32_t k)
{
  k *= MM_C1;
  k = rotl32(k, 15);
  k *= MM_C2;
  h ^= k;
  h = rotl32(h, 13);
  return h * 5 + MM_C3;
}

__host__ __device__ __forceinline__ uint32_t mm_tail(uint32_t h, uint32_t k)
{
  k *= MM_C1;
  k = rotl32(k, 15);
  k *= MM_C2;
  return h ^ k;
}

// Hash of a 4-byte value (one block, no tail).
__host__ __device__ __forceinline__ uint32_t mm_hash4(uint32_t v, uint32_t seed)
{
  uint32_t h = mm_mix_block(seed, v);
  return mm_fmix32(h ^ 4u);
}

// Hash of an 8-byte value: two little-endian blocks.
__host__ __device__ __forceinline__ uint32_t mm_hash8(uint64_t v, uint32_t seed)
{
  uint32_t h = mm_mix_block(seed, static_cast<uint32_t>(v));
  h = mm_mix_block(h, static_cast<uint32_t>(v >> 32));
  return mm_fmix32(h ^ 8u);
}

// Hash of a 1- or 2-byte value: tail only.
__host__ __device__ __forceinline__ uint32_t mm_hash_small(uint32_t v, int len, uint32_t seed)
{
  uint32_t h = mm_tail(seed, v);
  return mm_fmix32(h ^ static_cast<uint32_t>(len));
}

// Hash of `len` bytes at `p` (any alignment).  Word loads are used when the
// pointer is 4-byte aligned.
__host__ __device__ inline uint32_t mm_hash_bytes(uint8_t const* p, int64_t len, uint32_t seed)
{
  uint32_t h = seed;
  int64_t const nblocks = len >> 2;
  if ((reinterpret_cast<uintptr_t>(p) & 3) == 0) {
    uint32_t const* w = reinterpret_cast<uint32_t const*>(p);
    for (int64_t i = 0; i < nblocks; i++) h = mm_mix_block(h, w[i]);
  } else {
    for (int64_t i = 0; i < nblocks; i++) {
      uint8_t const* b = p + 4 * i;
      uint32_t k = uint32_t(b[0]) | (uint32_t(b[1]) << 8) | (uint32_t(b[2]) << 16) |
                   (uint32_t(b[3]) << 24);
      h = mm_mix_block(h, k);
    }
  }
  uint8_t const* t = p + 4 * nblocks;
  int const rem = static_cast<int>(len & 3);
  if (rem) {
    uint32_t k = 0;
    if (rem == 3) k ^= uint32_t(t[2]) << 16;
    if (rem >= 2) k ^= uint32_t(t[1]) << 8;
    k ^= t[0];
    h = mm_tail(h, k);
  }
  return mm_fmix32(h ^ static_cast<uint32_t>(len));
}

// MurmurHash3_x86_32 of the same bytes under two seeds, in one pass: the
// mixing of each input block is shared, each hash keeps its own state.
__host__ __device__ inline void mm_hash_bytes2(uint8_t const* p, int64_t len, uint32_t s1, uint32_t s2, uint32_t& o1,
                                               uint32_t& o2)
{
  uint32_t h1 = s1, h2 = s2;
  int64_t const nblocks = len >> 2;
  bool const aligned    = (reinterpret_cast<uintptr_t>(p) & 3) == 0;  // word loads when 4-byte aligned
  for (int64_t i = 0; i < nblocks; i++) {
    uint8_t const* b = p + 4 * i;
    uint32_t k = aligned ? reinterpret_cast<uint32_t const*>(p)[i]
                         : uint32_t(b[0]) | (uint32_t(b[1]) << 8) | (uint32_t(b[2]) << 16) | (uint32_t(b[3]) << 24);
    k *= MM_C1;
    k = rotl32(k, 15);
    k *= MM_C2;
    h1 ^= k;
    h2 ^= k;
    h1 = rotl32(h1, 13) * 5 + MM_C3;
    h2 = rotl32(h2, 13) * 5 + MM_C3;
  }
  uint8_t const* t = p + 4 * nblocks;
  int const rem = static_cast<int>(len & 3);
  if (rem) {
    uint32_t k = 0;
    if (rem == 3) k ^= uint32_t(t[2]) << 16;
    if (rem >= 2) k ^= uint32_t(t[1]) << 8;
    k ^= t[0];
    k *= MM_C1;
    k = rotl32(k, 15);
    k *= MM_C2;
    h1 ^= k;
    h2 ^= k;
  }
  o1 = mm_fmix32(h1 ^ static_cast<uint32_t>(len));
  o2 = mm_fmix32(h2 ^ static_cast<uint32_t>(len));
}

__host__ __device__ __forceinline__ uint32_t hash_combine(uint32_t lhs, uint32_t rhs)
{
  return lhs ^ (rhs + 0x9e3779b9u + (lhs << 6) + (lhs >> 2));
}

// -0.0 -> 0.0 and every NaN -> the canonical quiet NaN, on the bit patterns.
__host__ __device__ __forceinline__ uint32_t normalize_f32_bits(uint32_t b)
{
  if ((b & 0x7fffffffu) == 0) return 0u;
  if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu)) return 0x7fc00000u;
  return b;
}
__host__ __device__ __forceinline__ uint64_t normalize_f64_bits(uint64_t b)
{
  if ((b & 0x7fffffffffffffffull) == 0) return 0ull;
  if ((b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull))
    return 0x7ff8000000000000ull;
  return b;
}

}  // namespace mig
