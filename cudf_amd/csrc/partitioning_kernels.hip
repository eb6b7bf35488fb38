// Kernels of hash partitioning and table concatenation (partitioning.cpp).
//
// hash_partition (cpp/src/partitioning/partitioning.cu:469-711) assigns each
// row the partition `row_hash % num_partitions` and regroups the rows so that
// each partition is contiguous.  Here the regrouping is a stable LSD counting
// sort of the partition numbers, HP_DIGIT_BITS bits per pass:
//   k_hp_count    per-wave digit histograms, written digit-major so that one
//                 exclusive scan of the whole array gives every wave's first
//                 output position for every digit;
//   k_hp_scatter  every wave replays its rows in order; rows of one digit in
//                 a 64-row chunk are ranked with ballots, so the order inside
//                 a partition is the input row order (deterministic output).
// One pass covers up to 2048 partitions.  Each wave owns a contiguous range
// of rows and its own LDS counters, so no workgroup barrier orders waves.
#include "device_ops.hpp"
#include "partitioning.hpp"

namespace mig {

namespace {

constexpr int HP_BLOCK = 256;
constexpr int HP_WAVES = HP_BLOCK / 64;

__device__ __forceinline__ uint64_t lanes_below()
{
  return (1ull << (threadIdx.x & 63)) - 1ull;
}

// IdentityHash of one element (partitioning.cu:738-758): the value converted
// to uint32.  Integers wrap modulo 2^32; floats convert with truncation and
// saturation (NaN and negatives give 0), as the hardware conversion does.
__device__ __forceinline__ uint32_t identity_elem(dcol const& c, int64_t row)
{
  if (!dcol_valid(c, row)) return NULL_HASH;
  auto d = gp(c.data);
  switch (c.storage) {
    case ST_I8: return uint32_t(int32_t(reinterpret_cast<gmem<int8_t const>*>(d)[row]));
    case ST_I16: return uint32_t(int32_t(reinterpret_cast<gmem<int16_t const>*>(d)[row]));
    case ST_I32: return uint32_t(reinterpret_cast<gmem<int32_t const>*>(d)[row]);
    case ST_I64: return uint32_t(uint64_t(reinterpret_cast<gmem<int64_t const>*>(d)[row]));
    case ST_U8:
    case ST_BOOL: return uint32_t(reinterpret_cast<gmem<uint8_t const>*>(d)[row]);
    case ST_U16: return uint32_t(reinterpret_cast<gmem<uint16_t const>*>(d)[row]);
    case ST_U32: return reinterpret_cast<gmem<uint32_t const>*>(d)[row];
    case ST_U64: return uint32_t(reinterpret_cast<gmem<uint64_t const>*>(d)[row]);
    case ST_F32:
    case ST_F64: {
      double const x = c.storage == ST_F32 ? double(reinterpret_cast<gmem<float const>*>(d)[row])
                                           : reinterpret_cast<gmem<double const>*>(d)[row];
      if (!(x > 0.0)) return 0u;
      if (x >= 4294967295.0) return 0xFFFFFFFFu;
      return uint32_t(x);
    }
    default: return 0u;
  }
}

__global__ void __launch_bounds__(HP_BLOCK)
  k_hp_partition_ids(dcol const* __restrict__ cols, int ncols, int64_t n, uint32_t seed, int identity,
                     uint32_t num_partitions, uint32_t* __restrict__ pid)
{
  int64_t const i = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (i >= n) return;
  uint32_t h = seed;
  for (int c = 0; c < ncols; c++) {
    uint32_t const e = identity ? identity_elem(cols[c], i) : elem_hash(cols[c], i, seed);
    h = hash_combine(h, e);
  }
  gp(pid)[i] = (num_partitions & (num_partitions - 1)) == 0 ? (h & (num_partitions - 1)) : h % num_partitions;
}

__global__ void __launch_bounds__(HP_BLOCK)
  k_map_partition_ids(void const* __restrict__ map, int32_t map_storage, int64_t n, uint32_t num_partitions,
                      uint32_t* __restrict__ pid, uint32_t* __restrict__ bad)
{
  int64_t const i = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (i >= n) return;
  int64_t v;
  switch (map_storage) {
    case ST_I8: v = reinterpret_cast<gmem<int8_t const>*>(gp(map))[i]; break;
    case ST_I16: v = reinterpret_cast<gmem<int16_t const>*>(gp(map))[i]; break;
    case ST_I32: v = reinterpret_cast<gmem<int32_t const>*>(gp(map))[i]; break;
    case ST_U8: v = reinterpret_cast<gmem<uint8_t const>*>(gp(map))[i]; break;
    case ST_U16: v = reinterpret_cast<gmem<uint16_t const>*>(gp(map))[i]; break;
    case ST_U32: v = reinterpret_cast<gmem<uint32_t const>*>(gp(map))[i]; break;
    case ST_U64: {
      uint64_t const u = reinterpret_cast<gmem<uint64_t const>*>(gp(map))[i];
      v = u > uint64_t(INT64_MAX) ? -1 : int64_t(u);
      break;
    }
    default: v = reinterpret_cast<gmem<int64_t const>*>(gp(map))[i]; break;
  }
  bool const ok = v >= 0 && v < int64_t(num_partitions);
  if (!ok) atomicOr(bad, 1u);
  gp(pid)[i] = ok ? uint32_t(v) : 0u;
}

// Rows [w * per, min((w + 1) * per, n)) belong to wave w.
__device__ __forceinline__ void wave_range(int64_t n, int64_t per, int64_t w, int64_t& b, int64_t& e)
{
  b = w * per;
  e = b + per < n ? b + per : n;
  if (b > n) b = n;
}

__global__ void __launch_bounds__(HP_BLOCK)
  k_hp_count(uint32_t const* __restrict__ pid, int64_t n, int64_t per, int64_t W, int shift, uint32_t dmask,
             uint32_t* __restrict__ hist)
{
  extern __shared__ uint32_t hp_lds[];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t const nd = dmask + 1;
  uint32_t* cnt     = hp_lds + wave * nd;
  for (uint32_t d = lane; d < nd; d += 64) cnt[d] = 0;
  __syncthreads();
  int64_t const w = int64_t(blockIdx.x) * HP_WAVES + wave;
  int64_t b = n, e = n;
  if (w < W) wave_range(n, per, w, b, e);
  for (int64_t i = b + lane; i < e; i += 64) atomicAdd(&cnt[(gp(pid)[i] >> shift) & dmask], 1u);
  __syncthreads();
  if (w < W)
    for (uint32_t d = lane; d < nd; d += 64) gp(hist)[int64_t(d) * W + w] = cnt[d];
}

__global__ void __launch_bounds__(HP_BLOCK)
  k_hp_scatter(uint32_t const* __restrict__ pid_in, uint32_t const* __restrict__ rows_in, int64_t n, int64_t per,
               int64_t W, int shift, int bits, uint32_t const* __restrict__ start, uint32_t* __restrict__ pid_out,
               uint32_t* __restrict__ rows_out, hp_move_col const* __restrict__ move, int nmove)
{
  extern __shared__ uint32_t hp_lds[];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t const dmask = (1u << bits) - 1u;
  uint32_t const nd    = dmask + 1;
  volatile uint32_t* next = hp_lds + wave * nd;
  int64_t const w = int64_t(blockIdx.x) * HP_WAVES + wave;
  if (w >= W) return;
  for (uint32_t d = lane; d < nd; d += 64) next[d] = gp(start)[int64_t(d) * W + w];
  __builtin_amdgcn_wave_barrier();
  int64_t b, e;
  wave_range(n, per, w, b, e);
  uint64_t const below = lanes_below();
  for (int64_t i0 = b; i0 < e; i0 += 64) {
    int64_t const i  = i0 + lane;
    bool const valid = i < e;
    uint32_t const p   = valid ? gp(pid_in)[i] : 0u;
    uint32_t const row = valid ? (rows_in ? gp(rows_in)[i] : uint32_t(i)) : 0u;
    uint32_t const d   = (p >> shift) & dmask;
    // lanes holding the same digit
    uint64_t peers = __ballot(valid);
    for (int k = 0; k < bits; k++) {
      uint64_t const set = __ballot(valid && ((d >> k) & 1u));
      peers &= ((d >> k) & 1u) ? set : ~set;
    }
    uint32_t base = 0;
    if (valid) base = next[d];
    __builtin_amdgcn_wave_barrier();
    if (valid) {
      uint32_t const pos = base + uint32_t(__popcll(peers & below));
      if (pid_out) gp(pid_out)[pos] = p;  // needed by a later pass or the offsets kernel only
      if (rows_out) gp(rows_out)[pos] = row;
      // the last pass moves fixed-width values straight to their place
      for (int c = 0; c < nmove; c++) {
        hp_move_col const mc = move[c];
        switch (mc.width) {
          case 1: gp(static_cast<uint8_t*>(mc.dst))[pos] = gp(static_cast<uint8_t const*>(mc.src))[row]; break;
          case 2: gp(static_cast<uint16_t*>(mc.dst))[pos] = gp(static_cast<uint16_t const*>(mc.src))[row]; break;
          case 4: gp(static_cast<uint32_t*>(mc.dst))[pos] = gp(static_cast<uint32_t const*>(mc.src))[row]; break;
          case 8: gp(static_cast<uint64_t*>(mc.dst))[pos] = gp(static_cast<uint64_t const*>(mc.src))[row]; break;
          default: {
            uint64_t const lo = gp(static_cast<uint64_t const*>(mc.src))[2 * int64_t(row)];
            uint64_t const hi = gp(static_cast<uint64_t const*>(mc.src))[2 * int64_t(row) + 1];
            gp(static_cast<uint64_t*>(mc.dst))[2 * int64_t(pos)]     = lo;
            gp(static_cast<uint64_t*>(mc.dst))[2 * int64_t(pos) + 1] = hi;
          }
        }
      }
      if ((peers & below) == 0) next[d] = base + uint32_t(__popcll(peers));
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// offsets[q] = first output row of partition q (pid_sorted is non-decreasing).
__global__ void __launch_bounds__(HP_BLOCK)
  k_hp_offsets(uint32_t const* __restrict__ pid_sorted, int64_t n, uint32_t num_partitions,
               int32_t* __restrict__ offsets)
{
  int64_t const i = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (i > n) return;
  int64_t const prev = i == 0 ? -1 : int64_t(gp(pid_sorted)[i - 1]);
  int64_t const cur  = i == n ? int64_t(num_partitions) - 1 : int64_t(gp(pid_sorted)[i]);
  for (int64_t q = prev + 1; q <= cur; q++) gp(offsets)[q] = int32_t(i);
}

// One counting pass: the first row of partition q is the scanned histogram
// entry of digit q for wave 0.
__global__ void __launch_bounds__(HP_BLOCK)
  k_hp_offsets_from_hist(uint32_t const* __restrict__ start, int64_t W, uint32_t num_partitions,
                         int32_t* __restrict__ offsets)
{
  int64_t const q = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (q < num_partitions) gp(offsets)[q] = int32_t(gp(start)[q * W]);
}

template <typename T>
__global__ void __launch_bounds__(HP_BLOCK)
  k_gather_fixed(T const* __restrict__ src, uint32_t const* __restrict__ map, int64_t n, T* __restrict__ dst)
{
  int64_t const i = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (i < n) gp(dst)[i] = gp(src)[gp(map)[i]];
}

// 16-byte elements (DECIMAL128) as two 8-byte halves.
__global__ void __launch_bounds__(HP_BLOCK)
  k_gather_16(uint64_t const* __restrict__ src, uint32_t const* __restrict__ map, int64_t n,
              uint64_t* __restrict__ dst)
{
  int64_t const i = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (i >= n) return;
  int64_t const r = gp(map)[i];
  uint64_t const lo = gp(src)[2 * r], hi = gp(src)[2 * r + 1];
  gp(dst)[2 * i]     = lo;
  gp(dst)[2 * i + 1] = hi;
}

// Null count of a workgroup added once: the per-wave counts are summed in
// LDS first (one atomic per row would serialise on a single address).
__device__ __forceinline__ void add_block_nulls(uint32_t wave_nulls, uint32_t* null_count)
{
  __shared__ uint32_t part[HP_WAVES];
  int const wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) part[wave] = wave_nulls;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < HP_WAVES; w++) t += part[w];
    if (t) atomicAdd(null_count, t);
  }
}

// Output validity: each wave builds two mask words per 64 rows (grid-stride
// over the rows), nulls counted per workgroup.
__global__ void __launch_bounds__(HP_BLOCK)
  k_gather_mask(uint32_t const* __restrict__ mask, int64_t mask_offset, uint32_t const* __restrict__ map,
                int64_t n, uint32_t* __restrict__ out, uint32_t* __restrict__ null_count)
{
  uint32_t nulls = 0;
  for (int64_t base = int64_t(blockIdx.x) * HP_BLOCK; base < n; base += int64_t(gridDim.x) * HP_BLOCK) {
    int64_t const i = base + threadIdx.x;
    bool v = false;
    if (i < n) {
      int64_t const r = int64_t(gp(map)[i]) + mask_offset;
      v               = (gp(mask)[r >> 5] >> (r & 31)) & 1u;
    }
    uint64_t const bal = __ballot(v);
    uint64_t const in  = __ballot(i < n);
    if ((threadIdx.x & 63) == 0 && in) {
      int64_t const word = (i >> 5);
      gp(out)[word]      = uint32_t(bal);
      gp(out)[word + 1]  = uint32_t(bal >> 32);
      nulls += uint32_t(__popcll(in & ~bal));
    }
  }
  add_block_nulls(nulls, null_count);
}

__global__ void __launch_bounds__(HP_BLOCK)
  k_gather_str_sizes(dcol src, uint32_t const* __restrict__ map, int64_t n, int32_t* __restrict__ sizes)
{
  int64_t const i = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (i >= n) return;
  int64_t b, e;
  str_span(src, int64_t(gp(map)[i]), b, e);
  gp(sizes)[i] = int32_t(e - b);
}

// ---- concatenation -----------------------------------------------------------
__device__ __forceinline__ int find_piece(concat_piece const* __restrict__ pieces, int np, int64_t r)
{
  int lo = 0, hi = np - 1;  // last piece with row_start <= r
  while (lo < hi) {
    int const mid = (lo + hi + 1) >> 1;
    if (pieces[mid].row_start <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(HP_BLOCK)
  k_concat_mask(concat_piece const* __restrict__ pieces, int np, int64_t n, uint32_t* __restrict__ out,
                uint32_t* __restrict__ null_count)
{
  uint32_t nulls = 0;
  for (int64_t base = int64_t(blockIdx.x) * HP_BLOCK; base < n; base += int64_t(gridDim.x) * HP_BLOCK) {
    int64_t const i = base + threadIdx.x;
    bool v = false;
    if (i < n) {
      concat_piece const& P = pieces[find_piece(pieces, np, i)];
      if (P.mask == nullptr) v = true;
      else {
        int64_t const r = i - P.row_start + P.offset;
        v               = (gp(P.mask)[r >> 5] >> (r & 31)) & 1u;
      }
    }
    uint64_t const bal = __ballot(v);
    uint64_t const in  = __ballot(i < n);
    if ((threadIdx.x & 63) == 0 && in) {
      int64_t const word = (i >> 5);
      gp(out)[word]      = uint32_t(bal);
      gp(out)[word + 1]  = uint32_t(bal >> 32);
      nulls += uint32_t(__popcll(in & ~bal));
    }
  }
  add_block_nulls(nulls, null_count);
}

// String offsets of the concatenation: each piece's offsets rebased to where
// its chars land.  offsets[n] = total chars.
__global__ void __launch_bounds__(HP_BLOCK)
  k_concat_str_offsets(concat_piece const* __restrict__ pieces, int np, int64_t n, int64_t total_chars,
                       int32_t* __restrict__ out)
{
  int64_t const i = int64_t(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    gp(out)[n] = int32_t(total_chars);
    return;
  }
  concat_piece const& P = pieces[find_piece(pieces, np, i)];
  int64_t const r       = i - P.row_start + P.offset;
  int64_t o;
  if (P.offsets64) o = reinterpret_cast<gmem<int64_t const>*>(gp(P.offsets))[r];
  else o = reinterpret_cast<gmem<int32_t const>*>(gp(P.offsets))[r];
  gp(out)[i] = int32_t(o - P.chars_begin + P.chars_start);
}

// First and last offset of each string piece (its chars range).
__global__ void k_concat_char_ranges(concat_piece const* __restrict__ pieces, int np, int64_t* __restrict__ ranges)
{
  int const k = int(blockIdx.x) * HP_BLOCK + threadIdx.x;
  if (k >= np) return;
  concat_piece const& P = pieces[k];
  int64_t b, e;
  if (P.rows == 0 || P.offsets == nullptr) {  // an empty strings column may have no offsets
    b = 0;
    e = 0;
  } else if (P.offsets64) {
    auto o = reinterpret_cast<gmem<int64_t const>*>(gp(P.offsets));
    b = o[P.offset];
    e = o[P.offset + P.rows];
  } else {
    auto o = reinterpret_cast<gmem<int32_t const>*>(gp(P.offsets));
    b = o[P.offset];
    e = o[P.offset + P.rows];
  }
  gp(ranges)[2 * k]     = b;
  gp(ranges)[2 * k + 1] = e;
}

inline unsigned grid_of(int64_t n) { return unsigned((n + HP_BLOCK - 1) / HP_BLOCK); }
// mask kernels: enough workgroups to fill the device, each looping over rows
inline unsigned mask_grid(int64_t n) { return unsigned(std::min<int64_t>(grid_of(n), 4096)); }

}  // namespace

void launch_hp_partition_ids(dcol const* d_cols, int ncols, int64_t n, uint32_t seed, bool identity,
                             uint32_t num_partitions, uint32_t* pid, hipStream_t s)
{
  if (n == 0) return;
  hipLaunchKernelGGL(k_hp_partition_ids, dim3(grid_of(n)), dim3(HP_BLOCK), 0, s, d_cols, ncols, n, seed,
                     identity ? 1 : 0, num_partitions, pid);
}

void launch_map_partition_ids(void const* map, int32_t map_storage, int64_t n, uint32_t num_partitions,
                              uint32_t* pid, uint32_t* bad, hipStream_t s)
{
  if (n == 0) return;
  hipLaunchKernelGGL(k_map_partition_ids, dim3(grid_of(n)), dim3(HP_BLOCK), 0, s, map, map_storage, n,
                     num_partitions, pid, bad);
}

void launch_hp_count(uint32_t const* pid, int64_t n, int64_t per, int64_t W, int shift, int bits, uint32_t* hist,
                     hipStream_t s)
{
  size_t const lds = size_t(HP_WAVES) * (size_t(1) << bits) * sizeof(uint32_t);
  unsigned const blocks = unsigned((W + HP_WAVES - 1) / HP_WAVES);
  hipLaunchKernelGGL(k_hp_count, dim3(blocks), dim3(HP_BLOCK), lds, s, pid, n, per, W, shift,
                     (1u << bits) - 1u, hist);
}

void launch_hp_scatter(uint32_t const* pid_in, uint32_t const* rows_in, int64_t n, int64_t per, int64_t W,
                       int shift, int bits, uint32_t const* start, uint32_t* pid_out, uint32_t* rows_out,
                       hp_move_col const* move, int nmove, hipStream_t s)
{
  size_t const lds = size_t(HP_WAVES) * (size_t(1) << bits) * sizeof(uint32_t);
  unsigned const blocks = unsigned((W + HP_WAVES - 1) / HP_WAVES);
  hipLaunchKernelGGL(k_hp_scatter, dim3(blocks), dim3(HP_BLOCK), lds, s, pid_in, rows_in, n, per, W, shift,
                     bits, start, pid_out, rows_out, move, nmove);
}

void launch_hp_offsets(uint32_t const* pid_sorted, int64_t n, uint32_t num_partitions, int32_t* offsets,
                       hipStream_t s)
{
  hipLaunchKernelGGL(k_hp_offsets, dim3(grid_of(n + 1)), dim3(HP_BLOCK), 0, s, pid_sorted, n, num_partitions,
                     offsets);
}

void launch_hp_offsets_from_hist(uint32_t const* start, int64_t W, uint32_t num_partitions, int32_t* offsets,
                                 hipStream_t s)
{
  hipLaunchKernelGGL(k_hp_offsets_from_hist, dim3(grid_of(num_partitions)), dim3(HP_BLOCK), 0, s, start, W,
                     num_partitions, offsets);
}

void launch_gather_fixed(void const* src, int width, uint32_t const* map, int64_t n, void* dst, hipStream_t s)
{
  if (n == 0) return;
  dim3 const g(grid_of(n)), b(HP_BLOCK);
  switch (width) {
    case 1: hipLaunchKernelGGL(k_gather_fixed<uint8_t>, g, b, 0, s, static_cast<uint8_t const*>(src), map, n, static_cast<uint8_t*>(dst)); break;
    case 2: hipLaunchKernelGGL(k_gather_fixed<uint16_t>, g, b, 0, s, static_cast<uint16_t const*>(src), map, n, static_cast<uint16_t*>(dst)); break;
    case 4: hipLaunchKernelGGL(k_gather_fixed<uint32_t>, g, b, 0, s, static_cast<uint32_t const*>(src), map, n, static_cast<uint32_t*>(dst)); break;
    case 8: hipLaunchKernelGGL(k_gather_fixed<uint64_t>, g, b, 0, s, static_cast<uint64_t const*>(src), map, n, static_cast<uint64_t*>(dst)); break;
    default: hipLaunchKernelGGL(k_gather_16, g, b, 0, s, static_cast<uint64_t const*>(src), map, n, static_cast<uint64_t*>(dst)); break;
  }
}

void launch_gather_mask(uint32_t const* mask, int64_t mask_offset, uint32_t const* map, int64_t n, uint32_t* out,
                        uint32_t* null_count, hipStream_t s)
{
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_mask, dim3(mask_grid(n)), dim3(HP_BLOCK), 0, s, mask, mask_offset, map, n, out,
                     null_count);
}

void launch_gather_str_sizes(dcol const& src, uint32_t const* map, int64_t n, int32_t* sizes, hipStream_t s)
{
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_str_sizes, dim3(grid_of(n)), dim3(HP_BLOCK), 0, s, src, map, n, sizes);
}

void launch_concat_mask(concat_piece const* pieces, int np, int64_t n, uint32_t* out, uint32_t* null_count,
                        hipStream_t s)
{
  if (n == 0) return;
  hipLaunchKernelGGL(k_concat_mask, dim3(mask_grid(n)), dim3(HP_BLOCK), 0, s, pieces, np, n, out, null_count);
}

void launch_concat_char_ranges(concat_piece const* pieces, int np, int64_t* ranges, hipStream_t s)
{
  hipLaunchKernelGGL(k_concat_char_ranges, dim3(grid_of(np)), dim3(HP_BLOCK), 0, s, pieces, np, ranges);
}

void launch_concat_str_offsets(concat_piece const* pieces, int np, int64_t n, int64_t total_chars, int32_t* out,
                               hipStream_t s)
{
  hipLaunchKernelGGL(k_concat_str_offsets, dim3(grid_of(n + 1)), dim3(HP_BLOCK), 0, s, pieces, np, n,
                     total_chars, out);
}

}  // namespace mig
