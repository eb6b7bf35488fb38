// Hash partitioning and concatenation of tables: launchers (partitioning_kernels.hip)
// and the host drivers (partitioning.cpp) behind the C ABI.
#pragma once

#include <memory>
#include <vector>

#include "groupby.hpp"

// Columns owned by a handle; buffers come from `alloc` (the call's resource).
struct mig_table {
  int32_t num_rows = 0;
  std::vector<mig::owned_column> columns;
  mig::device_allocator alloc;
  ~mig_table();
};

namespace mig {

// Digit bits of one counting-sort pass (2^11 LDS counters per wave).
constexpr int HP_DIGIT_BITS = 11;

// One input piece of a concatenation (device copy read by the kernels).
struct concat_piece {
  const uint32_t* mask;  // validity words or nullptr (all valid)
  const void* offsets;   // strings: offsets child
  int64_t row_start;     // first output row of the piece
  int64_t offset;        // first row of the view in its buffers
  int64_t rows;
  int64_t chars_begin;   // strings: offsets[offset]
  int64_t chars_start;   // strings: first output char of the piece
  int32_t offsets64;
  int32_t pad;
};

void launch_hp_partition_ids(dcol const* d_cols, int ncols, int64_t n, uint32_t seed, bool identity,
                             uint32_t num_partitions, uint32_t* pid, hipStream_t s);
void launch_map_partition_ids(void const* map, int32_t map_storage, int64_t n, uint32_t num_partitions,
                              uint32_t* pid, uint32_t* bad, hipStream_t s);
void launch_hp_count(uint32_t const* pid, int64_t n, int64_t per, int64_t W, int shift, int bits, uint32_t* hist,
                     hipStream_t s);
// Fixed-width column moved by the last counting pass: dst[pos] = src[row].
struct hp_move_col {
  const void* src;  // element 0 of the view
  void* dst;
  int32_t width;    // 1, 2, 4, 8 or 16 bytes
  int32_t pad;
};
constexpr int HP_MAX_MOVE_COLS = 16;

// pid_out / rows_out may be null when nothing reads them later.
void launch_hp_scatter(uint32_t const* pid_in, uint32_t const* rows_in, int64_t n, int64_t per, int64_t W,
                       int shift, int bits, uint32_t const* start, uint32_t* pid_out, uint32_t* rows_out,
                       hp_move_col const* move, int nmove, hipStream_t s);
void launch_hp_offsets(uint32_t const* pid_sorted, int64_t n, uint32_t num_partitions, int32_t* offsets,
                       hipStream_t s);
void launch_hp_offsets_from_hist(uint32_t const* start, int64_t W, uint32_t num_partitions, int32_t* offsets,
                                 hipStream_t s);
void launch_gather_fixed(void const* src, int width, uint32_t const* map, int64_t n, void* dst, hipStream_t s);
void launch_gather_mask(uint32_t const* mask, int64_t mask_offset, uint32_t const* map, int64_t n, uint32_t* out,
                        uint32_t* null_count, hipStream_t s);
void launch_gather_str_sizes(dcol const& src, uint32_t const* map, int64_t n, int32_t* sizes, hipStream_t s);
void launch_concat_mask(concat_piece const* pieces, int np, int64_t n, uint32_t* out, uint32_t* null_count,
                        hipStream_t s);
void launch_concat_char_ranges(concat_piece const* pieces, int np, int64_t* ranges, hipStream_t s);
void launch_concat_str_offsets(concat_piece const* pieces, int np, int64_t n, int64_t total_chars, int32_t* out,
                               hipStream_t s);

// One key of a stable LSD sort of rows: a 32-bit value per input row (in
// input row order) of which the low `bits` bits are significant.
struct lsd_key {
  uint32_t const* data;
  int bits;
};
// Stable sort of the rows 0..n-1 by `keys`, least significant key first:
// map[i] = the input row at sorted position i.  `last_sorted`, when given,
// receives the most significant key's values in sorted order.  Counting-sort
// passes of at most HP_DIGIT_BITS bits per key; enqueued on `s` (scratch is
// released after synchronising `s`).
void lsd_sort_rows(lsd_key const* keys, int nkeys, int64_t n, uint32_t* map, uint32_t* last_sorted, hipStream_t s);

// cudf::hash_partition / cudf::partition: the regrouped table and the first
// row of every partition (`offsets`, num_partitions entries).
std::unique_ptr<mig_table> hash_partition(mig_column_view const* input, int32_t num_columns,
                                          int32_t const* columns_to_hash, int32_t num_columns_to_hash,
                                          int32_t num_partitions, int32_t hash_function, uint32_t seed,
                                          hipStream_t stream, mig_memory_resource const* mr,
                                          std::vector<int32_t>& offsets);
std::unique_ptr<mig_table> partition_by_map(mig_column_view const* input, int32_t num_columns,
                                            mig_column_view const& partition_map, int32_t num_partitions,
                                            hipStream_t stream, mig_memory_resource const* mr,
                                            std::vector<int32_t>& offsets);
// cudf::concatenate of tables (columns[t * num_columns + c]).
std::unique_ptr<mig_table> concatenate(mig_column_view const* columns, int32_t num_tables, int32_t num_columns,
                                       hipStream_t stream, mig_memory_resource const* mr);

}  // namespace mig
