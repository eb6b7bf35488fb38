// Parameter blocks and launcher entry points of the group-by kernels.
//
// The host side (groupby.cpp) fills a `gb_params` block once per call and
// copies it to device memory; every kernel of the call reads it through a
// uniform pointer, so the descriptors live in the scalar cache.
#pragma once

#include "common.hpp"

#include <vector>

namespace mig {

constexpr int MAX_KEYS   = 16;
constexpr int MAX_VALS   = 32;
constexpr int MAX_FIELDS = 128;
constexpr int MAX_OUTS   = 128;

// Empty marker of the inline key table.  Normalised float keys can never take
// this value (it is a signalling-NaN pattern and NaN keys are canonicalised);
// an integer key equal to it is routed to the dedicated slot `cap`.
constexpr uint64_t EMPTY_KEY = 0xFFF1A5A5A5A5A5A5ull;
// Empty marker of the general (row index) table: (hash << 32) | row.
constexpr uint64_t EMPTY_ROW = 0xFFFFFFFFFFFFFFFFull;

enum key_mode : int32_t {
  KM_INLINE  = 0,  // one fixed-width key column, normalised bits stored in the table
  KM_GENERAL = 1   // any keys: the table stores (row hash, representative row)
};

// Accumulator kinds.  Small integer sources are widened to 32 bits for
// min/max; floats use an order-preserving unsigned encoding for min/max.
enum field_kind : int32_t {
  FK_SUM_I64 = 0,
  FK_SUM_F32,
  FK_SUM_F64,
  FK_SUMSQ_I64,
  FK_SUMSQ_F32,
  FK_SUMSQ_F64,
  FK_MIN_S32,
  FK_MAX_S32,
  FK_MIN_S64,
  FK_MAX_S64,
  FK_MIN_U32,
  FK_MAX_U32,
  FK_MIN_U64,
  FK_MAX_U64,
  FK_MIN_F32,
  FK_MAX_F32,
  FK_MIN_F64,
  FK_MAX_F64,
  FK_COUNT,  // valid values (u32)
  FK_M2,     // sum of squared deviations from the group mean (f64, second pass)
  FK_ARGMIN, // row of the smallest valid value (u32, 0xFFFFFFFF = none)
  FK_ARGMAX, // row of the largest valid value
  FK_PROD_I64,  // products (compare-and-swap loops)
  FK_PROD_F32,
  FK_PROD_F64,
  FK_NUNIQUE,        // distinct valid values (u32; filled by the pre-sorted path's set insert)
  FK_NUNIQUE_NULLS,  // distinct values, nulls counted as one value
  FK_NTH             // row of the group's n-th value (u32, 0xFFFFFFFF = none); aux_sum = n, aux_count = include
};

// Bytes of one accumulator of kind `fk`.
__host__ __device__ inline int32_t field_kind_width(int32_t fk)
{
  switch (fk) {
    case FK_SUM_F32:
    case FK_SUMSQ_F32:
    case FK_MIN_S32:
    case FK_MAX_S32:
    case FK_MIN_U32:
    case FK_MAX_U32:
    case FK_MIN_F32:
    case FK_MAX_F32:
    case FK_COUNT:
    case FK_ARGMIN:
    case FK_ARGMAX:
    case FK_NUNIQUE:
    case FK_NUNIQUE_NULLS:
    case FK_NTH:
    case FK_PROD_F32: return 4;
    default: return 8;
  }
}

// Bytes of an accumulator in the partitioned path's LDS table.  With
// `wide_f32`, float32 sums and sums of squares accumulate in float64 there
// (LDS float64 atomic adds run several times faster than float32 ones, and
// the sum is rounded to float32 once at the end); everything else keeps its
// own width.
__host__ __device__ inline bool rpa_wide_sum(int32_t fk, bool wide_f32)
{
  return wide_f32 && (fk == FK_SUM_F32 || fk == FK_SUMSQ_F32);
}
__host__ __device__ inline int32_t rpa_lds_width(int32_t fk, bool wide_f32)
{
  return rpa_wide_sum(fk, wide_f32) ? 8 : field_kind_width(fk);
}

struct field_desc {
  void* ptr;        // global accumulator of slot 0 (cap + 2 slots)
  int32_t kind;     // field_kind
  int32_t vcol;     // value column it reads
  int32_t width;    // 4 or 8 bytes
  int32_t stride;   // bytes between the accumulators of consecutive slots
  int32_t lds_off;  // byte offset of this field's array in the LDS table
  int32_t aux_sum;    // FK_M2: index of the column's sum field
  int32_t aux_count;  // FK_M2: index of the column's count field
};

// Output program: how one result column is produced from the accumulators.
enum out_kind : int32_t {
  OK_COPY_I64 = 0,  // int64 sum / sum of squares
  OK_COPY_F32,
  OK_COPY_F64,
  OK_COUNT,      // int32 from a u32 count field
  OK_COUNT_ALL,  // int32 from the group row count
  OK_MINMAX,     // decode a min/max field into the source type
  OK_MEAN,       // double(sum) / count
  OK_VAR,        // m2 / (count - ddof)
  OK_STD,        // sqrt(var)
  OK_ARG,        // int32 row index from an ARGMIN / ARGMAX field, null when none
  OK_NTH,        // the value at an FK_NTH field's row, null when none or null
  OK_SUM_I32,    // DECIMAL32 SUM: the int64 sum wrapped to the 32-bit representation
  OK_MEAN_INT    // DECIMAL32/64 MEAN: representation sum (wrapped to out_width) / count, truncated
};

struct out_desc {
  void* data;         // G elements of the output type
  uint32_t* mask;     // validity words or nullptr
  int32_t kind;       // out_kind
  int32_t field;      // primary field (sum / min / max / count / m2)
  int32_t count_field;  // count field for validity and mean/var, -1 = always valid
  int32_t src_storage;  // source storage (min/max decode, mean of ints)
  int32_t ddof;
  int32_t out_width;
};

struct gb_params {
  int64_t num_rows;
  uint64_t cap;       // table slots (power of two); slots cap, cap+1 are special
  uint64_t cap_mask;
  uint32_t probe_limit;
  int32_t num_keys;
  int32_t num_vals;
  int32_t num_fields;
  int32_t num_outs;
  int32_t key_mode;
  int32_t skip_null_keys;  // null_policy EXCLUDE and some key column has nulls
  int32_t key_storage;     // inline mode: storage of the key column
  int32_t need_count_all;
  int32_t need_second_pass;  // some field is FK_M2
  uint64_t* slot_keys;   // inline table
  uint64_t* slot_rows;   // general table
  uint32_t* slot_count;  // group row count (COUNT_ALL) of slot 0, or null
  int32_t count_stride;  // bytes between the row counts of consecutive slots
  int32_t pad_stride;
  uint32_t* special;     // [2] occupancy of slots cap (sentinel key) and cap+1 (null key)
  int32_t* overflow;     // set when a probe sequence exceeds probe_limit
  dcol keys[MAX_KEYS];
  dcol vals[MAX_VALS];
  int32_t val_field_begin[MAX_VALS + 1];  // fields are grouped by value column
  field_desc fields[MAX_FIELDS];
  out_desc outs[MAX_OUTS];
  // key output (finalize)
  void* out_keys[MAX_KEYS];       // fixed-width key outputs (or string offsets)
  uint32_t* out_key_mask[MAX_KEYS];
  // low-cardinality pre-aggregation (inline keys): each workgroup keeps an
  // LDS table of lds_slots keys (at most lds_fill) with partials at
  // fields[f].lds_off (lds_slots + 2 elements of the field's width; slots
  // lds_slots and lds_slots + 1 hold the EMPTY_KEY and null-key groups) and
  // the row counts at lds_count_off (-1: none); the variance pass reuses the
  // bytes for a table of group slots and the M2 partials at lds_m2_off[f]
  int32_t lds_slots;
  int32_t lds_fill;
  int32_t lds_count_off;
  int32_t lds_bytes;
  int32_t lds_m2_off[MAX_FIELDS];
};

// ---------------------------------------------------------------------------
// Radix-partitioned aggregation (high cardinality, one fixed-width key).
// Rows are partitioned by the top bits of their row hash into B = 2^log_bins
// bins small enough that all groups of a bin fit in one workgroup's LDS
// (one or two scatter passes), then one workgroup aggregates each bin in LDS.
// Partition buffers are structure-of-arrays: normalised key bits (u64), the
// raw bytes of each value column, and a flags byte per row (bit v = value
// column v valid, bit 7 = key null under INCLUDE).
constexpr int RPA_MAX_FLAGGED = 7;
constexpr uint8_t RPA_KEY_NULL = 0x80;
constexpr int RPA_TILE = 2048;          // rows per scatter tile (at most)
constexpr int RPA_SMALL_BLOCK = 256;    // small helper kernels
constexpr int RPA_SCATTER_BLOCK = 512;  // scatter kernels
constexpr int RPA_HIST_BLOCK = 1024;    // histogram kernel
constexpr int RPA_AGG_BLOCK = 512;
constexpr int RPA_PLAN_COLS   = 4;   // value columns with accumulators (partitioned path)
constexpr int RPA_PLAN_FIELDS = 16;  // accumulator fields (partitioned path)
constexpr int RPA_PLAN_OUTS   = 64;  // result columns (partitioned path)
constexpr int RPA_CURSOR_STRIDE = 16;  // u64 elements between scatter cursors (128 bytes)

struct rpa_params {
  int32_t log_bins;   // final bins B = 1 << log_bins
  int32_t log_f1;     // first-level fanout F1 = 1 << log_f1 (== log_bins when one level)
  int32_t two_level;  // 1 when the second scatter pass runs
  int32_t lds_slots;  // LDS hash slots S per aggregating workgroup (power of two);
                      // accumulator rows S and S+1 hold the null-key group and
                      // the group of keys equal to EMPTY_KEY
  int32_t max_fill;   // keys a bin may insert before it counts as overflowing
  int32_t lds_bytes;  // dynamic LDS of the aggregate kernel
  int32_t count_all_off;  // LDS byte offsets of the aggregate kernel
  int32_t slot_of_j_off;
  int32_t misc_off;
  int32_t pad_agg;
  // scatter staging: a tile of rows sorted by bin in LDS before it is written
  int32_t tile_rows;
  int32_t stage_bytes;
  int32_t stage_flags_off;
  int32_t stage_bin_off;
  int32_t stage_val_off[MAX_VALS];
  int32_t hist_blocks;               // blocks of the histogram kernel
  int32_t pad0;
  uint32_t* hist_partial;            // [hist_blocks][B] per-block histograms
  uint32_t* hist;                    // [B] rows per final bin
  int64_t* bin_start;                // [B + 1]
  // next free row of each first-level bin (buffer A) / final bin (buffer
  // B), one per RPA_CURSOR_STRIDE elements so that every cursor has a cache
  // line of its own (each tile of a scatter reserves ranges on many of them)
  unsigned long long* cursor1;       // [F1 * RPA_CURSOR_STRIDE]
  unsigned long long* cursor2;       // [B * RPA_CURSOR_STRIDE]
  int64_t* tile_start1;              // [F1 + 1] scatter-2 tiles before each first-level bin
  int64_t* first_start;              // [F1] first record of each first-level bin (buffer A)
  // Rows each first-level / final bin may hold.  Exact layouts (from the
  // histogram) never reach them; speculative layouts (sized from the sampled
  // estimate, no histogram pass) flag a bin that would pass its capacity in
  // cap_overflow and write none of the rows that do not fit.
  int64_t cap1;
  int64_t cap2;
  int32_t* cap_overflow;
  uint64_t* keysA;
  uint64_t* keysB;
  uint8_t* flagsA;
  uint8_t* flagsB;
  uint8_t* valsA[MAX_VALS];
  uint8_t* valsB[MAX_VALS];
  int32_t val_flag_bit[MAX_VALS];    // flags bit of value column v, -1 if it has no nulls
  unsigned long long* group_counter; // groups emitted so far
  int32_t* overflow;                 // a bin had more groups than max_ids
  int32_t pad1;
  uint64_t out_capacity;             // rows allocated in every output column
  uint32_t* out_rep;                 // general keys: representative input row of each output row
  // General keys in hashed form (general_hashed = 1): the record word is
  // (row hash << 32) | a second, independent row hash, the row index travels
  // in rowsA / rowsB, and bins compare words instead of key rows.  A bin slot's
  // representative row is rep_rows[bin * rep_stride + slot]; rec_slot[row]
  // (input row; 0xFFFFFFFF for dropped rows and representatives) receives the
  // slot of the row's group, and k_rpa_verify compares every such row's key
  // row with its representative's exactly afterwards, in row order (any
  // difference reruns the call with general_hashed = 0, which compares key
  // rows while probing).
  int32_t general_hashed;
  int32_t stage_row_off;
  int64_t rep_stride;
  uint32_t* rowsA;
  uint32_t* rowsB;
  uint32_t* rep_rows;
  uint32_t* rec_slot;                // [input rows]
  int64_t rec_count;                 // input rows (the verification's range)
  unsigned long long* verified;      // rows compared with a representative
  int32_t* verify_failed;
  // Pass-A plan of the aggregate kernel, read into registers once per
  // workgroup: the accumulators of value column plan_col[i] are the fields
  // [plan_fb[i], plan_fb[i + 1]), field f described by plan_fd[f] =
  // kind | lds_off << 8.
  int32_t plan_ncols;
  int32_t plan_col[RPA_PLAN_COLS];
  int32_t plan_fb[RPA_PLAN_COLS + 1];
  uint32_t plan_fd[RPA_PLAN_FIELDS];
  // Variance finalisation plan: entry j describes M2 field j by its LDS
  // offset, its column's sum field (kind | lds_off << 8), the LDS offset of
  // its count (a count field or the group row counts) and the source storage.
  int32_t wide_f32;  // float32 sums accumulate in float64 (rpa_lds_width)
  int32_t m2_n;
  int32_t m2_off[RPA_PLAN_COLS];
  uint32_t m2_sum[RPA_PLAN_COLS];
  int32_t m2_cnt_off[RPA_PLAN_COLS];
  int32_t m2_storage[RPA_PLAN_COLS];
  // Output plan, read into registers once per workgroup: result column o is
  // out_op[o] = out kind | primary field kind << 8 | output width << 16, from
  // the field at LDS byte offset out_foff[o] (-1: none) and the count at
  // out_coff[o] (-1: always valid), VARIANCE/STD with out_ddof[o].
  int32_t out_n;
  uint32_t out_op[RPA_PLAN_OUTS];
  int32_t out_foff[RPA_PLAN_OUTS];
  int32_t out_coff[RPA_PLAN_OUTS];
  int32_t out_ddof[RPA_PLAN_OUTS];
};

// ---- launchers (groupby_kernels.hip) --------------------------------------
struct launch_cfg {
  int num_cus;
};

void launch_hash_rows(dcol const* d_cols, int ncols, int64_t n, uint32_t seed, uint32_t* out,
                      hipStream_t s);
// Distinct row hashes among `nsample` sampled rows: counters[0] = rows
// sampled (dropped rows excluded), counters[1] = distinct hashes,
// counters[2] = the largest number of sampled rows sharing one hash.
// `table` has table_mask + 1 zeroed slots.
void launch_sample_hash(gb_params const* dP, int64_t n, int64_t nsample, uint64_t* table,
                        uint32_t table_mask, uint32_t* counters, hipStream_t s);
void launch_init_state(gb_params const* dP, gb_params const& hP, hipStream_t s);
// Allows `bytes` of dynamic LDS for kernel `fn` (above the 64 KB default;
// gfx950 has 160 KB per CU).
void allow_dynamic_lds(void const* fn, int bytes);
void launch_aggregate(gb_params const* dP, gb_params const& hP, int num_cus, hipStream_t s);
// Low-cardinality form of launch_aggregate / launch_second_pass (inline keys,
// hP.lds_bytes > 0): `blocks` workgroups pre-aggregate their rows in LDS.
void launch_aggregate_lds(gb_params const* dP, gb_params const& hP, int64_t blocks, hipStream_t s);
void launch_second_pass_lds(gb_params const* dP, gb_params const& hP, int64_t blocks, hipStream_t s);
void launch_second_pass(gb_params const* dP, gb_params const& hP, int num_cus, hipStream_t s);
// Occupancy scan: block_counts has ceil((cap+2)/SCAN_TILE) entries.
constexpr int SCAN_TILE = 4096;
void launch_count_occupied(gb_params const* dP, gb_params const& hP, uint32_t* block_counts,
                           hipStream_t s);
void launch_scan_blocks(uint32_t* block_counts, int64_t nblocks, uint32_t* total, hipStream_t s);
// Writes map[g] = slot of dense group g, in slot order.
void launch_compact(gb_params const* dP, gb_params const& hP, uint32_t const* block_offsets,
                    uint32_t* map, hipStream_t s);
// Writes keys (inline table) or representative rows (general table) and all
// aggregation results of the G dense groups.
void launch_outputs(gb_params const* dP, gb_params const& hP, uint32_t const* map, int64_t G,
                    uint32_t* rep_rows, hipStream_t s);
// Gathers fixed-width key columns / string lengths by representative row;
// for string key columns also the source offset of each group's string.
void launch_gather_keys(gb_params const* dP, gb_params const& hP, uint32_t const* rep_rows,
                        int64_t G, int32_t* str_sizes_scratch, int64_t* str_begin_scratch, hipStream_t s);
// String chars from the source offsets launch_gather_keys recorded.
void launch_gather_chars_from(uint8_t const* src_chars, int64_t const* src_begin, int64_t G,
                              int32_t const* out_offsets, uint8_t* out_chars, hipStream_t s);
void launch_exclusive_scan_i32(int32_t* data, int64_t n, int32_t* total, hipStream_t s);
void launch_gather_string_chars(dcol const& src, uint32_t const* rep_rows, int64_t G,
                                int32_t const* out_offsets, uint8_t* out_chars, hipStream_t s);
void launch_count_nulls(dcol const& c, int64_t n, uint32_t* d_count, hipStream_t s);
// valid rows among the first n of each word-aligned mask, added into d_counts[i]
void launch_count_valid_list(uint32_t const* const* masks, int count, int64_t n, uint32_t* d_counts,
                             hipStream_t s);

// ---- pre-sorted keys: groups are runs of equal adjacent key rows ------------
int64_t sorted_tiles(int64_t n);
// run heads per tile of rows
void launch_sorted_count(gb_params const* dP, int64_t n, uint32_t* block_counts, hipStream_t s);
// group of every row (0xFFFFFFFF when dropped), slot_rows[g] = first row of group g
// and, when group_last is not null, group_last[g] = its last row
void launch_sorted_assign(gb_params const* dP, int64_t n, uint32_t const* block_offsets, uint32_t* gid,
                          uint32_t* group_last, hipStream_t s);
// Fields (and the group row count) that launch_sorted_stats reduces: value
// columns of int32 / int64 / float32 / float64 whose accumulators are all among SUM,
// SUM_OF_SQUARES, COUNT, MIN, MAX and M2, each by a kernel compiled for the
// column type.
struct sorted_stats_cover {
  std::vector<char> field;  // per field index: 1 = reduced by launch_sorted_stats
  bool rows      = false;   // the group row count too
  int m2_columns = 0;       // covered columns with M2 (each needs its per-group scratch)
};
sorted_stats_cover sorted_stats_plan(gb_params const& hP);
// The covered columns; M2 from the sums where exact (or within 1e-11 of the
// result for float64), otherwise flagged in m2_flag (G bytes) for
// launch_sorted_stats_m2, which runs after the sums and counts are final
// (m2_scratch: sorted_stats_scratch_bytes of device memory).
size_t sorted_stats_scratch_bytes(int64_t G, int m2_columns);
void launch_sorted_stats(gb_params const* dP, gb_params const& hP, int64_t n, int64_t G, uint32_t const* gid,
                         int num_cus, uint8_t* m2_scratch, hipStream_t s);
void launch_sorted_stats_m2(gb_params const* dP, gb_params const& hP, int64_t n, int64_t G, uint32_t const* gid,
                            int num_cus, uint8_t* m2_scratch, hipStream_t s);
// wave-range reductions with plain stores for groups inside one range (the
// fields `cover` names are left to launch_sorted_stats)
void launch_sorted_aggregate(gb_params const* dP, gb_params const& hP, int64_t n, uint32_t const* gid, int num_cus,
                             sorted_stats_cover const* cover, hipStream_t s);
void launch_sorted_second_pass(gb_params const* dP, gb_params const& hP, int64_t n, uint32_t const* gid,
                               int num_cus, sorted_stats_cover const* cover, hipStream_t s);
// NUNIQUE of field `f` on runs (`gid` = group of every row, NO_GROUP when
// dropped; G groups): a set of (group, value) in `table` (2n u32 slots,
// cleared here) counts each group's distinct values.
void launch_sorted_nunique(gb_params const* dP, int32_t f, int64_t n, int64_t G, uint32_t const* gid,
                           uint32_t* table, hipStream_t s);
// NTH_ELEMENT of field `f` on runs: the row of each group's n-th counted value,
// from the groups' first and last rows (launch_sorted_assign's group_last).
void launch_sorted_nth(gb_params const* dP, gb_params const& hP, int32_t f, int64_t G, uint32_t const* group_last,
                       uint32_t* large_scratch /* G + 1 entries */, hipStream_t s);
void launch_iota_u32(uint32_t* out, int64_t n, hipStream_t s);
// ARGMIN / ARGMAX fields of string columns on runs (the reduce passes skip them).
void launch_sorted_string_arg(gb_params const* dP, gb_params const& hP, int64_t n, uint32_t const* gid,
                              hipStream_t s);
// rows[g] = perm[rows[g]] where rows[g] >= 0.
void launch_map_rows(int32_t* rows, int64_t G, uint32_t const* perm, hipStream_t s);

// ---- sort-based aggregations (sort_kernels.hip, label kernels in
// groupby_kernels.hip) ---------------------------------------------------------
// Every row's slot in the device-wide table (inserting its key), UINT32_MAX
// for rows the null policy drops.
void launch_label_rows(gb_params const* dP, gb_params const& hP, int num_cus, uint32_t* slot_of_row, hipStream_t s);
// gid[r] = rank of row r's slot among the G occupied slots (map[j] = slot of
// group j), G for dropped rows; rank_of_slot has cap + 2 entries.
void launch_group_numbers(uint32_t const* map, int64_t G, uint32_t const* slot_of_row, int64_t n,
                          uint32_t* rank_of_slot, uint32_t* gid, hipStream_t s);
// Sort keys of a value column for quantiles: the order-preserving bits of
// each element split into lo / hi 32-bit words (hi only for 8-byte values)
// and the null flag (1 = null, so nulls sort after the values of a group).
// NaN orders after every number; -0.0 is taken as 0.0.
void launch_quantile_keys(dcol const& v, int64_t n, uint32_t* lo, uint32_t* hi, uint32_t* null_flag, hipStream_t s);
// Quantiles of every group: the group's rows are [off[g], off[g + 1]) of the
// order keys (ord_lo/ord_hi) and null flags sorted by (group, null, value);
// out[g * nq + j] and its validity bit (the mask is zeroed by the caller).
void launch_group_quantiles(uint32_t const* ord_lo, uint32_t const* ord_hi, uint32_t const* null_flag,
                            int32_t const* off, int64_t G, int32_t storage, double const* q, int32_t nq,
                            int32_t interp, double* out, uint32_t* out_mask, hipStream_t s);
// Group numbers above G (rows the null policy dropped) become G.
void launch_clamp_groups(uint32_t* gid, int64_t n, uint32_t G, hipStream_t s);
// rep[g] = map[off[g]]: the first input row of each group after a row sort.
void launch_first_rows(uint32_t const* map, int32_t const* off, int64_t G, uint32_t* rep, hipStream_t s);
// string MIN / MAX: lengths and validity of the values at the ARGMIN / ARGMAX rows (-1 = none)
void launch_str_arg_sizes(dcol const& src, int32_t const* rows, int64_t G, int32_t* sizes, uint32_t* mask,
                          hipStream_t s);

// ---- radix-partitioned aggregation (partition_kernels.hip) ------------------
void launch_rpa_hist(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int64_t n,
                     int num_cus, hipStream_t s);
void launch_rpa_scatter1(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int64_t n,
                         int num_cus, hipStream_t s);
// bin layout (bin_start, cursors, second-pass tiles) from the histogram
void launch_rpa_layout(rpa_params const* dR, hipStream_t s);
// Speculative layout: every first-level bin b1 starts at b1 * cap1, every
// final bin b at b * cap2 (no histogram).
void launch_rpa_spec_layout(rpa_params const* dR, rpa_params const& hR, hipStream_t s);
// Second-pass tiles of each first-level bin from the rows scatter 1 wrote.
void launch_rpa_tiles(rpa_params const* dR, hipStream_t s);
// second pass over the tiles [tile_begin, tile_end) of the first-level bins
// (tile_end < 0: to the last tile, counted on the device)
void launch_rpa_scatter2(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int num_cus,
                         int64_t tile_begin, int64_t tile_end, hipStream_t s);
// final bins [bin_begin, bin_end); key_form: 0 one fixed-width key inline,
// 1 general keys compared row by row, 2 general keys in hashed form
void launch_rpa_aggregate(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int key_form,
                          int64_t bin_begin, int64_t bin_end, hipStream_t s);
// exact key comparison of the (row, representative) pairs of the hashed form
void launch_rpa_verify(gb_params const* dP, rpa_params const* dR, int64_t records, hipStream_t s);

}  // namespace mig
