// Host drivers of hash partitioning (cudf::hash_partition, cudf::partition)
// and table concatenation (cudf::concatenate).
//
// hash_partition follows cpp/src/partitioning/partitioning.cu:759-830: the
// partition of a row is its row hash (MurmurHash3_x86_32 or IdentityHash of
// each element, combined from `seed` with hash_combine) modulo the number of
// partitions; rows are regrouped so that each partition is contiguous, and
// the first row of every partition is returned.  The regrouping is a stable
// counting sort (partitioning_kernels.hip), so rows keep their input order
// inside a partition and two calls on the same input give the same table.
#include "partitioning.hpp"

#include <algorithm>
#include <cstring>

#include "host_plan.hpp"
#include "kernels.hpp"

mig_table::~mig_table()
{
  for (auto& c : columns) {
    alloc.deallocate(c.col.data, c.data_bytes);
    alloc.deallocate(c.col.null_mask, c.mask_bytes);
    alloc.deallocate(c.col.offsets, c.offsets_bytes);
    c = mig::owned_column{};
  }
}

namespace mig {

namespace {

// Scratch blocks of one call, returned to the per-device cache once the
// call's stream has drained.
struct call_scratch {
  hipStream_t stream;
  int dev = 0;
  std::vector<std::pair<void*, size_t>> blocks;
  explicit call_scratch(hipStream_t s) : stream(s) { MIG_HIP_TRY(hipGetDevice(&dev)); }
  template <typename T>
  T* get(size_t count)
  {
    size_t const bytes = std::max<size_t>(256, (count * sizeof(T) + 255) & ~size_t(255));
    void* p            = scratch_take(dev, bytes);
    blocks.emplace_back(p, bytes);
    return static_cast<T*>(p);
  }
  ~call_scratch()
  {
    (void)hipStreamSynchronize(stream);
    for (auto& b : blocks) scratch_give(dev, b.first, b.second);
  }
};

// Element width of a fixed-width column (DECIMAL128 included), 0 otherwise.
int32_t element_width(int32_t type_id)
{
  if (type_id == MIG_TYPE_DECIMAL128) return 16;
  int32_t const st = storage_of(type_id);
  return st == ST_STRING ? 0 : storage_width(st);
}

void check_column_type(mig_column_view const& v)
{
  if (v.type_id != MIG_TYPE_STRING && element_width(v.type_id) == 0)
    fail(MIG_ERROR_DATA_TYPE, "unsupported column type " + std::to_string(v.type_id));
}

size_t mask_bytes_for(int64_t n) { return size_t((n + 511) / 512) * 64; }

void init_table(mig_table& t, hipStream_t stream, mig_memory_resource const* mr)
{
  t.alloc.stream = stream;
  if (mr) {
    t.alloc.mr     = *mr;
    t.alloc.has_mr = true;
  }
}

// Empty columns of the same types (cudf::empty_like of a table).
std::unique_ptr<mig_table> empty_like(mig_column_view const* input, int32_t num_columns, hipStream_t stream,
                                      mig_memory_resource const* mr)
{
  auto t = std::make_unique<mig_table>();
  init_table(*t, stream, mr);
  t->columns.resize(num_columns);
  for (int32_t c = 0; c < num_columns; c++) {
    owned_column& oc = t->columns[c];
    oc.col.type_id   = input[c].type_id;
    oc.col.scale     = input[c].scale;
    oc.col.size      = 0;
    if (input[c].type_id == MIG_TYPE_STRING) {
      oc.col.offsets_type_id = MIG_TYPE_INT32;
      oc.offsets_bytes       = sizeof(int32_t);
      oc.col.offsets         = t->alloc.allocate(oc.offsets_bytes);
      MIG_HIP_TRY(hipMemsetAsync(oc.col.offsets, 0, sizeof(int32_t), stream));
    }
  }
  MIG_HIP_TRY(hipStreamSynchronize(stream));
  return t;
}

// Regroups every column of `input` by the partition numbers in `pid` (one
// per row, below num_partitions) with a stable LSD counting sort; writes the
// first row of each partition to `offsets`.  The last counting pass moves up
// to HP_MAX_MOVE_COLS fixed-width columns itself (rows read in order, values
// written in runs per partition); validity, strings and further columns are
// gathered afterwards through the row map.
std::unique_ptr<mig_table> regroup(mig_column_view const* input, int32_t num_columns, int64_t n, uint32_t* pid,
                                   uint32_t num_partitions, call_scratch& S, hipStream_t stream,
                                   mig_memory_resource const* mr, std::vector<int32_t>& offsets)
{
  auto t = std::make_unique<mig_table>();
  init_table(*t, stream, mr);
  t->num_rows = int32_t(n);
  t->columns.resize(num_columns);
  std::vector<hp_move_col> moves;
  std::vector<int32_t> gathered;  // fixed-width columns beyond the moved ones
  std::vector<int32_t> str_cols;
  bool need_map = false;
  for (int32_t c = 0; c < num_columns; c++) {
    mig_column_view const& v = input[c];
    owned_column& oc         = t->columns[c];
    oc.col.type_id           = v.type_id;
    oc.col.scale             = v.scale;
    oc.col.size              = int32_t(n);
    if (v.null_mask) {
      oc.mask_bytes    = mask_bytes_for(n);
      oc.col.null_mask = static_cast<uint32_t*>(t->alloc.allocate(oc.mask_bytes));
      need_map         = true;
    }
    if (v.type_id == MIG_TYPE_STRING) {
      oc.col.offsets_type_id = MIG_TYPE_INT32;
      oc.offsets_bytes       = sizeof(int32_t) * size_t(n + 1);
      oc.col.offsets         = t->alloc.allocate(oc.offsets_bytes);
      str_cols.push_back(c);
      need_map = true;
      continue;
    }
    int32_t const w = element_width(v.type_id);
    oc.data_bytes   = size_t(w) * size_t(n);
    oc.col.data     = t->alloc.allocate(oc.data_bytes);
    if (int(moves.size()) < HP_MAX_MOVE_COLS) {
      moves.push_back(hp_move_col{static_cast<uint8_t const*>(v.data) + int64_t(w) * v.offset, oc.col.data, w, 0});
    } else {
      gathered.push_back(c);
      need_map = true;
    }
  }
  hp_move_col* d_moves = S.get<hp_move_col>(std::max<size_t>(1, moves.size()));
  if (!moves.empty())
    MIG_HIP_TRY(hipMemcpyAsync(d_moves, moves.data(), sizeof(hp_move_col) * moves.size(), hipMemcpyHostToDevice,
                               stream));

  int total_bits = 1;
  while (total_bits < 32 && (uint64_t(1) << total_bits) < num_partitions) total_bits++;
  int const passes        = (total_bits + HP_DIGIT_BITS - 1) / HP_DIGIT_BITS;
  int const bits_per_pass = (total_bits + passes - 1) / passes;

  uint32_t* pid_alt   = passes > 1 ? S.get<uint32_t>(n) : nullptr;
  uint32_t* rows_a    = (passes > 1 || need_map) ? S.get<uint32_t>(n) : nullptr;
  uint32_t* rows_b    = passes > 1 ? S.get<uint32_t>(n) : nullptr;
  uint32_t* pid_cur   = pid;
  uint32_t* rows_cur  = nullptr;  // identity before the first pass
  uint32_t* pid_next  = pid_alt;
  uint32_t* rows_next = rows_a;
  int shift           = 0;
  uint32_t const* last_hist = nullptr;
  int64_t last_W            = 0;
  for (int p = 0; p < passes; p++) {
    bool const last  = p == passes - 1;
    int const bits   = std::min(bits_per_pass, total_bits - shift);
    int64_t const nd = int64_t(1) << bits;
    // waves: enough to fill the device, few enough that the digit-major
    // histogram stays within 2^23 counters
    int64_t const w_cap = std::max<int64_t>(64, std::min<int64_t>(16384, (int64_t(1) << 23) / nd));
    int64_t W           = std::min<int64_t>(w_cap, std::max<int64_t>(1, (n + 255) / 256));
    int64_t const per   = (n + W - 1) / W;
    W                   = (n + per - 1) / per;
    uint32_t* hist = S.get<uint32_t>(size_t(nd * W + 1));
    launch_hp_count(pid_cur, n, per, W, shift, bits, hist, stream);
    launch_exclusive_scan_i32(reinterpret_cast<int32_t*>(hist), nd * W, reinterpret_cast<int32_t*>(hist) + nd * W,
                              stream);
    // partition numbers go out for a later pass or for the offsets of a
    // multi-pass sort (a single pass reads the offsets off its histogram);
    // the row map for a later pass or for the gathers
    launch_hp_scatter(pid_cur, rows_cur, n, per, W, shift, bits, hist, passes == 1 ? nullptr : pid_next,
                      (!last || need_map) ? rows_next : nullptr, last ? d_moves : nullptr,
                      last ? int(moves.size()) : 0, stream);
    MIG_HIP_TRY(hipGetLastError());
    last_hist = hist;
    last_W    = W;
    std::swap(pid_cur, pid_next);
    rows_cur  = rows_next;
    rows_next = rows_cur == rows_a ? rows_b : rows_a;
    shift += bits;
  }
  uint32_t const* map = rows_cur;

  int32_t* d_offsets = S.get<int32_t>(num_partitions);
  if (passes == 1) launch_hp_offsets_from_hist(last_hist, last_W, num_partitions, d_offsets, stream);
  else launch_hp_offsets(pid_cur, n, num_partitions, d_offsets, stream);

  uint32_t* d_nulls = S.get<uint32_t>(std::max(1, num_columns));
  MIG_HIP_TRY(hipMemsetAsync(d_nulls, 0, sizeof(uint32_t) * std::max(1, num_columns), stream));
  for (int32_t c = 0; c < num_columns; c++) {
    mig_column_view const& v = input[c];
    owned_column& oc         = t->columns[c];
    if (v.null_mask) launch_gather_mask(v.null_mask, v.offset, map, n, oc.col.null_mask, d_nulls + c, stream);
  }
  for (int32_t c : gathered) {
    mig_column_view const& v = input[c];
    int32_t const w          = element_width(v.type_id);
    launch_gather_fixed(static_cast<uint8_t const*>(v.data) + int64_t(w) * v.offset, w, map, n,
                        t->columns[c].col.data, stream);
  }
  for (int32_t c : str_cols) {
    auto* offs = static_cast<int32_t*>(t->columns[c].col.offsets);
    launch_gather_str_sizes(make_dcol(input[c], false), map, n, offs, stream);
    launch_exclusive_scan_i32(offs, n, offs + n, stream);
  }
  MIG_HIP_TRY(hipGetLastError());
  // chars of the string columns (sizes known once the scans are done)
  std::vector<int32_t> totals(str_cols.size());
  for (size_t k = 0; k < str_cols.size(); k++) {
    auto* offs = static_cast<int32_t*>(t->columns[str_cols[k]].col.offsets);
    MIG_HIP_TRY(hipMemcpyAsync(&totals[k], offs + n, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  }
  offsets.assign(num_partitions, 0);
  std::vector<uint32_t> nulls(std::max(1, num_columns));
  MIG_HIP_TRY(hipMemcpyAsync(offsets.data(), d_offsets, sizeof(int32_t) * num_partitions, hipMemcpyDeviceToHost,
                             stream));
  MIG_HIP_TRY(hipMemcpyAsync(nulls.data(), d_nulls, sizeof(uint32_t) * nulls.size(), hipMemcpyDeviceToHost, stream));
  MIG_HIP_TRY(hipStreamSynchronize(stream));
  for (size_t k = 0; k < str_cols.size(); k++) {
    mig_column_view const& v = input[str_cols[k]];
    owned_column& oc         = t->columns[str_cols[k]];
    MIG_EXPECTS(totals[k] >= 0, "partitioned string column exceeds 2^31 characters");
    oc.data_bytes = size_t(totals[k]);
    oc.col.data   = t->alloc.allocate(oc.data_bytes);
    launch_gather_string_chars(make_dcol(v, false), map, n, static_cast<int32_t const*>(oc.col.offsets),
                               static_cast<uint8_t*>(oc.col.data), stream);
  }
  for (int32_t c = 0; c < num_columns; c++)
    if (t->columns[c].col.null_mask) t->columns[c].col.null_count = int32_t(nulls[c]);
  MIG_HIP_TRY(hipGetLastError());
  MIG_HIP_TRY(hipStreamSynchronize(stream));
  return t;
}

}  // namespace

void lsd_sort_rows(lsd_key const* keys, int nkeys, int64_t n, uint32_t* map, uint32_t* last_sorted, hipStream_t s)
{
  MIG_EXPECTS(n >= 0 && n < (int64_t(1) << 32), "row count out of range for a row sort");
  if (n == 0) return;
  call_scratch S(s);
  uint32_t* pid_a    = S.get<uint32_t>(n);
  uint32_t* pid_b    = S.get<uint32_t>(n);
  uint32_t* rows_a   = S.get<uint32_t>(n);
  uint32_t* rows_b   = S.get<uint32_t>(n);
  uint32_t* rows_cur = nullptr;  // identity before the first pass
  for (int k = 0; k < nkeys; k++) {
    int const total_bits = std::max(1, std::min(32, keys[k].bits));
    // the key's values in the current order
    uint32_t* pid_cur = pid_a;
    if (rows_cur == nullptr)
      MIG_HIP_TRY(hipMemcpyAsync(pid_cur, keys[k].data, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    else
      launch_gather_fixed(keys[k].data, 4, rows_cur, n, pid_cur, s);
    uint32_t* pid_next      = pid_b;
    int const passes        = (total_bits + HP_DIGIT_BITS - 1) / HP_DIGIT_BITS;
    int const bits_per_pass = (total_bits + passes - 1) / passes;
    int shift               = 0;
    for (int p = 0; p < passes; p++) {
      int const bits   = std::min(bits_per_pass, total_bits - shift);
      int64_t const nd = int64_t(1) << bits;
      int64_t const w_cap = std::max<int64_t>(64, std::min<int64_t>(16384, (int64_t(1) << 23) / nd));
      int64_t W           = std::min<int64_t>(w_cap, std::max<int64_t>(1, (n + 255) / 256));
      int64_t const per   = (n + W - 1) / W;
      W                   = (n + per - 1) / per;
      uint32_t* hist      = S.get<uint32_t>(size_t(nd * W + 1));
      launch_hp_count(pid_cur, n, per, W, shift, bits, hist, s);
      launch_exclusive_scan_i32(reinterpret_cast<int32_t*>(hist), nd * W, reinterpret_cast<int32_t*>(hist) + nd * W,
                                s);
      uint32_t* rows_next = (rows_cur == rows_a) ? rows_b : rows_a;
      bool const final_pass = k == nkeys - 1 && p == passes - 1;
      if (final_pass) rows_next = map;
      bool const keep_pid = p < passes - 1 || (final_pass && last_sorted);
      uint32_t* pid_out   = final_pass && last_sorted ? last_sorted : pid_next;
      launch_hp_scatter(pid_cur, rows_cur, n, per, W, shift, bits, hist, keep_pid ? pid_out : nullptr, rows_next,
                        nullptr, 0, s);
      MIG_HIP_TRY(hipGetLastError());
      rows_cur = rows_next;
      std::swap(pid_cur, pid_next);
      shift += bits;
    }
  }
}

namespace {

void check_table(mig_column_view const* input, int32_t num_columns)
{
  MIG_EXPECTS(num_columns >= 0 && (num_columns == 0 || input != nullptr), "invalid table");
  for (int32_t c = 0; c < num_columns; c++) {
    check_column_type(input[c]);
    MIG_EXPECTS(input[c].size == input[0].size, "all columns must have the same number of rows");
    MIG_EXPECTS(input[c].size >= 0 && input[c].offset >= 0, "negative column size or offset");
  }
}

}  // namespace

std::unique_ptr<mig_table> hash_partition(mig_column_view const* input, int32_t num_columns,
                                          int32_t const* columns_to_hash, int32_t num_columns_to_hash,
                                          int32_t num_partitions, int32_t hash_function, uint32_t seed,
                                          hipStream_t stream, mig_memory_resource const* mr,
                                          std::vector<int32_t>& offsets)
{
  func_range range("cudf::hash_partition");
  check_table(input, num_columns);
  MIG_EXPECTS(num_columns_to_hash >= 0 && (num_columns_to_hash == 0 || columns_to_hash != nullptr),
              "invalid columns_to_hash");
  if (hash_function != MIG_HASH_IDENTITY && hash_function != MIG_HASH_MURMUR3)
    fail(MIG_ERROR_LOGIC, "Unsupported hash function in hash_partition");
  for (int32_t k = 0; k < num_columns_to_hash; k++) {
    int32_t const c = columns_to_hash[k];
    if (c < 0 || c >= num_columns)
      fail(MIG_ERROR_OUT_OF_RANGE, "column index " + std::to_string(c) + " out of range");
    if (hash_function == MIG_HASH_IDENTITY && !is_numeric_type(input[c].type_id))
      fail(MIG_ERROR_LOGIC, "IdentityHash does not support this data type");
    if (storage_of(input[c].type_id) == ST_NONE)
      fail(MIG_ERROR_DATA_TYPE, "unsupported column type " + std::to_string(input[c].type_id) + " for hashing");
  }
  int64_t const n = num_columns > 0 ? input[0].size : 0;
  if (num_partitions <= 0 || n == 0 || num_columns_to_hash == 0) {
    offsets.assign(size_t(std::max(0, num_partitions)), 0);
    return empty_like(input, num_columns, stream, mr);
  }
  MIG_EXPECTS(num_columns_to_hash <= MAX_KEYS, "too many columns to hash");

  call_scratch S(stream);
  std::vector<dcol> h(num_columns_to_hash);
  for (int32_t k = 0; k < num_columns_to_hash; k++) {
    mig_column_view const& v = input[columns_to_hash[k]];
    // honouring a mask that has no nulls hashes the same, so an unknown null
    // count needs no device count here
    h[k] = make_dcol(v, v.null_mask != nullptr && v.null_count != 0);
  }
  dcol* d_cols = S.get<dcol>(h.size());
  MIG_HIP_TRY(hipMemcpyAsync(d_cols, h.data(), sizeof(dcol) * h.size(), hipMemcpyHostToDevice, stream));
  uint32_t* pid = S.get<uint32_t>(n);
  launch_hp_partition_ids(d_cols, num_columns_to_hash, n, seed, hash_function == MIG_HASH_IDENTITY,
                          uint32_t(num_partitions), pid, stream);
  MIG_HIP_TRY(hipGetLastError());
  return regroup(input, num_columns, n, pid, uint32_t(num_partitions), S, stream, mr, offsets);
}

std::unique_ptr<mig_table> partition_by_map(mig_column_view const* input, int32_t num_columns,
                                            mig_column_view const& partition_map, int32_t num_partitions,
                                            hipStream_t stream, mig_memory_resource const* mr,
                                            std::vector<int32_t>& offsets)
{
  func_range range("cudf::partition");
  check_table(input, num_columns);
  int64_t const n = num_columns > 0 ? input[0].size : 0;
  MIG_EXPECTS(partition_map.size == n, "Size mismatch between table and partition map.");
  MIG_EXPECTS(partition_map.offset >= 0, "negative column offset");
  MIG_EXPECTS(null_count_of(partition_map, stream) == 0, "Unexpected null values in partition_map.");
  int32_t const st = storage_of(partition_map.type_id);
  bool const integral = st == ST_I8 || st == ST_I16 || st == ST_I32 || st == ST_I64 || st == ST_U8 ||
                        st == ST_U16 || st == ST_U32 || st == ST_U64;
  if (!integral || is_chrono_type(partition_map.type_id) || partition_map.type_id == MIG_TYPE_DECIMAL32 ||
      partition_map.type_id == MIG_TYPE_DECIMAL64)
    fail(MIG_ERROR_LOGIC, "Unexpected, non-integral partition map.");
  MIG_EXPECTS(num_partitions >= 0, "negative number of partitions");
  if (num_partitions == 0 || n == 0) {
    offsets.assign(size_t(num_partitions) + 1, 0);
    return empty_like(input, num_columns, stream, mr);
  }
  call_scratch S(stream);
  uint32_t* pid = S.get<uint32_t>(n);
  uint32_t* bad = S.get<uint32_t>(1);
  MIG_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(uint32_t), stream));
  auto const* map = static_cast<uint8_t const*>(partition_map.data) +
                    int64_t(storage_width(st)) * partition_map.offset;
  launch_map_partition_ids(map, st, n, uint32_t(num_partitions), pid, bad, stream);
  MIG_HIP_TRY(hipGetLastError());
  uint32_t h_bad = 0;
  MIG_HIP_TRY(hipMemcpyAsync(&h_bad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  MIG_HIP_TRY(hipStreamSynchronize(stream));
  MIG_EXPECTS(h_bad == 0, "partition_map values must lie in [0, num_partitions)");
  auto t = regroup(input, num_columns, n, pid, uint32_t(num_partitions), S, stream, mr, offsets);
  offsets.push_back(int32_t(n));
  return t;
}

std::unique_ptr<mig_table> concatenate(mig_column_view const* columns, int32_t num_tables, int32_t num_columns,
                                       hipStream_t stream, mig_memory_resource const* mr)
{
  func_range range("cudf::concatenate");
  MIG_EXPECTS(num_tables >= 0 && num_columns >= 0, "negative table or column count");
  auto t = std::make_unique<mig_table>();
  init_table(*t, stream, mr);
  if (num_tables == 0) return t;
  MIG_EXPECTS(num_columns == 0 || columns != nullptr, "invalid tables");
  auto at = [&](int32_t tb, int32_t c) -> mig_column_view const& { return columns[int64_t(tb) * num_columns + c]; };
  int64_t total = 0;
  for (int32_t tb = 0; tb < num_tables; tb++) {
    check_table(&at(tb, 0), num_columns);
    if (num_columns > 0) total += at(tb, 0).size;
  }
  if (total > int64_t(INT32_MAX))
    fail(MIG_ERROR_OVERFLOW, "Total number of concatenated rows exceeds the column size limit");
  for (int32_t c = 0; c < num_columns; c++)
    for (int32_t tb = 1; tb < num_tables; tb++)
      if (at(tb, c).type_id != at(0, c).type_id)
        fail(MIG_ERROR_DATA_TYPE, "Type mismatch in columns to concatenate.");
  t->num_rows = int32_t(total);
  t->columns.resize(num_columns);

  call_scratch S(stream);
  // host copies stay alive until the final synchronisation (the uploads are
  // asynchronous)
  std::vector<std::vector<concat_piece>> host_pieces(std::max(1, num_columns),
                                                     std::vector<concat_piece>(num_tables));
  concat_piece* d_pieces = S.get<concat_piece>(size_t(num_tables) * std::max(1, num_columns));
  int64_t* d_ranges      = S.get<int64_t>(2 * size_t(num_tables));
  uint32_t* d_nulls      = S.get<uint32_t>(std::max(1, num_columns));
  MIG_HIP_TRY(hipMemsetAsync(d_nulls, 0, sizeof(uint32_t) * std::max(1, num_columns), stream));
  for (int32_t c = 0; c < num_columns; c++) {
    owned_column& oc = t->columns[c];
    auto& pieces     = host_pieces[c];
    oc.col.type_id   = at(0, c).type_id;
    oc.col.scale     = at(0, c).scale;
    oc.col.size      = int32_t(total);
    bool any_nulls   = false;
    int64_t row      = 0;
    for (int32_t tb = 0; tb < num_tables; tb++) {
      mig_column_view const& v = at(tb, c);
      concat_piece& P          = pieces[tb];
      P                        = concat_piece{};
      // a piece with an unknown null count is read through its mask; a
      // result mask without nulls is dropped at the end
      P.mask                   = (v.null_mask != nullptr && v.null_count != 0) ? v.null_mask : nullptr;
      P.offsets                = v.offsets;
      P.offsets64              = v.offsets_type_id == MIG_TYPE_INT64;
      P.row_start              = row;
      P.offset                 = v.offset;
      P.rows                   = v.size;
      any_nulls                = any_nulls || P.mask != nullptr;
      row += v.size;
    }
    concat_piece* dp = d_pieces + int64_t(c) * num_tables;
    if (oc.col.type_id == MIG_TYPE_STRING) {
      MIG_HIP_TRY(hipMemcpyAsync(dp, pieces.data(), sizeof(concat_piece) * num_tables, hipMemcpyHostToDevice,
                                 stream));
      launch_concat_char_ranges(dp, num_tables, d_ranges, stream);
      std::vector<int64_t> ranges(2 * size_t(num_tables));
      MIG_HIP_TRY(hipMemcpyAsync(ranges.data(), d_ranges, sizeof(int64_t) * ranges.size(), hipMemcpyDeviceToHost,
                                 stream));
      MIG_HIP_TRY(hipStreamSynchronize(stream));
      int64_t chars = 0;
      for (int32_t tb = 0; tb < num_tables; tb++) {
        pieces[tb].chars_begin = ranges[2 * tb];
        pieces[tb].chars_start = chars;
        chars += ranges[2 * tb + 1] - ranges[2 * tb];
      }
      if (chars > int64_t(INT32_MAX))
        fail(MIG_ERROR_OVERFLOW, "Total number of concatenated chars exceeds the column size limit");
      MIG_HIP_TRY(hipMemcpyAsync(dp, pieces.data(), sizeof(concat_piece) * num_tables, hipMemcpyHostToDevice,
                                 stream));
      oc.col.offsets_type_id = MIG_TYPE_INT32;
      oc.offsets_bytes       = sizeof(int32_t) * size_t(total + 1);
      oc.col.offsets         = t->alloc.allocate(oc.offsets_bytes);
      launch_concat_str_offsets(dp, num_tables, total, chars, static_cast<int32_t*>(oc.col.offsets), stream);
      oc.data_bytes = size_t(chars);
      oc.col.data   = t->alloc.allocate(oc.data_bytes);
      for (int32_t tb = 0; tb < num_tables; tb++) {
        int64_t const len = ranges[2 * tb + 1] - ranges[2 * tb];
        if (len > 0)
          MIG_HIP_TRY(hipMemcpyAsync(static_cast<uint8_t*>(oc.col.data) + pieces[tb].chars_start,
                                     static_cast<uint8_t const*>(at(tb, c).data) + ranges[2 * tb], size_t(len),
                                     hipMemcpyDeviceToDevice, stream));
      }
    } else {
      MIG_HIP_TRY(hipMemcpyAsync(dp, pieces.data(), sizeof(concat_piece) * num_tables, hipMemcpyHostToDevice,
                                 stream));
      int32_t const w = element_width(oc.col.type_id);
      oc.data_bytes   = size_t(w) * size_t(total);
      oc.col.data     = t->alloc.allocate(oc.data_bytes);
      for (int32_t tb = 0; tb < num_tables; tb++) {
        mig_column_view const& v = at(tb, c);
        if (v.size > 0)
          MIG_HIP_TRY(hipMemcpyAsync(static_cast<uint8_t*>(oc.col.data) + int64_t(w) * pieces[tb].row_start,
                                     static_cast<uint8_t const*>(v.data) + int64_t(w) * v.offset,
                                     size_t(w) * size_t(v.size), hipMemcpyDeviceToDevice, stream));
      }
    }
    if (any_nulls) {
      oc.mask_bytes    = mask_bytes_for(total);
      oc.col.null_mask = static_cast<uint32_t*>(t->alloc.allocate(oc.mask_bytes));
      launch_concat_mask(dp, num_tables, total, oc.col.null_mask, d_nulls + c, stream);
    }
    MIG_HIP_TRY(hipGetLastError());
  }
  std::vector<uint32_t> nulls(std::max(1, num_columns));
  MIG_HIP_TRY(hipMemcpyAsync(nulls.data(), d_nulls, sizeof(uint32_t) * nulls.size(), hipMemcpyDeviceToHost, stream));
  MIG_HIP_TRY(hipStreamSynchronize(stream));
  for (int32_t c = 0; c < num_columns; c++) {
    owned_column& oc = t->columns[c];
    if (!oc.col.null_mask) continue;
    oc.col.null_count = int32_t(nulls[c]);
    if (nulls[c] == 0) {  // no piece had nulls: no mask, as for inputs without nulls
      t->alloc.deallocate(oc.col.null_mask, oc.mask_bytes);
      oc.col.null_mask = nullptr;
      oc.mask_bytes    = 0;
    }
  }
  return t;
}

}  // namespace mig
