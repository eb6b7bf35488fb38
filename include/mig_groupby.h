/*
 * mig_groupby.h -- C ABI of the MI355X (gfx950) hash group-by library.
 *
 * This is the drop-in boundary for libcudf's hash group-by path
 * (`cudf::groupby::groupby(keys).aggregate(requests)`).  Every entry point takes
 * plain pointers and sizes; no C++ or HIP types appear in the signatures, so the
 * header can be bound from ctypes, JNI, cgo or Cython alike.  The C++ facade in
 * `include/cudf/` (namespace `cudf`) is written on top of these functions.
 *
 * Conventions
 *   - Every function that can fail returns a `mig_status`; on failure the
 *     calling thread's message is available from `mig_last_error()`.
 *   - Device pointers are HIP device addresses on the current device.
 *   - `stream` arguments are `hipStream_t` values passed as `void*`; NULL means
 *     the device's null stream.
 *   - Validity bitmasks are Arrow/cudf style: 32-bit words, least significant
 *     bit first, bit set = row valid (cudf/utilities/bit.hpp:128-132).
 */
#ifndef MIG_GROUPBY_H
#define MIG_GROUPBY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIG_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
/* Error classes mirror cudf/utilities/error.hpp:64-127 (logic_error,
 * cuda_error, data_type_error) plus allocation failure. */
typedef enum mig_status {
  MIG_OK              = 0,
  MIG_ERROR_LOGIC     = 1, /* cudf::logic_error: invalid request / sizes      */
  MIG_ERROR_HIP       = 2, /* cudf::cuda_error equivalent: a HIP call failed  */
  MIG_ERROR_DATA_TYPE = 3, /* cudf::data_type_error: unsupported type combo   */
  MIG_ERROR_ALLOC     = 4, /* device allocation failed                        */
  MIG_ERROR_OVERFLOW  = 5, /* a fixed-capacity buffer would be exceeded, or a
                              result would exceed 2^31-1 rows (std::overflow_error) */
  MIG_ERROR_OUT_OF_RANGE = 6 /* std::out_of_range: a column index outside the table */
} mig_status;

/* ---- type ids: same numbering as cudf::type_id (cudf/types.hpp:203-235) -- */
enum {
  MIG_TYPE_EMPTY = 0,
  MIG_TYPE_INT8,
  MIG_TYPE_INT16,
  MIG_TYPE_INT32,
  MIG_TYPE_INT64,
  MIG_TYPE_UINT8,
  MIG_TYPE_UINT16,
  MIG_TYPE_UINT32,
  MIG_TYPE_UINT64,
  MIG_TYPE_FLOAT32,
  MIG_TYPE_FLOAT64,
  MIG_TYPE_BOOL8,
  MIG_TYPE_TIMESTAMP_DAYS,
  MIG_TYPE_TIMESTAMP_SECONDS,
  MIG_TYPE_TIMESTAMP_MILLISECONDS,
  MIG_TYPE_TIMESTAMP_MICROSECONDS,
  MIG_TYPE_TIMESTAMP_NANOSECONDS,
  MIG_TYPE_DURATION_DAYS,
  MIG_TYPE_DURATION_SECONDS,
  MIG_TYPE_DURATION_MILLISECONDS,
  MIG_TYPE_DURATION_MICROSECONDS,
  MIG_TYPE_DURATION_NANOSECONDS,
  MIG_TYPE_DICTIONARY32,
  MIG_TYPE_STRING,
  MIG_TYPE_LIST,
  MIG_TYPE_DECIMAL32,
  MIG_TYPE_DECIMAL64,
  MIG_TYPE_DECIMAL128,
  MIG_TYPE_STRUCT,
  MIG_NUM_TYPE_IDS
};

/* ---- aggregation kinds: same numbering as cudf::aggregation::Kind
 *      (cudf/aggregation.hpp:86-122).  The hash path accepts the 12 kinds
 *      listed in hash/groupby.cu:73-84, all implemented here: SUM, PRODUCT,
 *      MIN, MAX, COUNT_VALID, COUNT_ALL, SUM_OF_SQUARES, MEAN, VARIANCE, STD,
 *      ARGMIN, ARGMAX (string values: COUNT_*, MIN, MAX, ARGMIN, ARGMAX).
 *      MEDIAN, QUANTILE, NUNIQUE and NTH_ELEMENT are sort-path aggregations
 *      in the reference (groupby.cu:78-83, sort/aggregate.cpp:323-410); a
 *      request that holds one runs the sort-based group-by, for fixed-width
 *      values (a request there with string values returns
 *      MIG_ERROR_DATA_TYPE). --- */
enum {
  MIG_AGG_SUM            = 0,
  MIG_AGG_PRODUCT        = 1,
  MIG_AGG_MIN            = 2,
  MIG_AGG_MAX            = 3,
  MIG_AGG_COUNT_VALID    = 4,
  MIG_AGG_COUNT_ALL      = 5,
  MIG_AGG_ANY            = 6,
  MIG_AGG_ALL            = 7,
  MIG_AGG_SUM_OF_SQUARES = 8,
  MIG_AGG_MEAN           = 9,
  MIG_AGG_M2             = 10,
  MIG_AGG_VARIANCE       = 11,
  MIG_AGG_STD            = 12,
  MIG_AGG_MEDIAN         = 13,
  MIG_AGG_QUANTILE       = 14,
  MIG_AGG_ARGMAX         = 15,
  MIG_AGG_ARGMIN         = 16,
  MIG_AGG_NUNIQUE        = 17,
  MIG_AGG_NTH_ELEMENT    = 18
};

/* cudf::null_policy (cudf/types.hpp:126-129). */
enum { MIG_NULL_EXCLUDE = 0, MIG_NULL_INCLUDE = 1 };

/* cudf::interpolation (cudf/types.hpp:192-198), used by QUANTILE. */
enum {
  MIG_INTERP_LINEAR   = 0,
  MIG_INTERP_LOWER    = 1,
  MIG_INTERP_HIGHER   = 2,
  MIG_INTERP_MIDPOINT = 3,
  MIG_INTERP_NEAREST  = 4
};

/* Non-owning view of one column; mirrors cudf::column_view_base fields
 * (cudf/column/column_view.hpp:240-247).  For STRING columns `data` is the
 * chars buffer and `offsets` the offsets child (child 0, size+1 entries,
 * INT32 or INT64, indices absolute into `data`), as in
 * cudf/column/column_device_view.cuh:444-453. */
typedef struct mig_column_view {
  int32_t type_id;          /* MIG_TYPE_*                                      */
  int32_t scale;            /* fixed-point scale (unused by this library)      */
  int32_t size;             /* rows visible through the view (cudf::size_type) */
  int32_t offset;           /* first row of the view inside data / null_mask   */
  const void* data;         /* element array (fixed width) or chars (STRING)   */
  const uint32_t* null_mask;/* validity words, NULL = no mask (all valid)      */
  int32_t null_count;       /* nulls in [offset, offset+size); -1 = unknown    */
  int32_t offsets_type_id;  /* STRING only: MIG_TYPE_INT32 or MIG_TYPE_INT64   */
  const void* offsets;      /* STRING only: offsets child                      */
} mig_column_view;

/* One aggregation of a request (cudf::groupby_aggregation).  `ddof` is used
 * by VARIANCE and STD (cudf/aggregation.hpp make_variance_aggregation),
 * `null_handling` by NUNIQUE (make_nunique_aggregation: MIG_NULL_EXCLUDE
 * does not count nulls, MIG_NULL_INCLUDE counts them as one value) and by
 * NTH_ELEMENT (make_nth_element_aggregation, whose default is INCLUDE:
 * EXCLUDE counts valid values only), `n` by NTH_ELEMENT (the group's n-th
 * row, counted from the end when negative), `quantiles` /
 * `num_quantiles` / `interpolation` by QUANTILE
 * (make_quantile_aggregation(quantiles, interp); host memory, read during
 * the call).  Fields a kind does not use may be left zero. */
typedef struct mig_aggregation {
  int32_t kind; /* MIG_AGG_* */
  int32_t ddof;
  int32_t null_handling;
  int32_t n;
  int32_t interpolation; /* MIG_INTERP_* */
  int32_t num_quantiles;
  const double* quantiles;
} mig_aggregation;

/* cudf::groupby::aggregation_request (cudf/groupby.hpp:61-64). */
typedef struct mig_aggregation_request {
  mig_column_view values;
  const mig_aggregation* aggregations;
  int32_t num_aggregations;
} mig_aggregation_request;

/* Device memory resource used for result columns (the `mr` argument of
 * groupby::aggregate, cudf/groupby.hpp:187-199).  A NULL resource selects the
 * library's own stream-ordered allocator (hipMallocAsync). */
typedef struct mig_memory_resource {
  void* ctx;
  void* (*allocate)(void* ctx, size_t bytes, void* stream);
  void (*deallocate)(void* ctx, void* ptr, size_t bytes, void* stream);
} mig_memory_resource;

/* Column owned by a result handle (borrowed by the caller). */
typedef struct mig_column {
  int32_t type_id;
  int32_t size;
  void* data;               /* elements, or chars for STRING                  */
  uint32_t* null_mask;      /* NULL if the column has no mask                 */
  int32_t null_count;
  int32_t offsets_type_id;  /* STRING: MIG_TYPE_INT32                         */
  void* offsets;            /* STRING: size+1 offsets                         */
  int64_t data_bytes;       /* bytes allocated for `data` (>= size * width)   */
  int64_t mask_bytes;       /* bytes allocated for `null_mask`                */
  int64_t offsets_bytes;    /* bytes allocated for `offsets`                  */
  int32_t scale;            /* DECIMAL32 / DECIMAL64: fixed-point scale (the
                               value is data * 10^scale); 0 otherwise        */
} mig_column;

typedef struct mig_groupby_result mig_groupby_result;

/* ---- batched group-by -------------------------------------------------------
 * Many independent group-bys over different inputs with the same requests and
 * aggregations, executed together: one persistent work-queue kernel
 * aggregates every instance, captured with the kernels around it into one
 * hipGraph that each execute replays.  Instances may differ in row count,
 * key type and value types (fixed-width columns; one key column each).  No
 * counterpart in the reference, which runs one aggregate() per table
 * (cpp/src/groupby/groupby.cu:204-221); each instance's result equals what
 * mig_groupby_aggregate returns for it. */
typedef struct mig_batch_instance {
  const mig_column_view* keys;   /* num_keys key columns (one fixed-width key) */
  const mig_column_view* values; /* one values column per request             */
} mig_batch_instance;

/* Tuning knobs (0 = automatic); chosen by the autotuner in cudf_amd.autotune. */
typedef struct mig_batch_options {
  double load_factor;        /* target groups per table slot (default 0.5)      */
  int32_t lds_max_groups;    /* LDS pre-aggregation for instances estimated at
                                or below this many groups (default 512; -1 off) */
  int32_t block_size;        /* 256 or 512 (default 256)                        */
  int32_t tile_rows;         /* rows per work item (default 4096)               */
  int32_t use_graph;         /* -1 = launch the kernels one by one instead of
                                replaying the captured hipGraph (default on)    */
  int32_t partition_min_rows; /* instances with at least this many rows and an
                                estimated 65536 or more groups run the single-
                                call path (radix-partitioned aggregation) inside
                                mig_batch_execute (default 1 << 20; -1 never)   */
} mig_batch_options;

typedef struct mig_batch_plan mig_batch_plan;

/* Builds a plan: estimates every instance's group count on the device, sizes
 * tables and output regions, and (by default) captures the execution graph.
 * The instance columns must stay valid and unchanged while the plan is used. */
mig_status mig_batch_plan_create(const mig_batch_instance* instances,
                                 int32_t num_instances,
                                 int32_t num_keys,
                                 int32_t null_policy,
                                 const mig_aggregation* aggregations,
                                 const int32_t* num_aggregations,
                                 int32_t num_requests,
                                 const mig_batch_options* options,
                                 void* stream,
                                 mig_batch_plan** out);
/* Runs every instance on `stream` (enqueued; use mig_batch_synchronize or the
 * stream to wait). */
mig_status mig_batch_execute(mig_batch_plan* plan, void* stream);
/* Waits for the last execute and reads back the per-instance group counts. */
mig_status mig_batch_synchronize(mig_batch_plan* plan, void* stream);
int32_t mig_batch_num_groups(const mig_batch_plan* plan, int32_t instance);
/* Borrowed views of an instance's results, valid until the next execute or
 * mig_batch_plan_free. */
mig_status mig_batch_result_key(const mig_batch_plan* plan, int32_t instance, mig_column* out);
mig_status mig_batch_result_column(const mig_batch_plan* plan, int32_t instance, int32_t request,
                                   int32_t agg, mig_column* out);
void mig_batch_plan_free(mig_batch_plan* plan);

/* ---- library -------------------------------------------------------------- */

/* Message for the last failed call on this thread ("" if none). */
const char* mig_last_error(void);

/* MIG_ABI_VERSION of the loaded library. */
int32_t mig_abi_version(void);

/* Device memory of the library's default resource (stream-ordered
 * hipMallocAsync / hipFreeAsync): what result columns use when no resource
 * is passed, and what a caller uses to free them.  Plays the role of
 * rmm::mr::get_current_device_resource() in the reference API
 * (cudf/groupby.hpp:187-199).  Returns NULL on failure. */
void* mig_device_allocate(size_t bytes, void* stream);
void mig_device_deallocate(void* ptr, size_t bytes, void* stream);

/* ---- row hashing -----------------------------------------------------------
 * Replaces cudf::hashing::murmurhash3_x86_32(table_view, seed, stream, mr)
 * (cpp/include/cudf/hashing.hpp:61; src/hash/murmurhash3_x86_32.cu:34-57):
 * MurmurHash3_x86_32 of every row, elements combined with boost hash_combine,
 * null elements hashing to 0xFFFFFFFF.  `out` is a device array of
 * columns[0].size uint32 values. */
mig_status mig_hash_murmurhash3_x86_32(const mig_column_view* columns,
                                       int32_t num_columns,
                                       uint32_t seed,
                                       uint32_t* out,
                                       void* stream);

/* ---- hash group-by -----------------------------------------------------------
 * Replaces cudf::groupby::groupby(keys, null_policy).aggregate(requests, stream,
 * mr) on the hash path (cpp/src/groupby/groupby.cu:53-64, 204-221;
 * cpp/src/groupby/hash/groupby.cu:548-667).  On success `*out` owns the unique
 * keys (output row order unspecified, as in cudf/groupby.hpp:155-157) and one
 * column per requested aggregation, in request order.  Blocks until the
 * results are complete on `stream`. */
mig_status mig_groupby_aggregate(const mig_column_view* keys,
                                 int32_t num_keys,
                                 int32_t null_policy,
                                 const mig_aggregation_request* requests,
                                 int32_t num_requests,
                                 void* stream,
                                 const mig_memory_resource* mr,
                                 mig_groupby_result** out);

/* Replaces cudf::groupby::groupby(keys, null_policy, sorted::YES).aggregate(
 * requests, stream, mr): keys the caller promises are sorted take the
 * reference's sort path with keys_pre_sorted (cpp/src/groupby/groupby.cu:66-84;
 * sort/sort_helper.cu:55-72, 133-181).  Groups are the runs of equal adjacent
 * key rows (null keys equal each other), returned in input order, so sorted
 * keys come back sorted; unsorted keys give one group per run.  Rows with a
 * null key are dropped under MIG_NULL_EXCLUDE (the reference re-sorts such
 * keys first, which gives the same groups when the keys are sorted).  Same
 * aggregations, results and errors as mig_groupby_aggregate, except that
 * MIN / MAX / ARGMIN / ARGMAX of string values return MIG_ERROR_LOGIC. */
mig_status mig_groupby_aggregate_sorted(const mig_column_view* keys,
                                        int32_t num_keys,
                                        int32_t null_policy,
                                        const mig_aggregation_request* requests,
                                        int32_t num_requests,
                                        void* stream,
                                        const mig_memory_resource* mr,
                                        mig_groupby_result** out);

/* The hash group-by with no device allocation: no counterpart in the
 * reference (whose aggregate() allocates through its memory resource,
 * cpp/src/groupby/hash/groupby.cu:404, 478).  For callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same requests, results and errors as
 * mig_groupby_aggregate on its device-wide hash table path, for fixed-width
 * key and value columns and the hash aggregation kinds (sort-based kinds,
 * strings and pre-sorted keys return MIG_ERROR_LOGIC).
 *   - Scratch: the library's fixed device array of
 *     mig_groupby_into_scratch_bytes() bytes (128 MiB) per device, held until
 *     the process ends.  A call that needs more (a hash table for very many
 *     groups) returns MIG_ERROR_OVERFLOW without touching the outputs.  Every
 *     call on one device shares that array: calls that overlap on a device
 *     (several streams or threads at once) race and must not be made.
 *   - Outputs: the caller's buffers.  `out_keys` holds num_keys columns and
 *     `out_results` one column per aggregation, in request order (requests
 *     flattened).  On entry each column's `data` / `data_bytes` give a buffer
 *     for the result rows and `null_mask` / `mask_bytes` one for whole 64-bit
 *     validity words (8 * ceil(rows / 64) bytes); a buffer too small for the
 *     groups found returns MIG_ERROR_OVERFLOW (rows = the key count bounds
 *     the groups).  On success every column's type_id, size (= *num_groups)
 *     and null_count are set, and null_mask is set to NULL when the result has
 *     no validity mask (the caller's buffer is left alone).
 * Blocks until the results are complete on `stream`. */
mig_status mig_groupby_aggregate_into(const mig_column_view* keys,
                                      int32_t num_keys,
                                      int32_t null_policy,
                                      const mig_aggregation_request* requests,
                                      int32_t num_requests,
                                      mig_column* out_keys,
                                      mig_column* out_results,
                                      int32_t* num_groups,
                                      void* stream);
/* Bytes of the fixed device scratch mig_groupby_aggregate_into uses. */
size_t mig_groupby_into_scratch_bytes(void);

int32_t mig_result_num_groups(const mig_groupby_result* r);
int32_t mig_result_num_keys(const mig_groupby_result* r);
int32_t mig_result_num_requests(const mig_groupby_result* r);
int32_t mig_result_num_aggregations(const mig_groupby_result* r, int32_t request);
/* Borrow key column `i` / result column (`request`, `agg`). */
mig_status mig_result_key(const mig_groupby_result* r, int32_t i, mig_column* out);
mig_status mig_result_column(const mig_groupby_result* r,
                             int32_t request,
                             int32_t agg,
                             mig_column* out);
/* Take ownership of a key / result column: afterwards the caller frees its
 * buffers with the resource they came from and the handle forgets them. */
mig_status mig_result_release_key(mig_groupby_result* r, int32_t i, mig_column* out);
mig_status mig_result_release_column(mig_groupby_result* r,
                                     int32_t request,
                                     int32_t agg,
                                     mig_column* out);
/* Free the handle and every column it still owns. */
void mig_result_free(mig_groupby_result* r);

/* Statistics of the last mig_groupby_aggregate call on this thread, for
 * benchmarks and the autotuner (strategy chosen, table capacity, time per
 * phase).  Writes at most `n` doubles; returns the number available.
 * Index: 0 strategy, 1 key mode, 2 table slots, 3 estimated groups,
 * 4 groups, 5 table rebuilds, 6 state bytes, then phase times in ms
 * (7 estimate, 8 init, 9 accumulate, 10 second pass, 11 compact,
 * 12 outputs) when phase timing is on. */
int32_t mig_last_stats(double* values, int32_t n);

/* Turns hipEvent timing of each phase of mig_groupby_aggregate on (1) or off
 * (0) for the calling process, or (2) times only the main aggregation kernel
 * (two events per call); events are recorded on the call's stream.
 * Instrumentation only, with no counterpart in the reference (which marks
 * the call with an NVTX range, cpp/src/groupby/groupby.cu:209). */
void mig_set_phase_timing(int32_t enable);

/* ---- partitioning and concatenation ----------------------------------------
 * The building blocks of a multi-device group-by: each device aggregates its
 * rows, hash-partitions the partial results by key, exchanges partitions, and
 * concatenates what it receives before aggregating again (dask_cudf's
 * shuffle, python/dask_cudf/dask_cudf/groupby.py:367-432). */

/* cudf::hash_id (cpp/include/cudf/partitioning.hpp:41-44). */
enum { MIG_HASH_IDENTITY = 0, MIG_HASH_MURMUR3 = 1 };

/* Columns owned by a handle (a cudf::table). */
typedef struct mig_table mig_table;
int32_t mig_table_num_rows(const mig_table* t);
int32_t mig_table_num_columns(const mig_table* t);
/* Borrow column `i`. */
mig_status mig_table_column(const mig_table* t, int32_t i, mig_column* out);
/* Take ownership of column `i` (freed with the resource it came from). */
mig_status mig_table_release_column(mig_table* t, int32_t i, mig_column* out);
void mig_table_free(mig_table* t);

/* Replaces cudf::hash_partition(input, columns_to_hash, num_partitions,
 * hash_function, seed, stream, mr) (cpp/include/cudf/partitioning.hpp:105-112;
 * src/partitioning/partitioning.cu:759-830).  Row r goes to partition
 * row_hash(r) % num_partitions, where row_hash combines the MurmurHash3_x86_32
 * (or IdentityHash) of the hashed columns from `seed` with hash_combine, as
 * mig_hash_murmurhash3_x86_32 does.  `*out` holds the rows regrouped so that
 * each partition is contiguous; inside a partition rows keep their input
 * order.  `partition_offsets` (num_partitions entries, host memory) receives
 * the first row of each partition.  Errors: MIG_ERROR_OUT_OF_RANGE for a bad
 * column index; MIG_ERROR_LOGIC for an unknown hash function or IdentityHash
 * on a non-numeric column.  No partitions, rows or hashed columns give empty
 * columns of the input's types and zero offsets. */
mig_status mig_hash_partition(const mig_column_view* input,
                              int32_t num_columns,
                              const int32_t* columns_to_hash,
                              int32_t num_columns_to_hash,
                              int32_t num_partitions,
                              int32_t hash_function,
                              uint32_t seed,
                              void* stream,
                              const mig_memory_resource* mr,
                              mig_table** out,
                              int32_t* partition_offsets);

/* Replaces cudf::partition(t, partition_map, num_partitions, mr)
 * (cpp/include/cudf/partitioning.hpp:79-83; partitioning.cu:782-802): row r
 * goes to partition partition_map[r] (a non-null integer column).
 * `partition_offsets` receives num_partitions + 1 entries, the last being the
 * row count.  A map value outside [0, num_partitions) is an error. */
mig_status mig_partition(const mig_column_view* input,
                         int32_t num_columns,
                         const mig_column_view* partition_map,
                         int32_t num_partitions,
                         void* stream,
                         const mig_memory_resource* mr,
                         mig_table** out,
                         int32_t* partition_offsets);

/* Replaces cudf::concatenate(tables_to_concat, stream, mr)
 * (cpp/include/cudf/concatenate.hpp:95-98; src/copying/concatenate.cu):
 * `tables` holds num_tables * num_columns views, table-major
 * (tables[t * num_columns + c]).  Column types must match across tables
 * (MIG_ERROR_DATA_TYPE); more than 2^31-1 rows is MIG_ERROR_OVERFLOW.  An
 * output column has a validity mask when some input piece has nulls. */
mig_status mig_concatenate(const mig_column_view* tables,
                           int32_t num_tables,
                           int32_t num_columns,
                           void* stream,
                           const mig_memory_resource* mr,
                           mig_table** out);

/* Aggregation strategies.  AUTO picks per call from the key layout and a
 * sampled estimate of the number of groups; the others force one (used by
 * tests and the autotuner).  No counterpart in the reference, whose hash
 * path has a single strategy (cpp/src/groupby/hash/groupby.cu:548-622). */
enum {
  MIG_STRATEGY_AUTO        = 0,
  MIG_STRATEGY_GLOBAL      = 1, /* device-wide hash table, global atomics      */
  MIG_STRATEGY_PARTITIONED = 2, /* radix partition into LDS-sized bins          */
  MIG_STRATEGY_PARTITIONED_EXACT = 3 /* the same, general keys compared row by row
                                        while probing (the form the partitioned
                                        path falls back to when a verification
                                        finds different keys with equal hashes) */
};
void mig_set_strategy(int32_t strategy);

#ifdef __cplusplus
}
#endif

#endif /* MIG_GROUPBY_H */
