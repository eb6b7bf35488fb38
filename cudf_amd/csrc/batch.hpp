// Batched group-by: device structures and launchers (batch_kernels.hip),
// driven by batch.cpp.
//
// A batch is a template (value columns per request and the aggregations of
// each) applied to many instances.  Instances may differ in row count, key
// type and value types: every accumulator array has 8-byte elements and each
// instance carries its own accumulator kinds (an int32 sum and a float64 sum
// are both a "sum" field of the template), and every output column has an
// 8-byte-per-row region per instance holding elements of that instance's
// result width.
#pragma once

#include "kernels.hpp"

namespace mig {

constexpr int BATCH_LDS_SLOTS = 1024;  // LDS table of the pre-aggregating path
constexpr int BATCH_LDS_FILL  = 768;   // keys a tile may insert before rows bypass LDS
constexpr int BATCH_ELEM      = 8;     // bytes per accumulator / output element slot
constexpr int BATCH_CHUNK     = 4;     // rows per thread handled together in an LDS tile
constexpr int BATCH_MAX_PARTS = 16;    // hash parts of an LDS-aggregated instance (more: direct updates)

// One instance.  Its table occupies slots [slot_base, slot_base + cap + 2) of
// the concatenated slot arrays (slot_base + cap: key equal to EMPTY_KEY,
// slot_base + cap + 1: null key); its results occupy output rows
// [out_base, out_base + groups) (out_base is a multiple of 64).
struct batch_inst {
  int64_t rows;
  uint64_t cap;
  uint64_t slot_base;
  uint64_t out_base;
  int32_t lds;         // >= 1: pre-aggregate in LDS, the keys split over this many
                       // hash parts (one work item per tile and part); 0: direct
  int32_t pad;
  dcol key;
};

struct batch_params {
  int32_t n_inst;
  int32_t num_vals;
  int32_t num_fields;
  int32_t num_outs;
  int32_t skip_null_keys;   // null_policy EXCLUDE
  int32_t need_count_all;
  int32_t tile_rows;
  int32_t has_m2;           // some field is FK_M2 (second pass over the rows)
  int64_t n_items;          // aggregate work items (instance tiles of rows)
  int64_t n_slot_tiles;     // slot work items (SCAN_TILE slots of one instance)
  uint64_t total_slots;
  uint64_t out_rows;        // rows of the concatenated output regions
  uint32_t* queue;          // [0] next first-pass item, [1] next second-pass item
  uint64_t* items;          // (instance << 40) | (hash part << 32) | first row
  uint64_t* slot_tiles;     // (instance << 40) | first local slot
  uint32_t* tile_offsets;   // per slot tile: groups before it within its instance
  int64_t* inst_tile_begin; // per instance: first slot tile (n_inst + 1 entries)
  batch_inst* inst;
  dcol* vals;               // [n_inst * num_vals]
  int32_t* inst_kinds;      // [n_inst * num_fields] accumulator kind of each field
  int32_t* inst_widths;     // [n_inst * (num_outs + 1)] result element width (entry num_outs: key)
  uint64_t* slot_keys;
  uint32_t* slot_count;     // group row counts, or null
  uint32_t* special;        // [2 * n_inst] occupancy of the two reserved slots
  uint32_t* groups;         // [n_inst] output groups
  int32_t* overflow;        // [n_inst] a table ran out of slots
  uint32_t* valid_counts;   // [n_inst * (num_outs + 1)] set bits of each output mask
                            // (entry num_outs: the key mask)
  int32_t val_field_begin[MAX_VALS + 1];
  field_desc fields[MAX_FIELDS];  // ptr: total_slots elements of BATCH_ELEM bytes;
                                  // kind: FK_M2 or the template's class (unused otherwise)
  out_desc outs[MAX_OUTS];  // data: out_rows elements of BATCH_ELEM bytes; mask or null
  void* out_key;            // out_rows elements of BATCH_ELEM bytes
  uint32_t* out_key_mask;
  int32_t lds_field_off[MAX_FIELDS];  // first pass: LDS offsets of the partials (-1: none)
  int32_t lds_m2_off[MAX_FIELDS];     // second pass: LDS offsets of the M2 partials
  int32_t lds_count_off;              // -1 when no row count is kept
  int32_t lds_bytes;
};

void launch_batch_sample(batch_params const* dB, int32_t n_inst, int32_t per_inst, uint64_t* tables,
                         uint32_t table_slots, uint32_t* counters, hipStream_t s);
void launch_batch_init(batch_params const* dB, batch_params const& hB, int num_cus, hipStream_t s);
void prepare_batch_aggregate(int32_t lds_bytes);
void launch_batch_aggregate(batch_params const* dB, batch_params const& hB, int num_cus, int block,
                            hipStream_t s);
void launch_batch_second_pass(batch_params const* dB, batch_params const& hB, int num_cus, int block,
                              hipStream_t s);
void launch_batch_count(batch_params const* dB, batch_params const& hB, hipStream_t s);
void launch_batch_scan(batch_params const* dB, batch_params const& hB, hipStream_t s);
void launch_batch_outputs(batch_params const* dB, batch_params const& hB, hipStream_t s);
void launch_batch_valid_counts(batch_params const* dB, batch_params const& hB, hipStream_t s);

}  // namespace mig
