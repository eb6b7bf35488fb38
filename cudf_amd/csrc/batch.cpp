// Host driver of the batched group-by: many independent single-key group-bys
// (same requests and aggregations; any row counts, cardinalities, key types
// and value types) sized from one device-side sample pass, laid out in
// concatenated tables and output regions, and executed by persistent
// work-queue kernels plus a few compaction kernels captured in a hipGraph.
//
// Each instance's result is what the single-call path returns for it
// (groupby.cpp): same result types, nullability (a SUM/MIN/MAX/MEAN result
// has a mask iff that instance's values have nulls, VARIANCE/STD always, the
// key result iff its key view has a mask; an empty instance returns columns
// without masks, groupby.cu:150-172) and group set.
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "batch_host.hpp"
#include "groupby.hpp"
#include "host_plan.hpp"

struct mig_batch_plan {
  enum mask_mode : int32_t { MASK_NONE = 0, MASK_IF_NULLS = 1, MASK_ALWAYS = 2 };
  struct out_info {
    int32_t agg_kind;  // MIG_AGG_*
    int32_t vcol;
    int32_t mask;      // mask_mode
  };

  int dev     = 0;
  int num_cus = 256;
  int32_t n_inst = 0, null_policy = 0, num_requests = 0, num_vals = 0;
  mig_batch_options opt{};
  std::vector<mig_column_view> key_views;          // [n_inst]
  std::vector<mig_column_view> val_views;          // [n_inst * num_vals]
  std::vector<uint8_t> val_nulls;                  // [n_inst * num_vals]
  std::vector<int32_t> agg_begin;                  // [num_requests + 1]
  std::vector<int32_t> vcol_of;                    // [num_requests] value column of each request
  std::vector<mig_aggregation> aggs;               // the requests' aggregations, flattened
  std::vector<uint8_t> solo;                       // [n_inst] runs the single-call path
  std::vector<std::unique_ptr<mig_groupby_result>> solo_res;
  std::vector<int32_t> out_of_agg;                 // flattened (request, agg) -> output
  std::vector<out_info> outs;
  std::vector<int32_t> kinds;                      // [n_inst * num_fields]
  std::vector<int32_t> widths;                     // [n_inst * (num_outs + 1)]
  std::vector<double> est;                         // estimated groups per instance
  std::vector<uint64_t> cap;                       // table slots per instance
  std::vector<mig::batch_inst> hinst;
  mig::batch_params P{};
  mig::batch_params* dP = nullptr;
  std::vector<std::pair<void*, size_t>> blocks;  // from the scratch cache
  hipStream_t last_stream    = nullptr;           // stream of the last execute
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph           = nullptr;
  hipGraphExec_t exec        = nullptr;
  uint32_t* pinned           = nullptr;  // groups | overflow | valid counts
  std::vector<uint32_t> groups, valid;
  bool executed = false, synced = false;
  int32_t retries = 0;

  ~mig_batch_plan()
  {
    release_layout();
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    if (pinned) (void)hipHostFree(pinned);
  }
  void release_layout()
  {
    // blocks go back to the cache only once the last execute has finished
    if (executed && !synced && last_stream != nullptr) (void)hipStreamSynchronize(last_stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec  = nullptr;
    graph = nullptr;
    for (auto& b : blocks) mig::scratch_give(dev, b.first, b.second);
    blocks.clear();
    dP = nullptr;
  }
  template <typename T>
  T* alloc(size_t count)
  {
    size_t const bytes = std::max<size_t>(256, (count * sizeof(T) + 255) & ~size_t(255));
    void* p            = mig::scratch_take(dev, bytes);
    blocks.emplace_back(p, bytes);
    return static_cast<T*>(p);
  }
  uint64_t region_rows(int32_t i) const
  {
    mig::batch_inst const& I = hinst[size_t(i)];
    return (std::min<uint64_t>(I.cap + 2, uint64_t(I.rows)) + 63) & ~uint64_t(63);
  }
};

namespace mig {

namespace {

constexpr int32_t BATCH_SAMPLE = 4096;  // sampled rows per instance
// Instances this large with this many estimated groups run the single-call
// path: its radix-partitioned aggregation (LDS atomics per bin) beats the
// per-row device-wide atomics the batched kernels fall back to when an
// instance's groups exceed what LDS pre-aggregation takes.
constexpr int32_t BATCH_SOLO_ROWS   = 1 << 20;
constexpr double BATCH_SOLO_GROUPS  = 65536.0;

int32_t default_i(int32_t v, int32_t d) { return v == 0 ? d : v; }

// Table slots for an estimate of `e` groups over `rows` rows.
uint64_t table_slots(double e, int64_t rows, double load_factor)
{
  uint64_t const want = uint64_t(std::ceil(std::max(e, 1.0) / load_factor));
  uint64_t const top  = next_pow2(uint64_t(2 * rows + 2));
  return std::max<uint64_t>(16, std::min(next_pow2(want), top));
}

void enqueue(mig_batch_plan& p, hipStream_t s)
{
  launch_batch_init(p.dP, p.P, p.num_cus, s);
  launch_batch_aggregate(p.dP, p.P, p.num_cus, p.opt.block_size, s);
  launch_batch_second_pass(p.dP, p.P, p.num_cus, p.opt.block_size, s);
  launch_batch_count(p.dP, p.P, s);
  launch_batch_scan(p.dP, p.P, s);
  launch_batch_outputs(p.dP, p.P, s);
  launch_batch_valid_counts(p.dP, p.P, s);
  MIG_HIP_TRY(hipGetLastError());
}

// Tables, outputs, work lists and the parameter block for the current `cap`
// (uploaded on `stream`, which is synchronised before the graph is captured).
void layout(mig_batch_plan& p, hipStream_t stream)
{
  p.release_layout();
  batch_params& P = p.P;
  int32_t const n = p.n_inst;

  uint64_t total_slots = 0, out_rows = 0;
  int64_t n_items = 0, n_tiles = 0;
  std::vector<int64_t> inst_tile_begin(size_t(n) + 1, 0);
  for (int32_t i = 0; i < n; i++) {
    batch_inst& I = p.hinst[i];
    I.cap         = p.cap[i];
    I.slot_base   = total_slots;
    total_slots += I.cap + 2;
    I.out_base = out_rows;
    out_rows += p.region_rows(i);
    n_items += ((I.rows + P.tile_rows - 1) / P.tile_rows) * std::max(1, I.lds);
    inst_tile_begin[i] = n_tiles;
    n_tiles += int64_t((I.cap + 2 + SCAN_TILE - 1) / SCAN_TILE);
  }
  inst_tile_begin[n] = n_tiles;
  P.total_slots  = total_slots;
  P.out_rows     = out_rows;
  P.n_items      = n_items;
  P.n_slot_tiles = n_tiles;

  std::vector<uint64_t> items;
  items.reserve(size_t(n_items));
  std::vector<uint64_t> tiles;
  tiles.reserve(size_t(n_tiles));
  for (int32_t i = 0; i < n; i++) {
    // the parts of a tile are consecutive items, so the workgroups that read
    // the tile for its parts run at about the same time (cache reuse)
    for (int64_t r = 0; r < p.hinst[i].rows; r += P.tile_rows)
      for (int32_t part = 0; part < std::max(1, p.hinst[i].lds); part++)
        items.push_back((uint64_t(i) << 40) | (uint64_t(part) << 32) | uint64_t(r));
    for (uint64_t l = 0; l < p.hinst[i].cap + 2; l += SCAN_TILE) tiles.push_back((uint64_t(i) << 40) | l);
  }

  P.queue           = p.alloc<uint32_t>(2);
  P.items           = p.alloc<uint64_t>(std::max<size_t>(1, items.size()));
  P.slot_tiles      = p.alloc<uint64_t>(tiles.size());
  P.tile_offsets    = p.alloc<uint32_t>(tiles.size());
  P.inst_tile_begin = p.alloc<int64_t>(size_t(n) + 1);
  P.inst            = p.alloc<batch_inst>(size_t(n));
  P.vals            = p.alloc<dcol>(std::max<size_t>(1, size_t(n) * size_t(p.num_vals)));
  P.inst_kinds      = p.alloc<int32_t>(std::max<size_t>(1, p.kinds.size()));
  P.inst_widths     = p.alloc<int32_t>(p.widths.size());
  P.slot_keys       = p.alloc<uint64_t>(total_slots);
  P.slot_count      = P.need_count_all ? p.alloc<uint32_t>(total_slots) : nullptr;
  P.special         = p.alloc<uint32_t>(2 * size_t(n));
  P.groups          = p.alloc<uint32_t>(size_t(n));
  P.overflow        = p.alloc<int32_t>(size_t(n));
  P.valid_counts    = p.alloc<uint32_t>(size_t(n) * size_t(P.num_outs + 1));
  for (int f = 0; f < P.num_fields; f++) P.fields[f].ptr = p.alloc<uint64_t>(total_slots);
  size_t const mask_words = size_t((out_rows + 31) / 32);
  P.out_key      = p.alloc<uint64_t>(size_t(out_rows));
  P.out_key_mask = p.alloc<uint32_t>(mask_words);
  for (int o = 0; o < P.num_outs; o++) {
    P.outs[o].data = p.alloc<uint64_t>(size_t(out_rows));
    P.outs[o].mask = p.outs[o].mask != mig_batch_plan::MASK_NONE ? p.alloc<uint32_t>(mask_words) : nullptr;
  }

  std::vector<dcol> vals(size_t(n) * size_t(p.num_vals));
  for (int32_t i = 0; i < n; i++)
    for (int32_t v = 0; v < p.num_vals; v++) {
      size_t const k = size_t(i) * size_t(p.num_vals) + size_t(v);
      vals[k]        = make_dcol(p.val_views[k], p.val_nulls[k] != 0);
    }

  auto up = [&](void* dst, void const* src, size_t bytes) {
    if (bytes) MIG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
  };
  up(P.items, items.data(), items.size() * sizeof(uint64_t));
  up(P.slot_tiles, tiles.data(), tiles.size() * sizeof(uint64_t));
  up(P.inst_tile_begin, inst_tile_begin.data(), inst_tile_begin.size() * sizeof(int64_t));
  up(P.inst, p.hinst.data(), size_t(n) * sizeof(batch_inst));
  up(P.vals, vals.data(), vals.size() * sizeof(dcol));
  up(P.inst_kinds, p.kinds.data(), p.kinds.size() * sizeof(int32_t));
  up(P.inst_widths, p.widths.data(), p.widths.size() * sizeof(int32_t));
  p.dP = p.alloc<batch_params>(1);
  up(p.dP, &P, sizeof(batch_params));

  MIG_HIP_TRY(hipStreamSynchronize(stream));
  prepare_batch_aggregate(P.lds_bytes);
  if (p.opt.use_graph) {
    if (!p.capture_stream) MIG_HIP_TRY(hipStreamCreateWithFlags(&p.capture_stream, hipStreamNonBlocking));
    // relaxed: other threads (single calls, other plans) keep working while
    // this thread captures
    MIG_HIP_TRY(hipStreamBeginCapture(p.capture_stream, hipStreamCaptureModeRelaxed));
    try {
      enqueue(p, p.capture_stream);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(p.capture_stream, &g);
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();  // leave no sticky capture error to the thread's next call
      throw;
    }
    hipError_t const e = hipStreamEndCapture(p.capture_stream, &p.graph);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(MIG_ERROR_HIP, std::string("graph capture failed: ") + hipGetErrorName(e));
    }
    MIG_HIP_TRY(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
  }
}

// Per-instance group estimates from a strided sample of BATCH_SAMPLE rows
// (all rows of smaller instances, which makes their count exact up to
// 32-bit hash collisions).
void estimate(mig_batch_plan& p, hipStream_t stream)
{
  int32_t const n = p.n_inst;
  mig_batch_plan tmp;  // owns the sample's device memory
  tmp.dev = p.dev;
  batch_params& P = p.P;
  P.inst          = tmp.alloc<batch_inst>(size_t(n));
  MIG_HIP_TRY(hipMemcpyAsync(P.inst, p.hinst.data(), size_t(n) * sizeof(batch_inst), hipMemcpyHostToDevice,
                             stream));
  auto* dP = tmp.alloc<batch_params>(1);
  MIG_HIP_TRY(hipMemcpyAsync(dP, &P, sizeof(batch_params), hipMemcpyHostToDevice, stream));
  uint32_t const tslots = uint32_t(next_pow2(2 * BATCH_SAMPLE));
  auto* tables          = tmp.alloc<uint64_t>(size_t(n) * tslots);
  auto* counters        = tmp.alloc<uint32_t>(2 * size_t(n));
  MIG_HIP_TRY(hipMemsetAsync(tables, 0, size_t(n) * tslots * sizeof(uint64_t), stream));
  MIG_HIP_TRY(hipMemsetAsync(counters, 0, 2 * size_t(n) * sizeof(uint32_t), stream));
  launch_batch_sample(dP, n, BATCH_SAMPLE, tables, tslots, counters, stream);
  MIG_HIP_TRY(hipGetLastError());
  std::vector<uint32_t> h(2 * size_t(n));
  MIG_HIP_TRY(hipMemcpyAsync(h.data(), counters, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  MIG_HIP_TRY(hipStreamSynchronize(stream));
  P.inst = nullptr;
  p.est.assign(size_t(n), 1.0);
  for (int32_t i = 0; i < n; i++) {
    int64_t const rows   = p.hinst[i].rows;
    int64_t const ns     = std::min<int64_t>(rows, BATCH_SAMPLE);
    double const sampled = double(h[2 * i]), d = double(h[2 * i + 1]);
    if (ns == rows) p.est[i] = d;
    else if (sampled > 0) {
      double const pop = double(rows) * (sampled / double(ns));
      p.est[i]         = estimate_distinct(sampled, d, std::max(pop, d));
    } else p.est[i] = 1.0;
  }
}

bool fixed_width(int32_t type_id)
{
  int32_t const st = storage_of(type_id);
  return st != ST_NONE && st != ST_STRING;
}

}  // namespace

mig_batch_plan* batch_plan_create(mig_batch_instance const* instances, int32_t num_instances,
                                  int32_t num_keys, int32_t null_policy,
                                  mig_aggregation const* aggregations, int32_t const* num_aggregations,
                                  int32_t num_requests, mig_batch_options const* options,
                                  hipStream_t stream)
{
  MIG_EXPECTS(num_instances > 0, "a batch needs at least one instance");
  MIG_EXPECTS(num_instances < (1 << 24), "too many instances in one batch");
  MIG_EXPECTS(instances != nullptr, "null instance array");
  MIG_EXPECTS(num_keys == 1, "the batched group-by takes exactly one key column");
  MIG_EXPECTS(null_policy == MIG_NULL_EXCLUDE || null_policy == MIG_NULL_INCLUDE, "invalid null policy");
  MIG_EXPECTS(num_requests >= 0 && num_requests <= MAX_VALS, "number of requests out of range");
  MIG_EXPECTS(num_requests == 0 || (aggregations && num_aggregations), "null aggregation arrays");

  (void)hipGetLastError();  // a launch error reported below belongs to this call
  auto plan         = std::make_unique<mig_batch_plan>();
  mig_batch_plan& p = *plan;
  MIG_HIP_TRY(hipGetDevice(&p.dev));
  MIG_HIP_TRY(hipDeviceGetAttribute(&p.num_cus, hipDeviceAttributeMultiprocessorCount, p.dev));
  p.n_inst       = num_instances;
  p.null_policy  = null_policy;
  p.num_requests = num_requests;
  if (options) p.opt = *options;
  if (!(p.opt.load_factor > 0.0)) p.opt.load_factor = 0.5;
  MIG_EXPECTS(p.opt.load_factor <= 0.9, "load_factor must be in (0, 0.9]");
  p.opt.lds_max_groups = default_i(p.opt.lds_max_groups, 512);
  MIG_EXPECTS(p.opt.lds_max_groups <= BATCH_LDS_FILL, "lds_max_groups exceeds the LDS table");
  p.opt.block_size = default_i(p.opt.block_size, 256);
  MIG_EXPECTS(p.opt.block_size == 256 || p.opt.block_size == 512, "block_size must be 256 or 512");
  p.opt.tile_rows = default_i(p.opt.tile_rows, 4096);
  MIG_EXPECTS(p.opt.tile_rows >= 256 && p.opt.tile_rows <= (1 << 24), "tile_rows out of range");
  p.opt.use_graph = p.opt.use_graph < 0 ? 0 : 1;

  // ---- keys: one fixed-width column per instance ------------------------------
  p.key_views.resize(size_t(num_instances));
  p.hinst.assign(size_t(num_instances), batch_inst{});
  for (int32_t i = 0; i < num_instances; i++) {
    MIG_EXPECTS(instances[i].keys != nullptr, "null key array of an instance");
    mig_column_view const& k = instances[i].keys[0];
    if (!fixed_width(k.type_id) || is_decimal_type(k.type_id))
      fail(MIG_ERROR_DATA_TYPE, "the batched group-by takes one fixed-width, non-fixed-point key column per instance");
    MIG_EXPECTS(k.size >= 0 && k.offset >= 0, "negative column size or offset");
    p.key_views[i]  = k;
    p.hinst[i].rows = k.size;
    p.hinst[i].key  = make_dcol(k, k.null_mask != nullptr);
  }

  // ---- requests: one value column per distinct request column -------------
  p.agg_begin.assign(size_t(num_requests) + 1, 0);
  for (int32_t r = 0; r < num_requests; r++) {
    MIG_EXPECTS(num_aggregations[r] >= 0, "negative number of aggregations");
    p.agg_begin[r + 1] = p.agg_begin[r] + num_aggregations[r];
  }
  for (int32_t i = 0; i < num_instances; i++)
    MIG_EXPECTS(num_requests == 0 || instances[i].values != nullptr, "null values array of an instance");
  // request r reads value column vcol_of[r]; columns are shared when every
  // instance passes the same view for two requests
  std::vector<int32_t> vcol_of(size_t(num_requests), -1), req_of_vcol;
  for (int32_t r = 0; r < num_requests; r++) {
    for (size_t v = 0; v < req_of_vcol.size() && vcol_of[r] < 0; v++) {
      bool same = true;
      for (int32_t i = 0; i < num_instances && same; i++)
        same = same_view(instances[i].values[r], instances[i].values[req_of_vcol[v]]);
      if (same) vcol_of[r] = int32_t(v);
    }
    if (vcol_of[r] < 0) {
      vcol_of[r] = int32_t(req_of_vcol.size());
      req_of_vcol.push_back(r);
    }
  }
  p.num_vals = int32_t(req_of_vcol.size());
  p.vcol_of  = vcol_of;
  p.aggs.assign(aggregations, aggregations + p.agg_begin[size_t(num_requests)]);
  p.val_views.resize(size_t(num_instances) * size_t(p.num_vals));
  for (int32_t v = 0; v < p.num_vals; v++)
    for (int32_t i = 0; i < num_instances; i++) {
      mig_column_view const& c = instances[i].values[req_of_vcol[v]];
      if (!fixed_width(c.type_id) || is_decimal_type(c.type_id))
        fail(MIG_ERROR_DATA_TYPE, "the batched group-by takes fixed-width, non-fixed-point value columns");
      MIG_EXPECTS(c.size == p.key_views[i].size, "Size mismatch between request values and groupby keys.");
      MIG_EXPECTS(c.offset >= 0, "negative column offset");
      p.val_views[size_t(i) * p.num_vals + v] = c;
    }
  for (int32_t r = 0; r < num_requests; r++)
    for (int32_t a = 0; a < num_aggregations[r]; a++) {
      int32_t const kind = aggregations[p.agg_begin[r] + a].kind;
      if (kind == MIG_AGG_ARGMIN || kind == MIG_AGG_ARGMAX || kind == MIG_AGG_PRODUCT || kind == MIG_AGG_NUNIQUE ||
          kind == MIG_AGG_NTH_ELEMENT)
        fail(MIG_ERROR_LOGIC, std::string("aggregation ") + kind_name(kind) + " is not supported by the batched form");
      for (int32_t i = 0; i < num_instances; i++)
        check_supported(p.val_views[size_t(i) * p.num_vals + vcol_of[r]].type_id, kind);
    }

  // null counts of the value views (counted on the device where unknown)
  p.val_nulls.assign(p.val_views.size(), 0);
  {
    std::vector<size_t> unknown;
    for (size_t k = 0; k < p.val_views.size(); k++) {
      mig_column_view const& c = p.val_views[k];
      if (c.null_mask == nullptr || c.size == 0) continue;
      if (c.null_count >= 0) p.val_nulls[k] = c.null_count > 0;
      else unknown.push_back(k);
    }
    if (!unknown.empty()) {
      mig_batch_plan tmp;
      tmp.dev = p.dev;
      auto* d = tmp.alloc<uint32_t>(unknown.size());
      MIG_HIP_TRY(hipMemsetAsync(d, 0, unknown.size() * sizeof(uint32_t), stream));
      for (size_t u = 0; u < unknown.size(); u++) {
        mig_column_view const& c = p.val_views[unknown[u]];
        dcol dc{};
        dc.mask        = c.null_mask;
        dc.mask_offset = c.offset;
        launch_count_nulls(dc, c.size, d + u, stream);
      }
      std::vector<uint32_t> h(unknown.size());
      MIG_HIP_TRY(hipMemcpyAsync(h.data(), d, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
      MIG_HIP_TRY(hipStreamSynchronize(stream));
      for (size_t u = 0; u < unknown.size(); u++)
        p.val_nulls[unknown[u]] = int64_t(h[u]) < int64_t(p.val_views[unknown[u]].size);
    }
  }
  std::vector<uint8_t> any_nulls(size_t(p.num_vals), 0);
  for (int32_t i = 0; i < num_instances; i++)
    for (int32_t v = 0; v < p.num_vals; v++) any_nulls[v] |= p.val_nulls[size_t(i) * p.num_vals + v];

  // ---- accumulator template: fields grouped by value column ------------------
  batch_params& P = p.P;
  std::memset(&P, 0, sizeof(batch_params));
  P.n_inst         = num_instances;
  P.skip_null_keys = null_policy == MIG_NULL_EXCLUDE;
  P.tile_rows      = p.opt.tile_rows;
  enum field_class : int32_t { FC_SUM, FC_SUMSQ, FC_MIN, FC_MAX, FC_COUNT, FC_M2 };
  struct vplan {
    bool sum = false, sumsq = false, mn = false, mx = false, count = false, m2 = false;
    int32_t f_sum = -1, f_sumsq = -1, f_min = -1, f_max = -1, f_count = -1, f_m2 = -1;
  };
  std::vector<vplan> vp(size_t(p.num_vals));
  for (int32_t r = 0; r < num_requests; r++) {
    vplan& q      = vp[vcol_of[r]];
    bool const nn = any_nulls[vcol_of[r]] != 0;
    for (int32_t a = 0; a < num_aggregations[r]; a++) switch (aggregations[p.agg_begin[r] + a].kind) {
        case MIG_AGG_SUM: q.sum = true; q.count |= nn; break;
        case MIG_AGG_SUM_OF_SQUARES: q.sumsq = true; q.count |= nn; break;
        case MIG_AGG_MIN: q.mn = true; q.count |= nn; break;
        case MIG_AGG_MAX: q.mx = true; q.count |= nn; break;
        case MIG_AGG_COUNT_VALID: q.count = true; break;
        case MIG_AGG_COUNT_ALL: P.need_count_all = 1; break;
        case MIG_AGG_MEAN: q.sum = q.count = true; break;
        case MIG_AGG_VARIANCE:
        case MIG_AGG_STD: q.sum = q.count = q.m2 = true; break;
        default: break;
      }
  }
  std::vector<int32_t> fclass;  // class of each field
  int nf = 0;
  auto add_field = [&](int32_t v, int32_t cls) {
    MIG_EXPECTS(nf < MAX_FIELDS, "too many accumulators");
    field_desc& F = P.fields[nf];
    F.kind        = cls == FC_M2 ? FK_M2 : FK_SUM_I64;  // template marker; per-instance kinds below
    F.vcol        = v;
    F.width       = BATCH_ELEM;
    F.aux_sum     = -1;
    F.aux_count   = -1;
    fclass.push_back(cls);
    return nf++;
  };
  int32_t const COUNT_FROM_ROWS = -2;
  for (int32_t v = 0; v < p.num_vals; v++) {
    vplan& q             = vp[v];
    P.val_field_begin[v] = nf;
    if (q.sum) q.f_sum = add_field(v, FC_SUM);
    if (q.sumsq) q.f_sumsq = add_field(v, FC_SUMSQ);
    if (q.mn) q.f_min = add_field(v, FC_MIN);
    if (q.mx) q.f_max = add_field(v, FC_MAX);
    if (q.count) {
      if (any_nulls[v]) q.f_count = add_field(v, FC_COUNT);
      else {
        q.f_count        = COUNT_FROM_ROWS;
        P.need_count_all = 1;
      }
    }
    if (q.m2) {
      q.f_m2                     = add_field(v, FC_M2);
      P.fields[q.f_m2].aux_sum   = q.f_sum;
      P.fields[q.f_m2].aux_count = q.f_count;
      P.has_m2                   = 1;
    }
  }
  P.val_field_begin[p.num_vals] = nf;
  P.num_vals                    = p.num_vals;
  P.num_fields                  = nf;
  // per-instance accumulator kinds
  p.kinds.assign(size_t(num_instances) * size_t(nf), 0);
  for (int32_t i = 0; i < num_instances; i++)
    for (int f = 0; f < nf; f++) {
      int32_t const st = storage_of(p.val_views[size_t(i) * p.num_vals + P.fields[f].vcol].type_id);
      int32_t k        = FK_M2;
      switch (fclass[f]) {
        case FC_SUM: k = sum_field_kind(st); break;
        case FC_SUMSQ: k = sumsq_field_kind(st); break;
        case FC_MIN: k = minmax_field_kind(st, true); break;
        case FC_MAX: k = minmax_field_kind(st, false); break;
        case FC_COUNT: k = FK_COUNT; break;
        default: break;
      }
      p.kinds[size_t(i) * nf + f] = k;
    }

  // ---- output programs -----------------------------------------------------------
  p.out_of_agg.assign(size_t(p.agg_begin[num_requests]), -1);
  int no = 0;
  for (int32_t r = 0; r < num_requests; r++) {
    int32_t const v = vcol_of[r];
    vplan const& q  = vp[v];
    for (int32_t a = 0; a < num_aggregations[r]; a++) {
      int32_t const kind = aggregations[p.agg_begin[r] + a].kind;
      MIG_EXPECTS(no < MAX_OUTS, "too many aggregation results");
      out_desc& O   = P.outs[no];
      O             = out_desc{};
      O.count_field = -1;
      O.ddof        = aggregations[p.agg_begin[r] + a].ddof;
      int32_t mask  = mig_batch_plan::MASK_NONE;
      switch (kind) {
        case MIG_AGG_SUM:
        case MIG_AGG_SUM_OF_SQUARES:
          O.field = kind == MIG_AGG_SUM ? q.f_sum : q.f_sumsq;
          O.kind  = OK_COPY_I64;  // copies the instance's result width
          if (any_nulls[v]) { O.count_field = q.f_count; mask = mig_batch_plan::MASK_IF_NULLS; }
          break;
        case MIG_AGG_MIN:
        case MIG_AGG_MAX:
          O.field = kind == MIG_AGG_MIN ? q.f_min : q.f_max;
          O.kind  = OK_MINMAX;
          if (any_nulls[v]) { O.count_field = q.f_count; mask = mig_batch_plan::MASK_IF_NULLS; }
          break;
        case MIG_AGG_COUNT_VALID:
          if (q.f_count >= 0) { O.kind = OK_COUNT; O.field = q.f_count; }
          else O.kind = OK_COUNT_ALL;
          break;
        case MIG_AGG_COUNT_ALL: O.kind = OK_COUNT_ALL; break;
        case MIG_AGG_MEAN:
          O.kind        = OK_MEAN;
          O.field       = q.f_sum;
          O.count_field = q.f_count;
          if (any_nulls[v]) mask = mig_batch_plan::MASK_IF_NULLS;
          break;
        case MIG_AGG_VARIANCE:
        case MIG_AGG_STD:
          O.kind        = kind == MIG_AGG_VARIANCE ? OK_VAR : OK_STD;
          O.field       = q.f_m2;
          O.count_field = q.f_count;
          mask          = mig_batch_plan::MASK_ALWAYS;
          break;
        default: fail(MIG_ERROR_LOGIC, "unreachable aggregation kind");
      }
      p.outs.push_back({kind, v, mask});
      p.out_of_agg[size_t(p.agg_begin[r] + a)] = no++;
    }
  }
  P.num_outs = no;
  p.widths.assign(size_t(num_instances) * size_t(no + 1), 0);
  for (int32_t i = 0; i < num_instances; i++) {
    for (int o = 0; o < no; o++) {
      int32_t const src = p.val_views[size_t(i) * p.num_vals + p.outs[o].vcol].type_id;
      p.widths[size_t(i) * (no + 1) + o] = type_width(target_type_of(src, p.outs[o].agg_kind));
    }
    p.widths[size_t(i) * (no + 1) + no] = type_width(p.key_views[i].type_id);
  }

  // ---- LDS layout of the pre-aggregating tiles -----------------------------------
  // first pass: key table, one partial array per non-M2 field, row counts;
  // second pass (within the same bytes): slot table, one array per M2 field
  int32_t off = BATCH_LDS_SLOTS * 8, m2_off = BATCH_LDS_SLOTS * 8;
  for (int f = 0; f < nf; f++) {
    if (P.fields[f].kind == FK_M2) {
      P.lds_field_off[f] = -1;
      P.lds_m2_off[f]    = m2_off;
      m2_off += BATCH_LDS_SLOTS * 8;
    } else {
      P.lds_field_off[f] = off;
      P.lds_m2_off[f]    = -1;
      off += (BATCH_LDS_SLOTS + 2) * 8;
    }
  }
  P.lds_count_off = -1;
  if (P.need_count_all) {
    P.lds_count_off = off;
    off += ((BATCH_LDS_SLOTS + 2) * 4 + 7) & ~7;
  }
  P.lds_bytes = std::max(off, m2_off);
  // (the LDS path holds its field plan one field per lane: at most 64 fields)
  if (P.lds_bytes > 160 * 1024 || p.opt.lds_max_groups < 0 || nf > 64) {
    p.opt.lds_max_groups = -1;  // off, or does not fit: direct updates only
    P.lds_bytes          = 0;
  }

  // ---- estimates, tables, layout --------------------------------------------------
  estimate(p, stream);
  int32_t const solo_rows = p.opt.partition_min_rows == 0 ? BATCH_SOLO_ROWS : p.opt.partition_min_rows;
  p.solo.assign(size_t(num_instances), 0);
  p.solo_res.resize(size_t(num_instances));
  for (int32_t i = 0; i < num_instances; i++)
    if (p.opt.partition_min_rows >= 0 && p.hinst[i].rows >= solo_rows && p.est[i] >= BATCH_SOLO_GROUPS) {
      p.solo[i]       = 1;
      p.hinst[i].rows = 0;  // no rows in the batched kernels
    }
  p.cap.resize(size_t(num_instances));
  for (int32_t i = 0; i < num_instances; i++) {
    p.cap[i]       = table_slots(p.est[i], p.hinst[i].rows, p.opt.load_factor);
    // LDS pre-aggregation with the keys split over as many hash parts as
    // needed for each part to fit the LDS table (each part re-reads the rows,
    // mostly from cache); beyond BATCH_MAX_PARTS the table is updated directly
    int32_t parts = 0;
    if (p.opt.lds_max_groups > 0) {
      double const need = std::ceil(std::max(p.est[i], 1.0) / double(p.opt.lds_max_groups));
      if (need <= double(BATCH_MAX_PARTS)) parts = int32_t(need);
    }
    p.hinst[i].lds = parts;
  }
  layout(p, stream);

  size_t const words = size_t(num_instances) * (2 + size_t(no + 1));
  MIG_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p.pinned), words * sizeof(uint32_t), hipHostMallocDefault));
  return plan.release();
}

namespace {

void launch_batched(mig_batch_plan& p, hipStream_t stream)
{
  if (p.exec) MIG_HIP_TRY(hipGraphLaunch(p.exec, stream));
  else enqueue(p, stream);
}

// The single-call instances, one after the other on `stream` (each call
// returns once its group count is known).
void run_solo(mig_batch_plan& p, hipStream_t stream)
{
  std::vector<mig_aggregation_request> reqs(size_t(p.num_requests));
  for (int32_t i = 0; i < p.n_inst; i++) {
    if (!p.solo[size_t(i)]) continue;
    p.solo_res[size_t(i)].reset();
    for (int32_t r = 0; r < p.num_requests; r++) {
      reqs[r].values           = p.val_views[size_t(i) * size_t(p.num_vals) + size_t(p.vcol_of[r])];
      reqs[r].aggregations     = p.aggs.data() + p.agg_begin[r];
      reqs[r].num_aggregations = p.agg_begin[r + 1] - p.agg_begin[r];
    }
    p.solo_res[size_t(i)] =
      groupby_aggregate(&p.key_views[size_t(i)], 1, p.null_policy, reqs.data(), p.num_requests, stream, nullptr);
  }
}

}  // namespace

void batch_execute(mig_batch_plan* plan, hipStream_t stream)
{
  func_range range("groupby: batched plan");
  MIG_EXPECTS(plan != nullptr, "null batch plan");
  mig_batch_plan& p = *plan;
  launch_batched(p, stream);
  p.executed    = true;
  p.synced      = false;
  p.last_stream = stream;
  run_solo(p, stream);
}

void batch_synchronize(mig_batch_plan* plan, hipStream_t stream)
{
  MIG_EXPECTS(plan != nullptr, "null batch plan");
  mig_batch_plan& p = *plan;
  MIG_EXPECTS(p.executed, "mig_batch_synchronize before mig_batch_execute");
  size_t const n  = size_t(p.n_inst);
  size_t const nv = n * size_t(p.P.num_outs + 1);
  for (int attempt = 0;; attempt++) {
    uint32_t* h = p.pinned;
    MIG_HIP_TRY(hipMemcpyAsync(h, p.P.groups, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    MIG_HIP_TRY(hipMemcpyAsync(h + n, p.P.overflow, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    MIG_HIP_TRY(hipMemcpyAsync(h + 2 * n, p.P.valid_counts, nv * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    MIG_HIP_TRY(hipStreamSynchronize(stream));
    bool grew = false;
    for (size_t i = 0; i < n; i++) {
      if (h[n + i] == 0) continue;
      // a table ran out of slots (the estimate was low): grow it and run again
      uint64_t const top = next_pow2(uint64_t(2 * p.hinst[i].rows + 2));
      p.cap[i]           = std::min(top, p.cap[i] * 4);
      p.hinst[i].lds     = 0;
      grew               = true;
    }
    if (!grew) break;
    MIG_EXPECTS(attempt < 16, "batched group-by: tables kept overflowing");
    p.retries++;
    layout(p, stream);
    launch_batched(p, stream);
  }
  p.groups.assign(p.pinned, p.pinned + n);
  p.valid.assign(p.pinned + 2 * n, p.pinned + 2 * n + nv);
  p.synced = true;
}

int32_t batch_num_groups(mig_batch_plan const* plan, int32_t instance)
{
  if (!plan || !plan->synced || instance < 0 || instance >= plan->n_inst) return -1;
  if (plan->solo[size_t(instance)]) return plan->solo_res[size_t(instance)]->num_groups;
  return int32_t(plan->groups[size_t(instance)]);
}

mig_column batch_result_key(mig_batch_plan const* plan, int32_t instance)
{
  MIG_EXPECTS(plan != nullptr, "null batch plan");
  mig_batch_plan const& p = *plan;
  MIG_EXPECTS(p.synced, "results read before mig_batch_synchronize");
  MIG_EXPECTS(instance >= 0 && instance < p.n_inst, "instance index out of range");
  if (p.solo[size_t(instance)]) return p.solo_res[size_t(instance)]->keys[0].col;
  batch_inst const& I = p.hinst[size_t(instance)];
  uint32_t const G    = p.groups[size_t(instance)];
  int32_t const tid   = p.key_views[size_t(instance)].type_id;
  uint64_t const capr = p.region_rows(instance);
  mig_column c{};
  c.type_id    = tid;
  c.size       = int32_t(G);
  c.data       = static_cast<uint8_t*>(p.P.out_key) + I.out_base * BATCH_ELEM;
  c.data_bytes = int64_t(capr * uint64_t(type_width(tid)));
  // an empty instance returns empty columns without masks, as the single call does
  if (p.key_views[size_t(instance)].null_mask != nullptr && I.rows > 0) {
    c.null_mask  = p.P.out_key_mask + I.out_base / 32;
    c.mask_bytes = int64_t(capr / 8);
    c.null_count = int32_t(G - p.valid[size_t(instance) * size_t(p.P.num_outs + 1) + size_t(p.P.num_outs)]);
  }
  return c;
}

mig_column batch_result_column(mig_batch_plan const* plan, int32_t instance, int32_t request, int32_t agg)
{
  MIG_EXPECTS(plan != nullptr, "null batch plan");
  mig_batch_plan const& p = *plan;
  MIG_EXPECTS(p.synced, "results read before mig_batch_synchronize");
  MIG_EXPECTS(instance >= 0 && instance < p.n_inst, "instance index out of range");
  MIG_EXPECTS(request >= 0 && request < p.num_requests, "request index out of range");
  MIG_EXPECTS(agg >= 0 && agg < p.agg_begin[size_t(request) + 1] - p.agg_begin[size_t(request)],
              "aggregation index out of range");
  if (p.solo[size_t(instance)]) return p.solo_res[size_t(instance)]->results[size_t(request)][size_t(agg)].col;
  int32_t const o     = p.out_of_agg[size_t(p.agg_begin[size_t(request)] + agg)];
  auto const& oi      = p.outs[size_t(o)];
  batch_inst const& I = p.hinst[size_t(instance)];
  uint32_t const G    = p.groups[size_t(instance)];
  int32_t const src   = p.val_views[size_t(instance) * size_t(p.num_vals) + size_t(oi.vcol)].type_id;
  int32_t const tid   = target_type_of(src, oi.agg_kind);
  uint64_t const capr = p.region_rows(instance);
  mig_column c{};
  c.type_id    = tid;
  c.size       = int32_t(G);
  c.data       = static_cast<uint8_t*>(p.P.outs[o].data) + I.out_base * BATCH_ELEM;
  c.data_bytes = int64_t(capr * uint64_t(type_width(tid)));
  bool const nulls  = p.val_nulls[size_t(instance) * size_t(p.num_vals) + size_t(oi.vcol)] != 0;
  bool const masked = I.rows > 0 && (oi.mask == mig_batch_plan::MASK_ALWAYS ||
                                     (oi.mask == mig_batch_plan::MASK_IF_NULLS && nulls));
  if (masked) {
    c.null_mask  = p.P.outs[o].mask + I.out_base / 32;
    c.mask_bytes = int64_t(capr / 8);
    c.null_count = int32_t(G - p.valid[size_t(instance) * size_t(p.P.num_outs + 1) + size_t(o)]);
  }
  return c;
}

void batch_plan_free(mig_batch_plan* plan) { delete plan; }

}  // namespace mig
