// Sort-based group-by: MEDIAN and QUANTILE, and NUNIQUE / NTH_ELEMENT of
// keys that are not pre-sorted (the reference's sort path,
// cpp/src/groupby/groupby.cu:66-84 -> sort/aggregate.cpp:798).
//
// The reference sorts the key rows to find the groups (sort/sort_helper.cu:
// 92-104) and then, per aggregation, sorts or groups the values
// (sort/aggregate.cpp:323-410).  Here:
//   1. every row gets a dense group number from the device-wide hash table
//      (or, for pre-sorted keys, from its run), dropped rows number G;
//   2. one stable LSD counting sort orders the rows by group number
//      (rows keep their input order inside a group, as NTH_ELEMENT needs);
//   3. the value columns of the other aggregations are gathered in that
//      order and aggregated by the runs path (groups come out in group-number
//      order);
//   4. per value column with MEDIAN / QUANTILE, a second stable sort orders
//      the rows by (group, null, value) and k_group_quantiles reads every
//      group's segment;
//   5. the key columns are gathered from each group's first row.
// Output row order is unspecified, as for every group-by
// (cpp/include/cudf/groupby.hpp:155-157).
#include <algorithm>
#include <cstring>

#include "groupby.hpp"
#include "host_plan.hpp"
#include "kernels.hpp"
#include "partitioning.hpp"

namespace mig {

namespace {

// Device scratch of the call, back to the per-device cache at the end (after
// the stream has drained).
struct tmp_scratch {
  hipStream_t stream;
  int dev = 0;
  std::vector<std::pair<void*, size_t>> blocks;
  explicit tmp_scratch(hipStream_t s) : stream(s) { MIG_HIP_TRY(hipGetDevice(&dev)); }
  template <typename T>
  T* get(size_t count)
  {
    size_t const bytes = std::max<size_t>(256, (count * sizeof(T) + 255) & ~size_t(255));
    void* p            = scratch_take(dev, bytes);
    blocks.emplace_back(p, bytes);
    return static_cast<T*>(p);
  }
  ~tmp_scratch()
  {
    (void)hipStreamSynchronize(stream);
    for (auto& b : blocks) scratch_give(dev, b.first, b.second);
  }
};

int bits_for(uint64_t v)  // bits to represent every value in [0, v]
{
  int b = 1;
  while (b < 32 && (uint64_t(1) << b) <= v) b++;
  return b;
}

bool is_sort_only(int32_t k) { return k == MIG_AGG_MEDIAN || k == MIG_AGG_QUANTILE; }

size_t mask_words_bytes(int64_t n) { return std::max<size_t>(8, size_t((n + 63) / 64) * 8); }

}  // namespace

std::unique_ptr<mig_groupby_result> sort_groupby(mig_column_view const* keys, int32_t num_keys, int32_t null_policy,
                                                 mig_aggregation_request const* requests, int32_t num_requests,
                                                 hipStream_t stream, mig_memory_resource const* mr, bool keys_sorted)
{
  func_range range("groupby: sort-based aggregations");
  int64_t const N = keys[0].size;
  MIG_EXPECTS(N > 0 && N < (int64_t(1) << 31), "sort-based group-by: rows out of range");
  auto res = std::make_unique<mig_groupby_result>();
  res->alloc.stream = stream;
  if (mr) {
    res->alloc.mr     = *mr;
    res->alloc.has_mr = true;
  }
  device_allocator& A = res->alloc;
  tmp_scratch S(stream);
  double* stats = thread_stats();

  // ---- 1. group numbers ------------------------------------------------------
  auto P_ptr = std::make_unique<gb_params>();
  gb_params& P = *P_ptr;
  std::memset(&P, 0, sizeof(gb_params));
  P.num_rows = N;
  P.num_keys = num_keys;
  bool any_key_nulls = false;
  for (int i = 0; i < num_keys; i++) {
    bool const has_nulls = null_count_of(keys[i], stream) > 0;
    any_key_nulls |= has_nulls;
    P.keys[i] = make_dcol(keys[i], has_nulls);
  }
  P.skip_null_keys       = (null_policy == MIG_NULL_EXCLUDE) && any_key_nulls;
  bool const inline_keys = num_keys == 1 && storage_of(keys[0].type_id) != ST_STRING;
  P.key_mode             = inline_keys ? KM_INLINE : KM_GENERAL;
  P.key_storage          = inline_keys ? storage_of(keys[0].type_id) : ST_NONE;
  gb_params* dP          = S.get<gb_params>(1);
  auto upload = [&]() {
    MIG_HIP_TRY(hipMemcpyAsync(dP, &P, sizeof(gb_params), hipMemcpyHostToDevice, stream));
    MIG_HIP_TRY(hipStreamSynchronize(stream));  // P lives on the host stack frame
  };
  uint32_t* gid = S.get<uint32_t>(N);
  int64_t G     = 0;
  if (keys_sorted) {
    // groups are the runs of equal adjacent keys (sort_helper.cu:55-72)
    P.key_mode            = KM_GENERAL;
    int64_t const ntiles  = sorted_tiles(N);
    uint32_t* tile_counts = S.get<uint32_t>(ntiles + 1);
    upload();
    launch_sorted_count(dP, N, tile_counts, stream);
    launch_scan_blocks(tile_counts, ntiles, tile_counts + ntiles, stream);
    uint32_t hg = 0;
    MIG_HIP_TRY(hipMemcpyAsync(&hg, tile_counts + ntiles, 4, hipMemcpyDeviceToHost, stream));
    MIG_HIP_TRY(hipStreamSynchronize(stream));
    G = int64_t(hg);
    P.slot_rows = S.get<uint64_t>(G + 2);  // first row of each run (written by the assign kernel)
    upload();
    launch_sorted_assign(dP, N, tile_counts, gid, nullptr, stream);
    launch_clamp_groups(gid, N, uint32_t(G), stream);  // dropped rows: G
  } else {
    // the device-wide table at load <= 2/3, so no probe sequence can fail
    uint64_t const cap = next_pow2(uint64_t(N + N / 2 + 2));
    MIG_EXPECTS(cap + 2 < (uint64_t(1) << 32), "sort-based group-by: table too large");
    int32_t* d_flags = S.get<int32_t>(4);
    P.cap            = cap;
    P.cap_mask       = cap - 1;
    P.probe_limit    = uint32_t(std::min<uint64_t>(cap, 0xffffffffull));
    P.overflow       = d_flags;
    P.special        = reinterpret_cast<uint32_t*>(d_flags + 1);
    if (P.key_mode == KM_INLINE) P.slot_keys = S.get<uint64_t>(cap);
    else P.slot_rows = S.get<uint64_t>(cap);
    upload();
    launch_init_state(dP, P, stream);
    uint32_t* slot_of_row = S.get<uint32_t>(N);
    launch_label_rows(dP, P, num_cus_of_device(), slot_of_row, stream);
    int64_t const nb       = (int64_t(cap) + 2 + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t* block_counts = S.get<uint32_t>(nb + 1);
    launch_count_occupied(dP, P, block_counts, stream);
    launch_scan_blocks(block_counts, nb, block_counts + nb, stream);
    uint32_t hg[2] = {0, 0};
    MIG_HIP_TRY(hipMemcpyAsync(&hg[0], block_counts + nb, 4, hipMemcpyDeviceToHost, stream));
    MIG_HIP_TRY(hipMemcpyAsync(&hg[1], d_flags, 4, hipMemcpyDeviceToHost, stream));
    MIG_HIP_TRY(hipStreamSynchronize(stream));
    MIG_EXPECTS(hg[1] == 0, "sort-based group-by: hash table overflow");
    G             = int64_t(hg[0]);
    uint32_t* map = S.get<uint32_t>(std::max<int64_t>(G, 1));
    launch_compact(dP, P, block_counts, map, stream);
    uint32_t* rank_of_slot = S.get<uint32_t>(cap + 2);
    launch_group_numbers(map, G, slot_of_row, N, rank_of_slot, gid, stream);
  }
  MIG_HIP_TRY(hipGetLastError());
  int const gbits = bits_for(uint64_t(G));

  // ---- 2. rows by group ---------------------------------------------------------
  uint32_t* perm       = S.get<uint32_t>(N);
  uint32_t* gid_sorted = S.get<uint32_t>(N);
  lsd_key const by_group{gid, gbits};
  lsd_sort_rows(&by_group, 1, N, perm, gid_sorted, stream);
  int32_t* off = S.get<int32_t>(G + 2);
  launch_hp_offsets(gid_sorted, N, uint32_t(G + 2), off, stream);
  int32_t kept = 0;
  MIG_HIP_TRY(hipMemcpyAsync(&kept, off + G, 4, hipMemcpyDeviceToHost, stream));
  MIG_HIP_TRY(hipStreamSynchronize(stream));

  // ---- 3. the other aggregations on the rows in group order ---------------------
  std::vector<std::vector<owned_column>> results(static_cast<size_t>(num_requests));
  for (int r = 0; r < num_requests; r++) results[r].resize(size_t(requests[r].num_aggregations));
  struct gathered_col {
    mig_column_view src;
    mig_column_view view;
  };
  std::vector<gathered_col> gathered;
  auto gathered_view = [&](mig_column_view const& v) -> mig_column_view {
    for (auto const& g : gathered)
      if (same_view(g.src, v)) return g.view;
    mig_column_view out{};
    out.type_id = v.type_id;
    out.scale   = v.scale;
    out.size    = kept;
    if (storage_of(v.type_id) == ST_STRING) {
      // sizes in group order, their offsets, then the characters
      dcol const dv = make_dcol(v, false);
      auto* offs    = S.get<int32_t>(size_t(kept) + 1);
      launch_gather_str_sizes(dv, perm, kept, offs, stream);
      launch_exclusive_scan_i32(offs, kept, offs + kept, stream);
      int32_t total = 0;
      MIG_HIP_TRY(hipMemcpyAsync(&total, offs + kept, 4, hipMemcpyDeviceToHost, stream));
      MIG_HIP_TRY(hipStreamSynchronize(stream));
      MIG_EXPECTS(total >= 0, "string values exceed 2^31 characters");
      auto* chars = S.get<uint8_t>(std::max<size_t>(1, size_t(total)));
      launch_gather_string_chars(dv, perm, kept, offs, chars, stream);
      out.data            = chars;
      out.offsets         = offs;
      out.offsets_type_id = MIG_TYPE_INT32;
    } else {
      int32_t const w = type_width(v.type_id);
      auto* data      = S.get<uint8_t>(size_t(std::max(kept, 1)) * w);
      launch_gather_fixed(static_cast<uint8_t const*>(v.data) + int64_t(w) * v.offset, w, perm, kept, data, stream);
      out.data = data;
    }
    out.null_count = 0;
    if (v.null_mask && null_count_of(v, stream) > 0) {
      auto* mask  = S.get<uint32_t>(mask_words_bytes(kept) / 4);
      auto* nulls = S.get<uint32_t>(1);
      MIG_HIP_TRY(hipMemsetAsync(nulls, 0, 4, stream));
      launch_gather_mask(v.null_mask, v.offset, perm, kept, mask, nulls, stream);
      out.null_mask  = mask;
      out.null_count = -1;  // counted again on the device when needed
    }
    gathered.push_back({v, out});
    return out;
  };
  std::vector<mig_aggregation_request> sreqs;
  std::vector<std::vector<mig_aggregation>> saggs;
  std::vector<std::pair<int, int>> origin;  // (request, aggregation) of each sorted-call result
  std::vector<int> sreq_of;
  for (int r = 0; r < num_requests; r++) {
    std::vector<mig_aggregation> al;
    for (int a = 0; a < requests[r].num_aggregations; a++) {
      if (is_sort_only(requests[r].aggregations[a].kind)) continue;
      al.push_back(requests[r].aggregations[a]);
      origin.push_back({r, a});
    }
    if (al.empty()) continue;
    saggs.push_back(std::move(al));
    sreq_of.push_back(r);
  }
  if (!saggs.empty()) {
    mig_column_view gkey{};
    gkey.type_id    = MIG_TYPE_INT32;
    gkey.size       = kept;
    gkey.data       = gid_sorted;
    gkey.null_count = 0;
    for (size_t i = 0; i < saggs.size(); i++) {
      mig_aggregation_request q{};
      q.values           = gathered_view(requests[sreq_of[i]].values);
      q.aggregations     = saggs[i].data();
      q.num_aggregations = int32_t(saggs[i].size());
      sreqs.push_back(q);
    }
    auto sub = groupby_aggregate(&gkey, 1, MIG_NULL_INCLUDE, sreqs.data(), int32_t(sreqs.size()), stream, mr, true);
    MIG_EXPECTS(sub->num_groups == G, "sort-based group-by: group count mismatch");
    size_t k = 0;
    for (size_t i = 0; i < sub->results.size(); i++)
      for (auto& c : sub->results[i]) {
        auto const [r, a] = origin[k++];
        int32_t const kind = requests[r].aggregations[a].kind;
        // ARGMIN / ARGMAX name rows of the group-ordered columns: back to input rows
        if ((kind == MIG_AGG_ARGMIN || kind == MIG_AGG_ARGMAX) && c.col.data != nullptr)
          launch_map_rows(static_cast<int32_t*>(c.col.data), G, perm, stream);
        results[r][a] = c;
        c             = owned_column{};  // now owned by the result
      }
  }

  // ---- 4. MEDIAN / QUANTILE --------------------------------------------------------
  struct sorted_values {
    mig_column_view src;
    uint32_t *lo, *hi, *nul;
  };
  std::vector<sorted_values> sorted;
  auto sorted_of = [&](mig_column_view const& v) -> sorted_values {
    for (auto const& sv : sorted)
      if (same_view(sv.src, v)) return sv;
    int32_t const st  = storage_of(v.type_id);
    bool const wide   = storage_width(st) == 8;
    bool const nulls  = v.null_mask && null_count_of(v, stream) > 0;
    dcol const dv     = make_dcol(v, nulls);
    uint32_t* lo      = S.get<uint32_t>(N);
    uint32_t* hi      = wide ? S.get<uint32_t>(N) : nullptr;
    uint32_t* nul     = nulls ? S.get<uint32_t>(N) : nullptr;
    launch_quantile_keys(dv, N, lo, hi, nul, stream);
    // (group, null, value): least significant first
    std::vector<lsd_key> ks;
    ks.push_back({lo, 32});
    if (wide) ks.push_back({hi, 32});
    if (nulls) ks.push_back({nul, 1});
    ks.push_back({gid, gbits});
    uint32_t* vperm = S.get<uint32_t>(N);
    lsd_sort_rows(ks.data(), int(ks.size()), N, vperm, nullptr, stream);
    sorted_values out{v, S.get<uint32_t>(N), wide ? S.get<uint32_t>(N) : nullptr, nulls ? S.get<uint32_t>(N) : nullptr};
    launch_gather_fixed(lo, 4, vperm, kept, out.lo, stream);
    if (wide) launch_gather_fixed(hi, 4, vperm, kept, out.hi, stream);
    if (nulls) launch_gather_fixed(nul, 4, vperm, kept, out.nul, stream);
    sorted.push_back(out);
    return out;
  };
  std::vector<std::pair<int, int>> qpos;  // (request, aggregation) of the quantile results
  for (int r = 0; r < num_requests; r++) {
    for (int a = 0; a < requests[r].num_aggregations; a++) {
      mig_aggregation const& ag = requests[r].aggregations[a];
      if (!is_sort_only(ag.kind)) continue;
      sorted_values const sv = sorted_of(requests[r].values);
      int32_t const nq       = ag.kind == MIG_AGG_MEDIAN ? 1 : ag.num_quantiles;
      int32_t const interp   = ag.kind == MIG_AGG_MEDIAN ? int32_t(MIG_INTERP_LINEAR) : ag.interpolation;
      std::vector<double> qs = ag.kind == MIG_AGG_MEDIAN ? std::vector<double>{0.5}
                                                           : std::vector<double>(ag.quantiles, ag.quantiles + nq);
      double* dq = S.get<double>(std::max<size_t>(1, qs.size()));
      if (!qs.empty())
        MIG_HIP_TRY(hipMemcpyAsync(dq, qs.data(), sizeof(double) * qs.size(), hipMemcpyHostToDevice, stream));
      int64_t const rows = G * int64_t(nq);
      MIG_EXPECTS(rows < (int64_t(1) << 31), "quantile results exceed 2^31 rows");
      owned_column oc;
      oc.col.type_id    = MIG_TYPE_FLOAT64;
      oc.col.size       = int32_t(rows);
      oc.data_bytes     = sizeof(double) * size_t(rows);
      oc.col.data       = rows > 0 ? A.allocate(oc.data_bytes) : nullptr;
      oc.col.data_bytes = int64_t(oc.data_bytes);
      oc.mask_bytes     = mask_words_bytes(rows);
      oc.col.null_mask  = static_cast<uint32_t*>(A.allocate(oc.mask_bytes));
      MIG_HIP_TRY(hipMemsetAsync(oc.col.null_mask, 0, oc.mask_bytes, stream));
      launch_group_quantiles(sv.lo, sv.hi, sv.nul, off, G, storage_of(requests[r].values.type_id), dq, nq, interp,
                             static_cast<double*>(oc.col.data), oc.col.null_mask, stream);
      MIG_HIP_TRY(hipStreamSynchronize(stream));  // the host copy of the quantiles goes out of scope
      results[r][a] = oc;
      qpos.push_back({r, a});
    }
  }
  std::vector<mig_column*> quantile_cols;
  for (auto const& [r, a] : qpos) quantile_cols.push_back(&results[r][a].col);

  // ---- 5. keys from each group's first row ------------------------------------------
  uint32_t* rep = S.get<uint32_t>(std::max<int64_t>(G, 1));
  launch_first_rows(perm, off, G, rep, stream);
  uint32_t* knulls = S.get<uint32_t>(std::max(1, num_keys));
  MIG_HIP_TRY(hipMemsetAsync(knulls, 0, 4 * size_t(std::max(1, num_keys)), stream));
  for (int i = 0; i < num_keys; i++) {
    mig_column_view const& v = keys[i];
    owned_column oc;
    oc.col.type_id = v.type_id;
    oc.col.scale   = is_decimal_type(v.type_id) ? v.scale : 0;
    oc.col.size    = int32_t(G);
    if (v.null_mask) {
      oc.mask_bytes    = mask_words_bytes(G);
      oc.col.null_mask = static_cast<uint32_t*>(A.allocate(oc.mask_bytes));
      launch_gather_mask(v.null_mask, v.offset, rep, G, oc.col.null_mask, knulls + i, stream);
    }
    if (storage_of(v.type_id) == ST_STRING) {
      oc.col.offsets_type_id = MIG_TYPE_INT32;
      oc.offsets_bytes       = sizeof(int32_t) * size_t(G + 1);
      oc.col.offsets         = A.allocate(oc.offsets_bytes);
      auto* offs             = static_cast<int32_t*>(oc.col.offsets);
      dcol const dv          = make_dcol(v, false);
      launch_gather_str_sizes(dv, rep, G, offs, stream);
      launch_exclusive_scan_i32(offs, G, offs + G, stream);
      int32_t total = 0;
      MIG_HIP_TRY(hipMemcpyAsync(&total, offs + G, 4, hipMemcpyDeviceToHost, stream));
      MIG_HIP_TRY(hipStreamSynchronize(stream));
      MIG_EXPECTS(total >= 0, "unique string keys exceed 2^31 characters");
      oc.data_bytes     = size_t(total);
      oc.col.data       = A.allocate(oc.data_bytes);
      oc.col.data_bytes = total;
      launch_gather_string_chars(dv, rep, G, offs, static_cast<uint8_t*>(oc.col.data), stream);
    } else {
      int32_t const w   = type_width(v.type_id);
      oc.data_bytes     = size_t(w) * size_t(G);
      oc.col.data       = G > 0 ? A.allocate(oc.data_bytes) : nullptr;
      oc.col.data_bytes = int64_t(oc.data_bytes);
      launch_gather_fixed(static_cast<uint8_t const*>(v.data) + int64_t(w) * v.offset, w, rep, G, oc.col.data,
                          stream);
    }
    res->keys.push_back(oc);
  }
  // null counts: the keys' from their gathers, the quantile columns' counted
  std::vector<uint32_t> hk(size_t(std::max(1, num_keys)));
  MIG_HIP_TRY(hipMemcpyAsync(hk.data(), knulls, 4 * hk.size(), hipMemcpyDeviceToHost, stream));
  std::vector<uint32_t> hv(quantile_cols.size());
  if (!quantile_cols.empty()) {
    uint32_t* d_valid = S.get<uint32_t>(quantile_cols.size());
    MIG_HIP_TRY(hipMemsetAsync(d_valid, 0, 4 * quantile_cols.size(), stream));
    std::vector<uint32_t const*> masks;
    for (auto* c : quantile_cols) masks.push_back(c->null_mask);
    // every quantile column of a call has G * nq rows with its own nq
    for (size_t m = 0; m < masks.size(); m++)
      launch_count_valid_list(&masks[m], 1, quantile_cols[m]->size, d_valid + m, stream);
    MIG_HIP_TRY(hipMemcpyAsync(hv.data(), d_valid, 4 * hv.size(), hipMemcpyDeviceToHost, stream));
  }
  MIG_HIP_TRY(hipStreamSynchronize(stream));
  MIG_HIP_TRY(hipGetLastError());
  for (int i = 0; i < num_keys; i++)
    if (res->keys[i].col.null_mask) res->keys[i].col.null_count = int32_t(hk[i]);
  for (size_t m = 0; m < quantile_cols.size(); m++)
    quantile_cols[m]->null_count = quantile_cols[m]->size - int32_t(hv[m]);
  res->results    = std::move(results);
  res->num_groups = int32_t(G);
  std::fill(stats, stats + STAT_COUNT, 0.0);
  stats[STAT_STRATEGY] = 4;  // sort-based
  stats[STAT_GROUPS]   = double(G);
  return res;
}

}  // namespace mig
