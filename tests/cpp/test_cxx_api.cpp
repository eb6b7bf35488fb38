// C++ API test: the reference's group-by known-answer tests written against
// the cudf:: facade (include/cudf/groupby.hpp), the way
// cpp/tests/groupby/*_tests.cpp call cudf::groupby::groupby::aggregate.  Data
// moves with the HIP runtime; everything else goes through the facade.
// Exit status 0 when every case passes.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <cudf/concatenate.hpp>
#include <cudf/groupby.hpp>
#include <cudf/hashing.hpp>
#include <cudf/partitioning.hpp>

namespace {

int failures = 0;

#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);     \
      std::printf(__VA_ARGS__);                            \
      std::printf("\n");                                   \
      failures++;                                          \
    }                                                      \
  } while (0)

struct dev_buf {
  void* p = nullptr;
  explicit dev_buf(size_t bytes) { (void)hipMalloc(&p, bytes ? bytes : 1); }
  ~dev_buf() { (void)hipFree(p); }
};

template <typename T>
std::vector<T> to_host(void const* d, size_t n)
{
  std::vector<T> h(n);
  if (n) (void)hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
  return h;
}

std::vector<bool> validity(cudf::column const& c)
{
  std::vector<bool> v(c.size(), true);
  auto view = c.view();
  if (!view.null_mask()) return v;
  auto words = to_host<uint32_t>(view.null_mask(), (c.size() + 31) / 32);
  for (int i = 0; i < c.size(); i++) v[i] = (words[i / 32] >> (i % 32)) & 1;
  return v;
}

// Host data -> device column view (keeps the buffers alive in `keep`).
template <typename T>
cudf::column_view make_col(std::vector<T> const& vals, std::vector<int> const& valid, cudf::type_id t,
                           std::vector<std::unique_ptr<dev_buf>>& keep)
{
  keep.push_back(std::make_unique<dev_buf>(vals.size() * sizeof(T)));
  (void)hipMemcpy(keep.back()->p, vals.data(), vals.size() * sizeof(T), hipMemcpyHostToDevice);
  void const* data = keep.back()->p;
  cudf::bitmask_type const* mask = nullptr;
  cudf::size_type nulls = 0;
  if (!valid.empty()) {
    std::vector<uint32_t> words((vals.size() + 31) / 32, 0);
    for (size_t i = 0; i < valid.size(); i++) {
      if (valid[i]) words[i / 32] |= 1u << (i % 32);
      else nulls++;
    }
    keep.push_back(std::make_unique<dev_buf>(words.size() * 4));
    (void)hipMemcpy(keep.back()->p, words.data(), words.size() * 4, hipMemcpyHostToDevice);
    mask = static_cast<cudf::bitmask_type const*>(keep.back()->p);
  }
  return cudf::column_view{cudf::data_type{t}, cudf::size_type(vals.size()), data, mask, nulls};
}

// key -> (value, valid) of a single-aggregation result, int32 keys
template <typename R>
std::map<int, std::pair<R, bool>> result_by_key(cudf::table& keys, cudf::column const& res)
{
  auto kv    = keys.view().column(0);
  auto k     = to_host<int32_t>(kv.head(), kv.size());
  auto v     = to_host<R>(res.view().head(), res.size());
  auto valid = validity(res);
  std::map<int, std::pair<R, bool>> m;
  for (size_t i = 0; i < k.size(); i++) m[k[i]] = {v[i], valid[i]};
  return m;
}

template <typename R>
void run_case(char const* name, std::vector<int> keys, std::vector<int> key_valid, std::vector<int> vals,
              std::vector<int> val_valid, std::unique_ptr<cudf::groupby_aggregation> agg,
              std::map<int, std::pair<R, bool>> expect, cudf::null_policy policy = cudf::null_policy::EXCLUDE,
              cudf::sorted keys_sorted = cudf::sorted::NO)
{
  std::vector<std::unique_ptr<dev_buf>> keep;
  auto k = make_col(keys, key_valid, cudf::type_id::INT32, keep);
  auto v = make_col(vals, val_valid, cudf::type_id::INT32, keep);
  std::vector<cudf::groupby::aggregation_request> requests(1);
  requests[0].values = v;
  requests[0].aggregations.push_back(std::move(agg));
  cudf::groupby::groupby gb(cudf::table_view{{k}}, policy, keys_sorted);
  auto [out_keys, results] = gb.aggregate(requests);
  CHECK(out_keys->num_rows() == int(expect.size()), "%s: %d groups, expected %zu", name,
        out_keys->num_rows(), expect.size());
  auto got = result_by_key<R>(*out_keys, *results[0].results[0]);
  for (auto const& [key, ev] : expect) {
    auto it = got.find(key);
    CHECK(it != got.end(), "%s: key %d missing", name, key);
    if (it == got.end()) continue;
    CHECK(it->second.second == ev.second, "%s: key %d validity %d expected %d", name, key,
          int(it->second.second), int(ev.second));
    if (ev.second) {
      bool ok = std::is_floating_point<R>::value
                  ? std::fabs(double(it->second.first) - double(ev.first)) <= 1e-12 * std::fabs(double(ev.first))
                  : it->second.first == ev.first;
      CHECK(ok, "%s: key %d value %.17g expected %.17g", name, key, double(it->second.first),
            double(ev.first));
    }
  }
  std::printf("%s %s\n", failures ? "..." : "ok ", name);
}

}  // namespace

int main()
{
  std::vector<int> const K{1, 2, 3, 1, 2, 2, 1, 3, 3, 2};
  std::vector<int> const V{0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  std::vector<int> const KN{1, 2, 3, 1, 2, 2, 1, 3, 3, 2, 4};
  std::vector<int> const KNV{1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1};
  std::vector<int> const VN{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 4};
  std::vector<int> const VNV{0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0};
  using gba = cudf::groupby_aggregation;

  // sort-only kinds on pre-sorted keys: nunique_tests.cpp:70-88, 175-200 and
  // nth_element_tests.cpp:35-59, 94-119 with the rows stably sorted by key (null keys last)
  std::vector<int> const KS{1, 1, 1, 2, 2, 2, 2, 3, 3, 3};
  auto const SORTED = cudf::sorted::YES;
  auto const EXCL   = cudf::null_policy::EXCLUDE;
  run_case<int32_t>("nunique sorted duplicates", KS, {}, {0, 3, 3, 1, 4, 5, 9, 2, 2, 2}, {},
                    cudf::make_nunique_aggregation<gba>(), {{1, {2, true}}, {2, {4, true}}, {3, {1, true}}},
                    EXCL, SORTED);
  std::vector<int> const KS14{1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 4, 4, 3};
  std::vector<int> const KS14V{1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0};
  std::vector<int> const VS14{0, 3, 6, 1, 4, 5, 9, 2, 2, 2, 8, 4, 4, 7};
  std::vector<int> const VS14V{0, 1, 1, 1, 1, 0, 1, 0, 1, 1, 1, 0, 0, 1};
  run_case<int32_t>("nunique sorted include nulls", KS14, KS14V, VS14, VS14V,
                    cudf::make_nunique_aggregation<gba>(cudf::null_policy::INCLUDE),
                    {{1, {3, true}}, {2, {4, true}}, {3, {2, true}}, {4, {1, true}}}, EXCL, SORTED);
  std::vector<int> const VS{0, 3, 6, 1, 4, 5, 9, 2, 7, 8};
  run_case<int32_t>("nth_element 1 sorted", KS, {}, VS, {}, cudf::make_nth_element_aggregation<gba>(1),
                    {{1, {3, true}}, {2, {4, true}}, {3, {7, true}}}, EXCL, SORTED);
  run_case<int32_t>("nth_element -1 sorted", KS, {}, VS, {}, cudf::make_nth_element_aggregation<gba>(-1),
                    {{1, {6, true}}, {2, {9, true}}, {3, {8, true}}}, EXCL, SORTED);
  run_case<int32_t>("nth_element 1 exclude nulls sorted", KS14, KS14V, VS14, VS14V,
                    cudf::make_nth_element_aggregation<gba>(1, cudf::null_policy::EXCLUDE),
                    {{1, {6, true}}, {2, {4, true}}, {3, {2, true}}, {4, {0, false}}}, EXCL, SORTED);
  // sum_tests.cpp:38-54, 110-132
  run_case<int64_t>("sum basic", K, {}, V, {}, cudf::make_sum_aggregation<gba>(),
                    {{1, {9, true}}, {2, {19, true}}, {3, {17, true}}});
  // argmin_tests.cpp:94-118, argmax_tests.cpp:93-116, product_tests.cpp:115-139 (int32 values)
  run_case<int32_t>("argmin null keys and values", {1, 2, 3, 1, 2, 2, 1, 3, 3, 2, 4},
                    {1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1}, {9, 8, 7, 6, 5, 4, 3, 2, 1, 0, 4},
                    {1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 0}, cudf::make_argmin_aggregation<gba>(),
                    {{1, {3, true}}, {2, {9, true}}, {3, {8, true}}, {4, {0, false}}});
  run_case<int32_t>("argmax null keys and values", {1, 2, 3, 1, 2, 2, 1, 3, 3, 2, 4},
                    {1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1}, {9, 8, 7, 6, 5, 4, 3, 2, 1, 0, 4},
                    {0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0}, cudf::make_argmax_aggregation<gba>(),
                    {{1, {3, true}}, {2, {4, true}}, {3, {7, true}}, {4, {0, false}}});
  run_case<int64_t>("product null keys and values", {1, 2, 3, 1, 2, 2, 1, 3, 3, 2, 4},
                    {1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1}, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 3},
                    {0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0}, cudf::make_product_aggregation<gba>(),
                    {{1, {18, true}}, {2, {36, true}}, {3, {16, true}}, {4, {0, false}}});
  run_case<int64_t>("sum null keys and values", KN, KNV, VN, VNV, cudf::make_sum_aggregation<gba>(),
                    {{1, {9, true}}, {2, {14, true}}, {3, {10, true}}, {4, {0, false}}});
  // count_tests.cpp:32-51, 114-140
  run_case<int32_t>("count basic", K, {}, V, {}, cudf::make_count_aggregation<gba>(),
                    {{1, {3, true}}, {2, {4, true}}, {3, {3, true}}});
  run_case<int32_t>("count all with nulls", KN, KNV, VN, VNV,
                    cudf::make_count_aggregation<gba>(cudf::null_policy::INCLUDE),
                    {{1, {3, true}}, {2, {4, true}}, {3, {2, true}}, {4, {1, true}}});
  // min/max_tests.cpp null_keys_and_values
  run_case<int32_t>("min null keys and values", KN, KNV, VN, VNV, cudf::make_min_aggregation<gba>(),
                    {{1, {3, true}}, {2, {1, true}}, {3, {2, true}}, {4, {0, false}}});
  // mean_tests.cpp:113-136
  run_case<double>("mean null keys and values", KN, KNV, VN, VNV, cudf::make_mean_aggregation<gba>(),
                   {{1, {4.5, true}}, {2, {14.0 / 3, true}}, {3, {5.0, true}}, {4, {0, false}}});
  // var_tests.cpp:35-53, 126-148
  run_case<double>("var basic", K, {}, V, {}, cudf::make_variance_aggregation<gba>(),
                   {{1, {9.0, true}}, {2, {131.0 / 12, true}}, {3, {31.0 / 3, true}}});
  std::vector<int> const VV{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 3};
  std::vector<int> const VVV{0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1};
  run_case<double>("var ddof 2", KN, KNV, VV, VVV, cudf::make_variance_aggregation<gba>(2),
                   {{1, {0, false}}, {2, {98.0 / 3, true}}, {3, {0, false}}, {4, {0, false}}});

  // keys_tests.cpp:94-120: null keys form a group under INCLUDE
  {
    std::vector<std::unique_ptr<dev_buf>> keep;
    auto k = make_col(KN, KNV, cudf::type_id::INT32, keep);
    auto v = make_col(VN, {}, cudf::type_id::INT32, keep);
    std::vector<cudf::groupby::aggregation_request> requests(1);
    requests[0].values = v;
    requests[0].aggregations.push_back(cudf::make_sum_aggregation<gba>());
    cudf::groupby::groupby gb(cudf::table_view{{k}}, cudf::null_policy::INCLUDE);
    auto [out_keys, results] = gb.aggregate(requests);
    auto kv    = to_host<int32_t>(out_keys->view().column(0).head(), out_keys->num_rows());
    auto kval  = validity(out_keys->get_column(0));
    auto sums  = to_host<int64_t>(results[0].results[0]->view().head(), out_keys->num_rows());
    int64_t null_sum = -1;
    for (size_t i = 0; i < kv.size(); i++)
      if (!kval[i]) null_sum = sums[i];
    CHECK(out_keys->num_rows() == 5 && null_sum == 7, "include null keys: %d groups, null-key sum %lld",
          out_keys->num_rows(), (long long)null_sum);
    std::printf("%s include null keys\n", failures ? "..." : "ok ");
  }

  // size mismatch -> cudf::logic_error (groupby.cu:210-214)
  {
    std::vector<std::unique_ptr<dev_buf>> keep;
    auto k = make_col(std::vector<int>{1, 2, 3}, {}, cudf::type_id::INT32, keep);
    auto v = make_col(std::vector<int>{1, 2}, {}, cudf::type_id::INT32, keep);
    std::vector<cudf::groupby::aggregation_request> requests(1);
    requests[0].values = v;
    requests[0].aggregations.push_back(cudf::make_sum_aggregation<gba>());
    bool threw = false;
    try {
      cudf::groupby::groupby(cudf::table_view{{k}}).aggregate(requests);
    } catch (cudf::logic_error const&) {
      threw = true;
    }
    CHECK(threw, "size mismatch did not throw cudf::logic_error");
    std::printf("%s size mismatch throws\n", failures ? "..." : "ok ");
  }

  // hashing docstring vector (python/cudf/cudf/core/indexed_frame.py:2894-2897)
  {
    std::vector<std::unique_ptr<dev_buf>> keep;
    auto c = make_col(std::vector<int64_t>{10, 120, 30}, {}, cudf::type_id::INT64, keep);
    auto h = cudf::hashing::murmurhash3_x86_32(cudf::table_view{{c}});
    auto hv = to_host<int32_t>(h->view().head(), 3);
    CHECK(hv[0] == -1930516747 && hv[1] == 422619251 && hv[2] == -941520876, "murmur3 row hashes %d %d %d",
          hv[0], hv[1], hv[2]);
    std::printf("%s murmurhash3_x86_32\n", failures ? "..." : "ok ");
  }

  // hash_partition (cpp/tests/partitioning/hash_partition_test.cpp:155-179, 53-64)
  {
    std::vector<std::unique_ptr<dev_buf>> keep;
    auto to_hash = make_col(std::vector<int>{1, 2, 3, 4, 5, 6}, {}, cudf::type_id::INT32, keep);
    auto first   = make_col(std::vector<int>{7, 8, 9, 10, 11, 12}, {}, cudf::type_id::INT32, keep);
    auto second  = make_col(std::vector<int>{13, 14, 15, 16, 17, 18}, {}, cudf::type_id::INT32, keep);
    auto [r1, o1] = cudf::hash_partition(cudf::table_view{{to_hash, first}}, {0}, 3);
    auto [r2, o2] = cudf::hash_partition(cudf::table_view{{to_hash, second}}, {0}, 3);
    CHECK(o1.size() == 3 && o1 == o2, "hash_partition offsets differ");
    auto a = to_host<int32_t>(r1->view().column(0).head(), 6);
    auto b = to_host<int32_t>(r2->view().column(0).head(), 6);
    CHECK(a == b, "hash_partition hashed columns differ");
    // every row sits in the partition its row hash names, rows in input order inside it
    auto h  = cudf::hashing::murmurhash3_x86_32(cudf::table_view{{r1->view().column(0)}});
    auto hv = to_host<uint32_t>(h->view().head(), 6);
    auto f  = to_host<int32_t>(r1->view().column(1).head(), 6);
    for (int i = 0; i < 6; i++) {
      int p = 0;
      while (p + 1 < 3 && o1[p + 1] <= i) p++;
      CHECK(int(hv[i] % 3) == p, "row %d in partition %d, hash says %u", i, p, hv[i] % 3);
      if (i > 0 && int(hv[i - 1] % 3) == p) CHECK(f[i - 1] < f[i], "rows out of input order in partition %d", p);
    }
    bool threw = false;
    try {
      cudf::hash_partition(cudf::table_view{{to_hash}}, {-1}, 3);
    } catch (std::out_of_range const&) {
      threw = true;
    }
    CHECK(threw, "invalid column index did not throw std::out_of_range");
    auto pmap = make_col(std::vector<int>{2, 0, 1, 2, 0, 0}, {}, cudf::type_id::INT32, keep);
    auto [r3, o3] = cudf::partition(cudf::table_view{{first}}, pmap, 3);
    auto g = to_host<int32_t>(r3->view().column(0).head(), 6);
    CHECK((o3 == std::vector<int>{0, 3, 4, 6}) && (g == std::vector<int>{8, 11, 12, 9, 7, 10}),
          "partition by map");
    auto cat = cudf::concatenate({cudf::table_view{{to_hash, first}}, cudf::table_view{{second, to_hash}}});
    auto c0  = to_host<int32_t>(cat->view().column(0).head(), 12);
    CHECK(cat->num_rows() == 12 && c0[5] == 6 && c0[6] == 13 && c0[11] == 18, "concatenate");
    std::printf("%s hash_partition / partition / concatenate\n", failures ? "..." : "ok ");
  }

  std::printf("%d failure(s)\n", failures);
  return failures == 0 ? 0 : 1;
}
