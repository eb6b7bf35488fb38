// Host-side planning helpers shared by the single-call driver (groupby.cpp)
// and the batched driver (batch.cpp).
#pragma once

#include "common.hpp"

namespace mig {

// Device blocks of the per-device scratch cache (kept across calls: mapping
// fresh device memory costs far more than the kernels).  A block is given
// back only once no queued work uses it.
void* scratch_take(int dev, size_t bytes);
void scratch_give(int dev, void* p, size_t bytes);

bool same_view(mig_column_view const& a, mig_column_view const& b);
// Nulls of a view (its null_count when known, else counted on the device;
// synchronises `s` in that case).
int32_t null_count_of(mig_column_view const& v, hipStream_t s);
dcol make_dcol(mig_column_view const& v, bool use_mask);
bool is_numeric_type(int32_t t);
bool is_decimal_type(int32_t t);  // DECIMAL32 / DECIMAL64
bool is_chrono_type(int32_t t);
char const* kind_name(int32_t k);
// Result type of (source type, aggregation): aggregation.hpp:1211-1331.
int32_t target_type_of(int32_t src, int32_t kind);
// Throws MIG_ERROR_LOGIC / MIG_ERROR_DATA_TYPE for unsupported combinations.
void check_supported(int32_t src, int32_t kind);
int32_t sum_field_kind(int32_t st);
int32_t sumsq_field_kind(int32_t st);
int32_t minmax_field_kind(int32_t st, bool is_min);
int32_t field_width(int32_t fk);
uint64_t next_pow2(uint64_t x);
// Distinct values of a population from `d` distinct among `n` samples.
double estimate_distinct(double n, double d, double population);
// max(estimate_distinct, Chao's bound from the sample's singletons f1 and doubletons f2)
// (*lower_bound, when given: Chao's bound decided, so the true count may be far larger)
double estimate_distinct_skewed(double n, double d, double f1, double f2, double population,
                                bool* lower_bound = nullptr);

}  // namespace mig
