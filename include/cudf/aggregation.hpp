// Aggregation requests (drop-in for cpp/include/cudf/aggregation.hpp:86-337:
// aggregation::Kind numbering and the make_*_aggregation factories used by
// group-by).
#pragma once

#include <memory>
#include <vector>

#include "types.hpp"

namespace cudf {

class aggregation {
 public:
  enum Kind {
    SUM,
    PRODUCT,
    MIN,
    MAX,
    COUNT_VALID,
    COUNT_ALL,
    ANY,
    ALL,
    SUM_OF_SQUARES,
    MEAN,
    M2,
    VARIANCE,
    STD,
    MEDIAN,
    QUANTILE,
    ARGMAX,
    ARGMIN,
    NUNIQUE,
    NTH_ELEMENT
  };
  explicit aggregation(Kind k, size_type ddof = 1, null_policy nulls = null_policy::EXCLUDE, size_type n = 0)
    : kind{k}, _ddof{ddof}, _null_handling{nulls}, _n{n}
  {
  }
  virtual ~aggregation() = default;
  [[nodiscard]] virtual std::unique_ptr<aggregation> clone() const
  {
    return std::make_unique<aggregation>(*this);
  }
  Kind kind;
  size_type _ddof;  // VARIANCE / STD
  null_policy _null_handling;  // NUNIQUE, NTH_ELEMENT
  size_type _n;                // NTH_ELEMENT
  std::vector<double> _quantiles;                     // QUANTILE
  interpolation _interpolation = interpolation::LINEAR;  // QUANTILE
};

/// Aggregations valid in a group-by (aggregation.hpp groupby_aggregation).
class groupby_aggregation : public virtual aggregation {
 public:
  explicit groupby_aggregation(Kind k, size_type ddof = 1, null_policy nulls = null_policy::EXCLUDE,
                               size_type n = 0)
    : aggregation(k, ddof, nulls, n)
  {
  }
  [[nodiscard]] std::unique_ptr<aggregation> clone() const override
  {
    return std::make_unique<groupby_aggregation>(*this);
  }
};

template <typename Base = aggregation>
std::unique_ptr<Base> make_sum_aggregation()
{
  return std::make_unique<Base>(aggregation::SUM);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_min_aggregation()
{
  return std::make_unique<Base>(aggregation::MIN);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_max_aggregation()
{
  return std::make_unique<Base>(aggregation::MAX);
}
/// COUNT_VALID under EXCLUDE, COUNT_ALL under INCLUDE (aggregation.cpp:452-463).
template <typename Base = aggregation>
std::unique_ptr<Base> make_count_aggregation(null_policy null_handling = null_policy::EXCLUDE)
{
  return std::make_unique<Base>(null_handling == null_policy::EXCLUDE ? aggregation::COUNT_VALID
                                                                       : aggregation::COUNT_ALL);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_sum_of_squares_aggregation()
{
  return std::make_unique<Base>(aggregation::SUM_OF_SQUARES);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_product_aggregation()
{
  return std::make_unique<Base>(aggregation::PRODUCT);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_argmax_aggregation()
{
  return std::make_unique<Base>(aggregation::ARGMAX);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_argmin_aggregation()
{
  return std::make_unique<Base>(aggregation::ARGMIN);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_mean_aggregation()
{
  return std::make_unique<Base>(aggregation::MEAN);
}
/// Distinct values per group (aggregation.hpp make_nunique_aggregation);
/// taken with keys_are_sorted == sorted::YES.
template <typename Base = aggregation>
std::unique_ptr<Base> make_nunique_aggregation(null_policy null_handling = null_policy::EXCLUDE)
{
  return std::make_unique<Base>(aggregation::NUNIQUE, 1, null_handling);
}
/// The n-th value of each group (aggregation.hpp make_nth_element_aggregation);
/// taken with keys_are_sorted == sorted::YES.
template <typename Base = aggregation>
std::unique_ptr<Base> make_nth_element_aggregation(size_type n, null_policy null_handling = null_policy::INCLUDE)
{
  return std::make_unique<Base>(aggregation::NTH_ELEMENT, 1, null_handling, n);
}
/// The 0.5 quantile, linear interpolation (aggregation.hpp make_median_aggregation).
template <typename Base = aggregation>
std::unique_ptr<Base> make_median_aggregation()
{
  return std::make_unique<Base>(aggregation::MEDIAN);
}
/// Quantiles of each group's valid values; the result column holds
/// num_groups * quantiles.size() doubles, group-major
/// (aggregation.hpp make_quantile_aggregation).
template <typename Base = aggregation>
std::unique_ptr<Base> make_quantile_aggregation(std::vector<double> const& quantiles,
                                                interpolation interp = interpolation::LINEAR)
{
  auto a            = std::make_unique<Base>(aggregation::QUANTILE);
  a->_quantiles     = quantiles;
  a->_interpolation = interp;
  return a;
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_variance_aggregation(size_type ddof = 1)
{
  return std::make_unique<Base>(aggregation::VARIANCE, ddof);
}
template <typename Base = aggregation>
std::unique_ptr<Base> make_std_aggregation(size_type ddof = 1)
{
  return std::make_unique<Base>(aggregation::STD, ddof);
}

}  // namespace cudf
