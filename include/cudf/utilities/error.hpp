// Exceptions of the cudf API (drop-in for cpp/include/cudf/utilities/
// error.hpp:64-127) and the conversion of the C ABI's status codes into them.
#pragma once

#include <stdexcept>
#include <string>

#include "../../mig_groupby.h"

namespace cudf {

/// An invalid request: sizes, policies, unsupported aggregation kinds.
struct logic_error : public std::logic_error {
  using std::logic_error::logic_error;
};

/// A device runtime call failed.  The reference names this cuda_error
/// (error.hpp:92); here it reports HIP errors.
struct cuda_error : public std::runtime_error {
  using std::runtime_error::runtime_error;
};

/// A type is not supported by the operation (error.hpp:127).
struct data_type_error : public std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

namespace detail {

/// Throw the exception matching a failed C ABI status.
inline void throw_on_error(mig_status s)
{
  if (s == MIG_OK) return;
  std::string msg = mig_last_error();
  switch (s) {
    case MIG_ERROR_HIP: throw cuda_error(msg);
    case MIG_ERROR_DATA_TYPE: throw data_type_error(msg);
    case MIG_ERROR_ALLOC: throw std::bad_alloc();
    case MIG_ERROR_OUT_OF_RANGE: throw std::out_of_range(msg);
    case MIG_ERROR_OVERFLOW: throw std::overflow_error(msg);
    default: throw logic_error(msg);
  }
}

}  // namespace detail
}  // namespace cudf
