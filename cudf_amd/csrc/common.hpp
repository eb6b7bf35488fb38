// Shared host/device definitions of the gfx950 group-by library.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/mig_groupby.h"

namespace mig {

// A named range on the ROCm tools timeline (roctx) for the enclosing scope:
// the reference's CUDF_FUNC_RANGE (NVTX, cpp/src/groupby/groupby.cu:209).
// rocprofv3 --marker-trace records it; without a tool attached it costs a
// call into an empty library function.
struct func_range {
  explicit func_range(char const* name) { roctxRangePushA(name); }
  ~func_range() { roctxRangePop(); }
  func_range(func_range const&)            = delete;
  func_range& operator=(func_range const&) = delete;
};

// ---------------------------------------------------------------------------
// Errors.  The C ABI converts these to mig_status codes; the classes follow
// cudf/utilities/error.hpp:64-127 (logic_error, cuda_error, data_type_error).
struct mig_error : std::runtime_error {
  mig_status status;
  mig_error(mig_status s, std::string const& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(mig_status s, std::string const& msg) { throw mig_error(s, msg); }

#define MIG_EXPECTS(cond, msg) \
  do {                         \
    if (!(cond)) ::mig::fail(MIG_ERROR_LOGIC, std::string("logic error: ") + (msg)); \
  } while (0)

#define MIG_HIP_TRY(call)                                                                \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      ::mig::fail(MIG_ERROR_HIP, std::string("HIP error ") + hipGetErrorName(e_) + " at " + \
                                   __FILE__ + ":" + std::to_string(__LINE__) + ": " +   \
                                   hipGetErrorString(e_));                               \
  } while (0)

// ---------------------------------------------------------------------------
// Physical storage classes of cudf types (cudf's dispatch_storage_type).
enum storage : int32_t {
  ST_NONE = 0,
  ST_I8,
  ST_I16,
  ST_I32,
  ST_I64,
  ST_U8,
  ST_U16,
  ST_U32,
  ST_U64,
  ST_F32,
  ST_F64,
  ST_BOOL,
  ST_STRING
};

__host__ __device__ inline int32_t storage_of(int32_t type_id)
{
  switch (type_id) {
    case MIG_TYPE_INT8: return ST_I8;
    case MIG_TYPE_INT16: return ST_I16;
    case MIG_TYPE_INT32:
    case MIG_TYPE_TIMESTAMP_DAYS:
    case MIG_TYPE_DURATION_DAYS:
    case MIG_TYPE_DECIMAL32: return ST_I32;
    case MIG_TYPE_INT64:
    case MIG_TYPE_TIMESTAMP_SECONDS:
    case MIG_TYPE_TIMESTAMP_MILLISECONDS:
    case MIG_TYPE_TIMESTAMP_MICROSECONDS:
    case MIG_TYPE_TIMESTAMP_NANOSECONDS:
    case MIG_TYPE_DURATION_SECONDS:
    case MIG_TYPE_DURATION_MILLISECONDS:
    case MIG_TYPE_DURATION_MICROSECONDS:
    case MIG_TYPE_DURATION_NANOSECONDS:
    case MIG_TYPE_DECIMAL64: return ST_I64;
    case MIG_TYPE_UINT8: return ST_U8;
    case MIG_TYPE_UINT16: return ST_U16;
    case MIG_TYPE_UINT32: return ST_U32;
    case MIG_TYPE_UINT64: return ST_U64;
    case MIG_TYPE_FLOAT32: return ST_F32;
    case MIG_TYPE_FLOAT64: return ST_F64;
    case MIG_TYPE_BOOL8: return ST_BOOL;
    case MIG_TYPE_STRING: return ST_STRING;
    default: return ST_NONE;
  }
}

__host__ __device__ inline int32_t storage_width(int32_t st)
{
  switch (st) {
    case ST_I8:
    case ST_U8:
    case ST_BOOL: return 1;
    case ST_I16:
    case ST_U16: return 2;
    case ST_I32:
    case ST_U32:
    case ST_F32: return 4;
    case ST_I64:
    case ST_U64:
    case ST_F64: return 8;
    default: return 0;
  }
}

inline bool is_signed_storage(int32_t st) { return st == ST_I8 || st == ST_I16 || st == ST_I32 || st == ST_I64; }
inline bool is_unsigned_storage(int32_t st)
{
  return st == ST_U8 || st == ST_U16 || st == ST_U32 || st == ST_U64 || st == ST_BOOL;
}
inline bool is_float_storage(int32_t st) { return st == ST_F32 || st == ST_F64; }

inline int32_t type_width(int32_t type_id) { return storage_width(storage_of(type_id)); }

// ---------------------------------------------------------------------------
// Device-side column descriptor.  `data` already points at row 0 of the view
// for fixed-width columns; for strings `data` is the chars buffer and
// `offsets` points at the view's first offset (indices stay absolute).
struct dcol {
  const uint8_t* data;
  const uint32_t* mask;  // nullptr when the column has no nulls to honour
  const void* offsets;   // strings only
  int32_t mask_offset;   // bit index of row 0 inside `mask`
  int32_t storage;       // storage enum
  int32_t offsets64;     // strings: 1 if offsets are int64
  int32_t pad;
};

// Global-memory view of a device pointer.  Pointers read from parameter
// structs in device memory are generic to the compiler, and a generic (flat)
// access is counted by the LDS / scalar-load wait counter as well as the
// vector-memory one: every wait for such a load would also wait for the LDS
// operations in flight, and every wait for LDS data for the loads.  Accesses
// through gp(...) compile to global_load / global_store / global_atomic.
template <typename T>
using gmem = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ gmem<T>* gp(T* p)
{
  return (gmem<T>*)p;
}

__device__ __forceinline__ bool dcol_valid(dcol const& c, int64_t row)
{
  if (c.mask == nullptr) return true;
  int64_t b = row + c.mask_offset;
  return (gp(c.mask)[b >> 5] >> (b & 31)) & 1u;
}

}  // namespace mig
