// The allocation-free group-by (mig_groupby_aggregate_into): the hash
// group-by on the device-wide table, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops).  Its scratch —
// the sample table, the hash table, accumulators, the compaction map and the
// parameter block — is carved from one fixed array in device memory, and its
// results go into buffers the caller passes.  No counterpart in the
// reference, whose aggregate() allocates through its memory resource
// (cpp/src/groupby/hash/groupby.cu:404, 478).
//
// Every call on a device shares that device's array, which keeps the last
// call's contents; each call initialises what it uses.  Calls that overlap on
// one device race on it.
#include <mutex>
#include <unordered_map>

#include "groupby.hpp"

namespace mig {
namespace {

// Device scratch of the allocation-free entry point, per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
static __device__ __align__(256) uint8_t g_noalloc_scratch[NOALLOC_SCRATCH_BYTES];

// The array's device address on the current device, looked up once per
// device.  The host-side name of a __device__ array is not a device address:
// only the address hipGetSymbolAddress returns is passed to kernels and
// copies.
uint8_t* noalloc_scratch()
{
  static std::mutex mu;
  static std::unordered_map<int, uint8_t*> ptrs;
  int dev = 0;
  MIG_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  void* p       = nullptr;
  hipError_t e  = hipGetSymbolAddress(&p, HIP_SYMBOL(g_noalloc_scratch));
  if (e != hipSuccess || p == nullptr)
    fail(MIG_ERROR_HIP, std::string("hipGetSymbolAddress of the group-by scratch failed: ") + hipGetErrorString(e));
  ptrs.emplace(dev, static_cast<uint8_t*>(p));
  return static_cast<uint8_t*>(p);
}

}  // namespace

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

int32_t groupby_aggregate_noalloc(mig_column_view const* keys, int32_t num_keys, int32_t null_policy,
                                  mig_aggregation_request const* requests, int32_t num_requests,
                                  mig_column* out_keys, mig_column* out_results, hipStream_t stream)
{
  call_memory m;
  m.arena       = noalloc_scratch();
  m.arena_bytes = NOALLOC_SCRATCH_BYTES;
  m.out_keys    = out_keys;
  m.out_results = out_results;
  return groupby_aggregate_into(keys, num_keys, null_policy, requests, num_requests, stream, m);
}

}  // namespace mig
