// Stream and device-memory-resource handles of the group-by API.  They take
// the place of rmm::cuda_stream_view and rmm::device_async_resource_ref in the
// reference signatures (cpp/include/cudf/groupby.hpp:187-199): a stream is a
// hipStream_t carried as void*, a resource allocates stream-ordered device
// memory.
#pragma once

#include <cstddef>

#include "../../mig_groupby.h"

namespace cudf {

/// Non-owning stream handle; the default is the device's null stream.
class stream_view {
 public:
  constexpr stream_view() = default;
  constexpr explicit stream_view(void* hip_stream) : _stream{hip_stream} {}
  [[nodiscard]] constexpr void* value() const noexcept { return _stream; }

 private:
  void* _stream{nullptr};
};

inline stream_view get_default_stream() { return stream_view{}; }

/// Stream-ordered device memory resource (the role of
/// rmm::mr::device_memory_resource).  Result columns keep a pointer to the
/// resource they were allocated from, so it must outlive them.
class device_memory_resource {
 public:
  virtual ~device_memory_resource() = default;
  virtual void* allocate(std::size_t bytes, stream_view stream)               = 0;
  virtual void deallocate(void* p, std::size_t bytes, stream_view stream)     = 0;

  /// The C ABI view of this resource.
  [[nodiscard]] mig_memory_resource c_resource()
  {
    mig_memory_resource r{};
    r.ctx        = this;
    r.allocate   = [](void* ctx, std::size_t bytes, void* s) -> void* {
      return static_cast<device_memory_resource*>(ctx)->allocate(bytes, stream_view{s});
    };
    r.deallocate = [](void* ctx, void* p, std::size_t bytes, void* s) {
      static_cast<device_memory_resource*>(ctx)->deallocate(p, bytes, stream_view{s});
    };
    return r;
  }
};

/// The library's own stream-ordered allocator (mig_device_allocate).
class default_memory_resource final : public device_memory_resource {
 public:
  void* allocate(std::size_t bytes, stream_view stream) override
  {
    return mig_device_allocate(bytes, stream.value());
  }
  void deallocate(void* p, std::size_t bytes, stream_view stream) override
  {
    mig_device_deallocate(p, bytes, stream.value());
  }
};

inline device_memory_resource* get_current_device_resource()
{
  static default_memory_resource r;
  return &r;
}

}  // namespace cudf
