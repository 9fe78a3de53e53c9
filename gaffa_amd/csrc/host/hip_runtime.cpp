// RAII device memory / streams over the C ABI, and the mapping of C-ABI
// status codes onto the exception classes gaffa throws.
#include "gaffa/gpu_memory.h"

#include <stdexcept>
#include <string>

namespace gaffa {

void throw_hip_status(int status, const char* what) {
  const std::string msg = gaffa_hip_last_error();
  const std::string full = msg.empty() ? std::string(what) : msg;
  switch (status) {
    case GAFFA_HIP_INVALID_ARGUMENT:
      throw std::invalid_argument(full);
    case GAFFA_HIP_OVERFLOW:
      throw std::overflow_error(full);
    case GAFFA_HIP_OUT_OF_RANGE:
      throw std::out_of_range(full);
    case GAFFA_HIP_LOGIC:
      throw std::logic_error(full);
    default:
      throw std::runtime_error(std::string(what) + ": " + full);
  }
}

GpuDeviceScope::GpuDeviceScope(int device_id) {
  int32_t cur = 0;
  hip_ok(gaffa_hip_get_device(&cur), "hipGetDevice");
  previous_ = cur;
  if (cur != device_id) {
    hip_ok(gaffa_hip_set_device(device_id), "hipSetDevice");
    restore_ = true;
  }
}

GpuDeviceScope::~GpuDeviceScope() {
  if (restore_) (void)gaffa_hip_set_device(previous_);
}

GpuDeviceMemory::GpuDeviceMemory(std::size_t bytes) {
  int32_t dev = 0;
  hip_ok(gaffa_hip_get_device(&dev), "hipGetDevice");
  if (bytes != 0) hip_ok(gaffa_hip_malloc(&data_, bytes), "hipMalloc");
  bytes_ = bytes;
  device_id_ = dev;
}

GpuDeviceMemory::GpuDeviceMemory(GpuDeviceMemory&& o) noexcept
    : data_(std::exchange(o.data_, nullptr)),
      bytes_(std::exchange(o.bytes_, 0)),
      device_id_(std::exchange(o.device_id_, -1)) {}

GpuDeviceMemory& GpuDeviceMemory::operator=(GpuDeviceMemory&& o) noexcept {
  if (this != &o) {
    release_noexcept();
    data_ = std::exchange(o.data_, nullptr);
    bytes_ = std::exchange(o.bytes_, 0);
    device_id_ = std::exchange(o.device_id_, -1);
  }
  return *this;
}

GpuDeviceMemory::~GpuDeviceMemory() { release_noexcept(); }

void GpuDeviceMemory::reset() {
  if (data_ != nullptr) {
    GpuDeviceScope scope(device_id_);
    hip_ok(gaffa_hip_free(data_), "hipFree");
  }
  data_ = nullptr;
  bytes_ = 0;
}

void GpuDeviceMemory::release_noexcept() noexcept {
  if (data_ == nullptr) return;
  int32_t cur = 0;
  const bool switch_dev = gaffa_hip_get_device(&cur) == 0 && cur != device_id_;
  if (switch_dev) (void)gaffa_hip_set_device(device_id_);
  (void)gaffa_hip_free(data_);
  if (switch_dev) (void)gaffa_hip_set_device(cur);
  data_ = nullptr;
  bytes_ = 0;
}

GpuStream GpuStream::create() {
  GpuStream s;
  hip_ok(gaffa_hip_stream_create(&s.s_), "hipStreamCreate");
  return s;
}

void GpuStream::synchronize() const {
  hip_ok(gaffa_hip_stream_synchronize(s_), "hipStreamSynchronize");
}

void GpuStream::destroy() noexcept {
  if (s_ != nullptr) (void)gaffa_hip_stream_destroy(s_);
  s_ = nullptr;
}

GpuPinnedBuffer::GpuPinnedBuffer(std::size_t bytes) : bytes_(bytes) {
  hip_ok(gaffa_hip_host_malloc(&p_, bytes), "hipHostMalloc");
}

void GpuPinnedBuffer::release() noexcept {
  if (p_ != nullptr) (void)gaffa_hip_host_free(p_);
  p_ = nullptr;
  bytes_ = 0;
}

GpuEvent::GpuEvent() { hip_ok(gaffa_hip_event_create(&e_), "hipEventCreate"); }

GpuEvent::~GpuEvent() {
  if (e_ != nullptr) (void)gaffa_hip_event_destroy(e_);
}

void GpuEvent::record(gaffa_hip_stream stream) {
  hip_ok(gaffa_hip_event_record(e_, stream), "hipEventRecord");
}

void GpuEvent::wait(gaffa_hip_stream stream) {
  hip_ok(gaffa_hip_stream_wait_event(stream, e_), "hipStreamWaitEvent");
}

void GpuEvent::synchronize() { hip_ok(gaffa_hip_event_synchronize(e_), "hipEventSynchronize"); }

void copy_to_device(void* dst, const void* src, std::size_t bytes, gaffa_hip_stream stream) {
  hip_ok(gaffa_hip_memcpy(dst, src, bytes, stream), "hipMemcpy host-to-device");
}

void copy_to_host(void* dst, const void* src, std::size_t bytes, gaffa_hip_stream stream) {
  hip_ok(gaffa_hip_memcpy(dst, src, bytes, stream), "hipMemcpy device-to-host");
  if (stream != nullptr) hip_ok(gaffa_hip_stream_synchronize(stream), "hipStreamSynchronize");
}

int gpu_device_count() {
  int32_t n = 0;
  hip_ok(gaffa_hip_device_count(&n), "hipGetDeviceCount");
  return n;
}

int gpu_runtime_version() {
  gaffa_hip_device_info info{};
  if (gpu_device_count() == 0) return 0;
  hip_ok(gaffa_hip_device_info_get(0, &info), "hipGetDeviceProperties");
  return info.runtime_version;
}

}  // namespace gaffa
