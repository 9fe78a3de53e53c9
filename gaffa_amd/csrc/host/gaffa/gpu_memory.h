// Device memory and stream ownership over the C ABI (the HIP counterpart of
// include/gaffa/cuda_memory.h and launch_cuda.h).  Move-only RAII owners that
// remember their device; views are plain (pointer, count, device) triples.
#pragma once

#include "gaffa_hip.h"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>

namespace gaffa {

// Maps a C-ABI status to the exception class gaffa throws for it.
void throw_hip_status(int status, const char* what);

inline void hip_ok(int status, const char* what) {
  if (status != GAFFA_HIP_OK) throw_hip_status(status, what);
}

template <typename T>
struct GpuSpan {
  T* data = nullptr;
  std::size_t count = 0;
  int device_id = 0;

  std::size_t size() const noexcept { return count; }
  std::size_t bytes() const noexcept { return count * sizeof(T); }
  bool empty() const noexcept { return data == nullptr || count == 0; }
};

// Restores the previously current device on scope exit
// (ffa_bindings.cpp:31-53 in gaffa).
class GpuDeviceScope {
 public:
  explicit GpuDeviceScope(int device_id);
  ~GpuDeviceScope();
  GpuDeviceScope(const GpuDeviceScope&) = delete;
  GpuDeviceScope& operator=(const GpuDeviceScope&) = delete;

 private:
  int previous_ = 0;
  bool restore_ = false;
};

class GpuDeviceMemory {
 public:
  GpuDeviceMemory() = default;
  explicit GpuDeviceMemory(std::size_t bytes);
  GpuDeviceMemory(const GpuDeviceMemory&) = delete;
  GpuDeviceMemory& operator=(const GpuDeviceMemory&) = delete;
  GpuDeviceMemory(GpuDeviceMemory&& other) noexcept;
  GpuDeviceMemory& operator=(GpuDeviceMemory&& other) noexcept;
  ~GpuDeviceMemory();

  void reset();
  void* data() noexcept { return data_; }
  const void* data() const noexcept { return data_; }
  std::size_t bytes() const noexcept { return bytes_; }
  int device_id() const noexcept { return device_id_; }

 private:
  void release_noexcept() noexcept;
  void* data_ = nullptr;
  std::size_t bytes_ = 0;
  int device_id_ = -1;
};

template <typename T>
class GpuDeviceBuffer {
 public:
  GpuDeviceBuffer() = default;
  explicit GpuDeviceBuffer(std::size_t count) : memory_(checked_bytes(count)), count_(count) {}

  T* data() noexcept { return static_cast<T*>(memory_.data()); }
  const T* data() const noexcept { return static_cast<const T*>(memory_.data()); }
  std::size_t size() const noexcept { return count_; }
  std::size_t bytes() const noexcept { return memory_.bytes(); }
  int device_id() const noexcept { return memory_.device_id(); }

  GpuSpan<T> as_span() noexcept { return GpuSpan<T>{data(), count_, device_id()}; }
  GpuSpan<const T> as_span() const noexcept {
    return GpuSpan<const T>{data(), count_, device_id()};
  }

 private:
  static std::size_t checked_bytes(std::size_t count) {
    if (count != 0 && count > std::numeric_limits<std::size_t>::max() / sizeof(T)) {
      throw std::overflow_error("GPU device buffer size overflow");
    }
    return count * sizeof(T);
  }
  GpuDeviceMemory memory_{};
  std::size_t count_ = 0;
};

class GpuStream {
 public:
  GpuStream() = default;
  static GpuStream create();
  GpuStream(const GpuStream&) = delete;
  GpuStream& operator=(const GpuStream&) = delete;
  GpuStream(GpuStream&& o) noexcept : s_(std::exchange(o.s_, nullptr)) {}
  GpuStream& operator=(GpuStream&& o) noexcept {
    if (this != &o) {
      destroy();
      s_ = std::exchange(o.s_, nullptr);
    }
    return *this;
  }
  ~GpuStream() { destroy(); }
  gaffa_hip_stream get() const noexcept { return s_; }
  void synchronize() const;

 private:
  void destroy() noexcept;
  gaffa_hip_stream s_ = nullptr;
};

// Page-locked host buffer (asynchronous copies to and from the device).
class GpuPinnedBuffer {
 public:
  GpuPinnedBuffer() = default;
  explicit GpuPinnedBuffer(std::size_t bytes);
  GpuPinnedBuffer(const GpuPinnedBuffer&) = delete;
  GpuPinnedBuffer& operator=(const GpuPinnedBuffer&) = delete;
  GpuPinnedBuffer(GpuPinnedBuffer&& o) noexcept
      : p_(std::exchange(o.p_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) {}
  GpuPinnedBuffer& operator=(GpuPinnedBuffer&& o) noexcept {
    if (this != &o) {
      release();
      p_ = std::exchange(o.p_, nullptr);
      bytes_ = std::exchange(o.bytes_, 0);
    }
    return *this;
  }
  ~GpuPinnedBuffer() { release(); }
  void* data() noexcept { return p_; }
  const void* data() const noexcept { return p_; }
  std::size_t bytes() const noexcept { return bytes_; }

 private:
  void release() noexcept;
  void* p_ = nullptr;
  std::size_t bytes_ = 0;
};

// An event created on construction (default flags), for ordering work
// across streams.
class GpuEvent {
 public:
  GpuEvent();
  GpuEvent(const GpuEvent&) = delete;
  GpuEvent& operator=(const GpuEvent&) = delete;
  ~GpuEvent();
  void record(gaffa_hip_stream stream);
  void wait(gaffa_hip_stream stream);  // `stream` waits for the last record
  void synchronize();                  // the host waits for the last record
  gaffa_hip_event get() const noexcept { return e_; }

 private:
  gaffa_hip_event e_ = nullptr;
};

struct GpuLaunchOptions {
  int device_id = 0;
  std::size_t threads_per_block = 256;
  gaffa_hip_stream stream = nullptr;
  bool synchronize_after_call = true;
};

void copy_to_device(void* dst, const void* src, std::size_t bytes, gaffa_hip_stream stream);
void copy_to_host(void* dst, const void* src, std::size_t bytes, gaffa_hip_stream stream);

template <typename T>
GpuDeviceBuffer<T> upload(const T* host, std::size_t count) {
  GpuDeviceBuffer<T> buf(count);
  if (count != 0) copy_to_device(buf.data(), host, count * sizeof(T), nullptr);
  return buf;
}

int gpu_device_count();
int gpu_runtime_version();

}  // namespace gaffa
