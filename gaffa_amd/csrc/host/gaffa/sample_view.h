// Filterbank sample views (include/gaffa/sample_view.h:18-92 in gaffa): a
// row-major (time, if, channel) array, channels fastest.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <stdexcept>

namespace gaffa {

struct SampleShape {
  std::size_t nsamples = 0;
  std::size_t nifs = 0;
  std::size_t nchans = 0;
};

inline std::size_t checked_sample_multiply(std::size_t lhs, std::size_t rhs) {
  if (lhs != 0 && rhs > std::numeric_limits<std::size_t>::max() / lhs) {
    throw std::overflow_error("sample shape size overflow");
  }
  return lhs * rhs;
}

inline std::size_t sample_element_count(const SampleShape& shape) {
  return checked_sample_multiply(checked_sample_multiply(shape.nsamples, shape.nifs),
                                 shape.nchans);
}

template <typename T>
struct HostSampleView {
  std::span<const T> data;
  SampleShape shape{};

  std::size_t size() const { return sample_element_count(shape); }
  const T* data_ptr() const noexcept { return data.data(); }
  bool empty() const noexcept { return data.empty(); }
};

template <typename T>
HostSampleView<T> make_host_sample_view(std::span<const T> data, SampleShape shape) {
  if (data.size() != sample_element_count(shape)) {
    throw std::invalid_argument("sample view data size does not match shape");
  }
  return HostSampleView<T>{data, shape};
}

}  // namespace gaffa
