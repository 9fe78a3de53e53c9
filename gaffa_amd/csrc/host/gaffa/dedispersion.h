// Dedispersion plans, results and GPU entry points.  The plan and result
// types keep the layout and defaults of include/gaffa/dedispersion.h:14-114;
// the device entry points replace include/gaffa/dedispersion_cuda.h with
// Gpu-named equivalents that run the HIP kernels through the C ABI.
#pragma once

#include "gaffa/gpu_memory.h"
#include "gaffa/sample_view.h"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <stdexcept>
#include <vector>

namespace gaffa {

struct SingleDmDedispersionPlan {
  double dm = 0.0;
  double ref_frequency_mhz = 0.0;
  double tsamp = 0.0;
  std::size_t chan_begin = 0;
  std::size_t chan_end = 0;
};

struct MultiDmDedispersionPlan {
  double dm_low = 0.0;
  double dm_step = 1.0;
  std::size_t ndm = 0;
  double ref_frequency_mhz = 0.0;
  double tsamp = 0.0;
  std::size_t chan_begin = 0;
  std::size_t chan_end = 0;
};

struct SubbandDedispersionOptions {
  std::size_t subband_channels = 32;
  std::size_t ndm_per_nominal = 32;
};

struct DedispersedShape {
  std::size_t ndm = 0;
  std::size_t nsamples = 0;
};

inline std::size_t dedispersed_element_count(DedispersedShape shape) {
  if (shape.ndm != 0 && shape.nsamples > std::numeric_limits<std::size_t>::max() / shape.ndm) {
    throw std::overflow_error("dedispersed shape size overflow");
  }
  return shape.ndm * shape.nsamples;
}

template <typename T>
struct DedispersedResultView {
  std::span<const T> data;
  DedispersedShape shape{};

  [[nodiscard]] std::size_t size() const { return dedispersed_element_count(shape); }
  [[nodiscard]] bool empty() const noexcept { return data.empty(); }
  [[nodiscard]] std::span<const T> dm_series(std::size_t dm_index) const {
    if (data.size() != size()) {
      throw std::invalid_argument("dedispersed result view data size does not match shape");
    }
    if (dm_index >= shape.ndm) {
      throw std::out_of_range("dedispersed result dm_index is out of range");
    }
    return data.subspan(dm_index * shape.nsamples, shape.nsamples);
  }
};

template <typename T>
struct DedispersedResult {
  std::vector<T> data;
  DedispersedShape shape{};

  [[nodiscard]] DedispersedResultView<T> view() const noexcept {
    return DedispersedResultView<T>{std::span<const T>(data), shape};
  }
};

// ---------------------------------------------------------------- delays --
// dedispersion_delay.h:37-50: round half away from zero of
// 4148.808 * DM * (f^-2 - ref^-2) / tsamp in FP64, range-checked to int32.
std::int32_t dedispersion_delay_bins(double dm, double frequency_mhz,
                                     double ref_frequency_mhz, double tsamp);

std::vector<std::int32_t> make_multi_dm_delay_table(std::span<const double> frequency_mhz,
                                                    const MultiDmDedispersionPlan& plan);

// n_out = nsamples - (delay of the largest DM at the lowest selected
// frequency) (dedispersion_delay.h:95-139).
std::size_t multi_dm_output_nsamples(std::size_t input_nsamples,
                                     std::span<const double> frequency_mhz,
                                     const MultiDmDedispersionPlan& plan);

// Host-built tables of the two-stage subband algorithm
// (dedispersion.cpp:141-189, 351-395).
struct SubbandTables {
  std::size_t nsub = 0;
  std::size_t n_nominal = 0;
  std::vector<std::int32_t> coarse;    // [n_nominal][nchan_sel]
  std::vector<std::int32_t> residual;  // [ndm][nsub]
  std::int32_t max_residual = 0;
  std::int32_t max_coarse = 0;
  // largest, over nominal groups, sum over subbands of the group's largest
  // residual rounded up to 8: the GPU kernel's LDS window budget
  std::int64_t residual_window = 0;
};

SubbandTables make_subband_tables(std::span<const double> frequency_mhz,
                                  const MultiDmDedispersionPlan& plan,
                                  const SubbandDedispersionOptions& options);

// ------------------------------------------------------------- GPU path --
struct GpuDedispersionOptions {
  int device_id = 0;
  std::size_t threads_per_block = 256;  // kept for API parity; the kernels fix 256
  std::size_t time_tile_samples = 81920;  // kept for API parity; LDS tiles are chosen per plan
  gaffa_hip_stream stream = nullptr;
};

template <typename T>
struct GpuDedispersedView {
  T* data = nullptr;
  DedispersedShape shape{};
  std::size_t stride = 0;  // elements between DM rows (>= shape.nsamples)
  int device_id = 0;
};

// Device-resident DM x time plane, rows padded to a multiple of 64 elements.
template <typename T>
struct GpuDedispersedResult {
  GpuDeviceBuffer<T> data;
  DedispersedShape shape{};
  std::size_t stride = 0;
  int device_id = 0;

  GpuDedispersedView<T> view() noexcept { return {data.data(), shape, stride, device_id}; }
  GpuDedispersedView<const T> view() const noexcept {
    return {data.data(), shape, stride, device_id};
  }
};

// Device inputs: the filterbank already resident on the GPU.
struct GpuSampleView {
  const std::uint8_t* data = nullptr;
  SampleShape shape{};
  int device_id = 0;
};

GpuDedispersedResult<std::uint32_t> dedisperse_multi_dm_gpu_device(
    GpuSampleView samples, std::span<const double> frequency_mhz,
    const MultiDmDedispersionPlan& plan, const GpuDedispersionOptions& options = {});

GpuDedispersedResult<std::uint32_t> dedisperse_subband_gpu_device(
    GpuSampleView samples, std::span<const double> frequency_mhz,
    const MultiDmDedispersionPlan& plan, const SubbandDedispersionOptions& subband_options = {},
    const GpuDedispersionOptions& options = {});

GpuDedispersedResult<std::uint32_t> dedisperse_single_dm_gpu_device(
    GpuSampleView samples, std::span<const double> frequency_mhz,
    const SingleDmDedispersionPlan& plan, const GpuDedispersionOptions& options = {});

// Writes into caller storage: ndm rows of `out_stride` elements, the first
// `out_nsamples` of each (0 = the plan's own valid length; a smaller value
// truncates, as a multi-segment DM plan needs a common length).
// Aligned single-DM dynamic spectrum (dedispersion.h:117-142 in gaffa):
// data [nsamples][nchans] of the input element type, valid samples only.
template <typename T>
struct DedispersedSpectrum {
  std::vector<T> data;
  SampleShape shape{};
  double dm = 0.0;
  double tsamp = 0.0;
  std::size_t chan_begin = 0;
  std::size_t chan_end = 0;
};

// out[t][c] = samples[t + delay(c)][chan_begin + c] for the plan's DM
// (dedispersion.cpp:249-289), on the device.  element_size 1, 2 or 4.
void dedisperse_spectrum_gpu_raw(const void* samples, std::size_t element_size,
                                 const SampleShape& shape, std::span<const double> frequency_mhz,
                                 const SingleDmDedispersionPlan& plan, void* out,
                                 std::size_t out_nsamples, const GpuDedispersionOptions& options);
std::size_t spectrum_output_nsamples(std::size_t nsamples, std::span<const double> frequency_mhz,
                                     const SingleDmDedispersionPlan& plan);

template <typename T>
DedispersedSpectrum<T> dedisperse_spectrum_gpu(HostSampleView<T> samples,
                                               std::span<const double> frequency_mhz,
                                               const SingleDmDedispersionPlan& plan,
                                               const GpuDedispersionOptions& options = {}) {
  DedispersedSpectrum<T> r;
  const std::size_t n_out = spectrum_output_nsamples(samples.shape.nsamples, frequency_mhz, plan);
  const std::size_t nc = plan.chan_end - plan.chan_begin;
  r.shape = SampleShape{n_out, 1, nc};
  r.dm = plan.dm;
  r.tsamp = plan.tsamp;
  r.chan_begin = plan.chan_begin;
  r.chan_end = plan.chan_end;
  GpuDeviceScope scope(options.device_id);
  GpuDeviceBuffer<T> in = upload(samples.data.data(), samples.data.size());
  GpuDeviceBuffer<T> out(n_out * nc);
  dedisperse_spectrum_gpu_raw(in.data(), sizeof(T), samples.shape, frequency_mhz, plan, out.data(),
                              n_out, options);
  r.data.resize(n_out * nc);
  copy_to_host(r.data.data(), out.data(), r.data.size() * sizeof(T), options.stream);
  return r;
}

struct GpuChannelMajorView;

// Device-resident delay tables of one DM segment, built once and reused by
// every launch over it: launches then need no host synchronisation.
struct GpuSegmentTables {
  bool subband = false;
  GpuDeviceBuffer<std::int32_t> first, second;  // brute: delays; subband: coarse, residual
  std::int32_t bound_first = 0, bound_second = 0;  // brute: block spread; subband: maxima
  std::int64_t residual_window = 0;                 // subband: SubbandTables::residual_window
  std::vector<std::int32_t> host_first, host_second;  // subband: host copies of the tables
  std::size_t n_nominal = 0;                          // subband: nominal groups
  std::size_t ndm = 0, chan_begin = 0, chan_end = 0;
  std::size_t subband_channels = 0, ndm_per_nominal = 0;
  std::size_t lookahead = 0;  // input samples past an output range its kernel reads
};

// Every nominal group of several subband segments (same channel range and
// subband width) in one table set, so one launch dedisperses all of them:
// concatenated coarse rows and residual rows, and one group entry per
// nominal group naming its coarse row, first output row and DM count.
struct GpuSubbandBatch {
  GpuDeviceBuffer<std::int32_t> coarse, residual;
  GpuDeviceBuffer<gaffa_hip_subband_group> groups;
  std::size_t ngroups = 0, n_coarse_rows = 0, ndm = 0, max_group = 0;
  std::size_t chan_begin = 0, chan_end = 0, subband_channels = 0;
  std::int32_t max_coarse = 0, max_residual = 0;
  std::int64_t residual_window = 0;
};

GpuSubbandBatch make_gpu_subband_batch(const std::vector<GpuSegmentTables>& segments);

// Enqueues output samples [begin, end) of every segment of the batch (rows in
// segment order from `out`).
void enqueue_subband_batch_range(const GpuSubbandBatch& batch, const GpuChannelMajorView& channels,
                                 std::size_t nsamples, std::uint32_t* out, std::size_t out_stride,
                                 std::size_t begin, std::size_t end, gaffa_hip_stream stream);

GpuSegmentTables make_gpu_segment_tables(std::span<const double> frequency_mhz,
                                         const MultiDmDedispersionPlan& plan, bool subband,
                                         const SubbandDedispersionOptions& subband_options);

// Enqueues output samples [begin, end) of one segment on `stream` (no
// synchronisation; the caller checks gaffa_hip_dedisperse_check once).
// Brute force reads `samples` ([nsamples][nchans]); subband reads `channels`
// (the full channel-major copy, rows of channels.stride bytes).
void enqueue_segment_range(const GpuSegmentTables& tables, const std::uint8_t* samples,
                           const GpuChannelMajorView& channels, std::size_t nsamples,
                           std::size_t nchans, std::uint32_t* out, std::size_t out_stride,
                           std::size_t begin, std::size_t end, gaffa_hip_stream stream);

// Channel-major device copy of a filterbank's channels [chan_begin,
// chan_begin + nchans): row c holds channel chan_begin + c over time, zero
// padded from nsamples to stride (the layout subband dedispersion reads).
struct GpuChannelMajorView {
  const std::uint8_t* data = nullptr;
  std::size_t nsamples = 0;
  std::size_t chan_begin = 0;
  std::size_t nchans = 0;
  std::size_t stride = 0;
  int device_id = 0;
  std::size_t length = 0;  // readable bytes per row from data (0 = stride)
};

// Row length a channel-major copy needs for subband dedispersion of `plan`.
std::size_t subband_channel_stride(std::size_t nsamples, std::span<const double> frequency_mhz,
                                   const MultiDmDedispersionPlan& plan,
                                   const SubbandDedispersionOptions& subband_options);

// Transposes channels [chan_begin, chan_end) of `samples` into `out`
// ([chan_end - chan_begin][stride] bytes), enqueued on options.stream.
void transpose_filterbank_gpu_into(GpuSampleView samples, std::size_t chan_begin,
                                   std::size_t chan_end, std::uint8_t* out, std::size_t stride,
                                   const GpuDedispersionOptions& options);

// Subband dedispersion from a channel-major copy (the plan's channel range
// must lie inside the copy's).
void dedisperse_subband_gpu_cm_into(const GpuChannelMajorView& channels,
                                    std::span<const double> frequency_mhz,
                                    const MultiDmDedispersionPlan& plan,
                                    const SubbandDedispersionOptions& subband_options,
                                    std::uint32_t* out, std::size_t out_stride,
                                    const GpuDedispersionOptions& options,
                                    std::size_t out_nsamples = 0);

void dedisperse_subband_gpu_into(GpuSampleView samples, std::span<const double> frequency_mhz,
                                 const MultiDmDedispersionPlan& plan,
                                 const SubbandDedispersionOptions& subband_options,
                                 std::uint32_t* out, std::size_t out_stride,
                                 const GpuDedispersionOptions& options = {},
                                 std::size_t out_nsamples = 0);
void dedisperse_multi_dm_gpu_into(GpuSampleView samples, std::span<const double> frequency_mhz,
                                  const MultiDmDedispersionPlan& plan, std::uint32_t* out,
                                  std::size_t out_stride,
                                  const GpuDedispersionOptions& options = {},
                                  std::size_t out_nsamples = 0);

// Host-input conveniences: upload, run, copy back (dedispersion_cuda.h:163-209).
DedispersedResult<std::uint32_t> dedisperse_single_dm_gpu(
    HostSampleView<std::uint8_t> samples, std::span<const double> frequency_mhz,
    const SingleDmDedispersionPlan& plan, const GpuDedispersionOptions& options = {});
DedispersedResult<std::uint32_t> dedisperse_multi_dm_gpu(
    HostSampleView<std::uint8_t> samples, std::span<const double> frequency_mhz,
    const MultiDmDedispersionPlan& plan, const GpuDedispersionOptions& options = {});
DedispersedResult<std::uint32_t> dedisperse_subband_gpu(
    HostSampleView<std::uint8_t> samples, std::span<const double> frequency_mhz,
    const MultiDmDedispersionPlan& plan, const SubbandDedispersionOptions& subband_options = {},
    const GpuDedispersionOptions& options = {});

DedispersedResult<std::uint32_t> copy_to_host(const GpuDedispersedResult<std::uint32_t>& result);

}  // namespace gaffa
