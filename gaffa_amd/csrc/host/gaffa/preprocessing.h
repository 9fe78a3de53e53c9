// Preprocessing plans (include/gaffa/preprocessing.h:11-42) and the GPU
// batch program that replaces CudaPreprocessProgram
// (include/gaffa/preprocessing_cuda.h:19-89).
#pragma once

#include "gaffa/gpu_memory.h"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace gaffa {

struct NormaliseOptions {
  bool reject_constant = true;
};

struct DetrendRunningMedianOptions {
  std::size_t window_samples = 101;
  std::size_t min_points = 101;
};

enum class PreprocessStepKind {
  DetrendRunningMedian,
  Normalise,
};

struct PreprocessStep {
  PreprocessStepKind kind = PreprocessStepKind::Normalise;
  DetrendRunningMedianOptions detrend_running_median{};
  NormaliseOptions normalise{};
};

struct PreprocessPlan {
  std::vector<PreprocessStep> steps;
};

struct RiptidePreprocessOptions {
  double running_median_width_seconds = 5.0;
  std::size_t running_median_min_points = 101;
  bool normalise = true;
  NormaliseOptions normalise_options{};
};

DetrendRunningMedianOptions running_median_window_seconds(double window_seconds, double tsamp);
PreprocessPlan make_riptide_preprocess_plan(double tsamp,
                                            const RiptidePreprocessOptions& options = {});

struct GpuTimeSeriesBatchView {
  const float* data = nullptr;
  std::size_t nseries = 0;
  std::size_t nsamples = 0;
  std::size_t stride = 0;  // 0 means nsamples
  int device_id = 0;
};

struct MutableGpuTimeSeriesBatchView {
  float* data = nullptr;
  std::size_t nseries = 0;
  std::size_t nsamples = 0;
  std::size_t stride = 0;
  int device_id = 0;
};

struct GpuPreprocessProgramOptions {
  int device_id = 0;
};

struct GpuPreprocessExecutionOptions {
  std::size_t series_tile_size = 1024;
  std::size_t max_nsamples = 0;
  std::size_t workspace_bytes_limit = 0;  // 0: unlimited
  gaffa_hip_stream stream = nullptr;
};

// Owns only metadata and workspace; callers provide the device plane, which
// is transformed in place.  synchronize() reports deferred data errors (non
// finite input, constant series) with the messages of gaffa's CPU path.
class GpuPreprocessProgram {
 public:
  GpuPreprocessProgram(PreprocessPlan plan, const GpuPreprocessProgramOptions& program_options = {},
                       const GpuPreprocessExecutionOptions& execution_options = {});
  GpuPreprocessProgram(GpuPreprocessProgram&&) noexcept = default;
  GpuPreprocessProgram& operator=(GpuPreprocessProgram&&) noexcept = default;

  int device_id() const noexcept { return program_.device_id; }
  std::size_t tile_capacity() const noexcept { return exec_.series_tile_size; }
  std::size_t max_nsamples() const noexcept { return exec_.max_nsamples; }
  std::size_t workspace_bytes() const noexcept { return workspace_.bytes(); }
  void synchronize();

  // In place over [nseries][stride]; input_is_u32 converts uint32 sums first.
  void run(void* data, std::size_t nseries, std::size_t nsamples, std::size_t stride,
           bool input_is_u32);

 private:
  PreprocessPlan plan_;
  GpuPreprocessProgramOptions program_;
  GpuPreprocessExecutionOptions exec_;
  bool detrend_ = false;
  bool normalise_ = false;
  bool reject_constant_ = true;
  std::size_t window_samples_ = 0;
  std::size_t min_points_ = 0;
  GpuDeviceBuffer<std::uint8_t> workspace_;
  GpuDeviceBuffer<std::int32_t> status_;
  std::size_t pending_series_ = 0;
  bool poisoned_ = false;
};

void preprocess_time_series_batch_inplace_gpu(GpuPreprocessProgram& program,
                                              MutableGpuTimeSeriesBatchView batch);

}  // namespace gaffa
