// Preprocessing plans (preprocessing.cpp:164-210 in gaffa) and the GPU batch
// program that runs them through gaffa_hip_detrend_normalise_batch.
#include "gaffa/preprocessing.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace gaffa {
namespace {

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

}  // namespace

DetrendRunningMedianOptions running_median_window_seconds(double window_seconds, double tsamp) {
  check_arg(window_seconds > 0.0 && std::isfinite(window_seconds),
            "running median window_seconds must be finite and > 0");
  check_arg(tsamp > 0.0 && std::isfinite(tsamp), "running median tsamp must be finite and > 0");
  const double rounded = std::round(window_seconds / tsamp);
  constexpr double max_size = static_cast<double>(std::numeric_limits<std::size_t>::max());
  check_arg(std::isfinite(rounded) && !(rounded > max_size - 1.0),
            "running median window_samples is too large");
  auto w = static_cast<std::size_t>(std::max(1.0, rounded));
  if (w % 2 == 0) ++w;
  DetrendRunningMedianOptions o;
  o.window_samples = w;
  return o;
}

PreprocessPlan make_riptide_preprocess_plan(double tsamp, const RiptidePreprocessOptions& o) {
  PreprocessPlan plan;
  PreprocessStep detrend;
  detrend.kind = PreprocessStepKind::DetrendRunningMedian;
  detrend.detrend_running_median = running_median_window_seconds(o.running_median_width_seconds, tsamp);
  detrend.detrend_running_median.min_points = o.running_median_min_points;
  plan.steps.push_back(detrend);
  if (o.normalise) {
    PreprocessStep n;
    n.kind = PreprocessStepKind::Normalise;
    n.normalise = o.normalise_options;
    plan.steps.push_back(n);
  }
  return plan;
}

GpuPreprocessProgram::GpuPreprocessProgram(PreprocessPlan plan,
                                           const GpuPreprocessProgramOptions& program_options,
                                           const GpuPreprocessExecutionOptions& execution_options)
    : plan_(std::move(plan)), program_(program_options), exec_(execution_options) {
  check_arg(!plan_.steps.empty(), "preprocess plan must contain at least one step");
  // The GPU runs the shape make_riptide_preprocess_plan emits: an optional
  // detrend followed by an optional normalise, each at most once.
  std::size_t i = 0;
  if (i < plan_.steps.size() && plan_.steps[i].kind == PreprocessStepKind::DetrendRunningMedian) {
    const auto& d = plan_.steps[i].detrend_running_median;
    check_arg(d.window_samples != 0 && d.window_samples % 2 == 1,
              "running median detrend window_samples must be odd and > 0");
    check_arg(d.min_points != 0 && d.min_points % 2 == 1,
              "running median detrend min_points must be odd and > 0");
    detrend_ = true;
    window_samples_ = d.window_samples;
    min_points_ = d.min_points;
    ++i;
  }
  if (i < plan_.steps.size() && plan_.steps[i].kind == PreprocessStepKind::Normalise) {
    normalise_ = true;
    reject_constant_ = plan_.steps[i].normalise.reject_constant;
    ++i;
  }
  if (i != plan_.steps.size()) {
    throw std::invalid_argument(
        "GPU preprocessing supports [detrend][, normalise] plans in that order");
  }
  check_arg(exec_.series_tile_size > 0, "preprocess series_tile_size must be > 0");
  GpuDeviceScope scope(program_.device_id);
  status_ = GpuDeviceBuffer<std::int32_t>(exec_.series_tile_size);
  if (exec_.max_nsamples > 0) {
    const std::size_t f = detrend_ ? std::max<std::size_t>(
                                         1, static_cast<std::size_t>(
                                                static_cast<double>(window_samples_) /
                                                static_cast<double>(min_points_)))
                                   : 1;
    const std::size_t bytes =
        gaffa_hip_preprocess_workspace_bytes(exec_.series_tile_size, exec_.max_nsamples, f);
    if (exec_.workspace_bytes_limit != 0 && bytes > exec_.workspace_bytes_limit) {
      throw std::invalid_argument("preprocess workspace exceeds workspace_bytes_limit");
    }
    workspace_ = GpuDeviceBuffer<std::uint8_t>(bytes);
  }
}

void GpuPreprocessProgram::run(void* data, std::size_t nseries, std::size_t nsamples,
                               std::size_t stride, bool input_is_u32) {
  if (poisoned_) throw std::runtime_error("GPU preprocess program is poisoned; recreate it");
  check_arg(data != nullptr, "preprocess batch data must not be null");
  check_arg(nseries > 0 && nsamples > 0, "preprocess batch must not be empty");
  if (stride == 0) stride = nsamples;
  check_arg(stride >= nsamples, "preprocess stride must be >= nsamples");
  const std::size_t f =
      detrend_ ? std::max<std::size_t>(1, static_cast<std::size_t>(
                                              static_cast<double>(window_samples_) /
                                              static_cast<double>(min_points_)))
               : 1;
  const std::size_t median_window = f == 1 ? window_samples_ : min_points_;
  GpuDeviceScope scope(program_.device_id);
  const std::size_t tile = std::min(nseries, exec_.series_tile_size);
  const std::size_t need = gaffa_hip_preprocess_workspace_bytes(tile, nsamples, f);
  if (workspace_.bytes() < need) {
    if (exec_.workspace_bytes_limit != 0 && need > exec_.workspace_bytes_limit) {
      throw std::invalid_argument("preprocess workspace exceeds workspace_bytes_limit");
    }
    workspace_ = GpuDeviceBuffer<std::uint8_t>(need);
  }
  hip_ok(gaffa_hip_memset(status_.data(), 0, status_.bytes(), exec_.stream), "hipMemset");
  for (std::size_t s0 = 0; s0 < nseries; s0 += tile) {
    const std::size_t ns = std::min(tile, nseries - s0);
    gaffa_hip_preprocess_args a{};
    a.data = static_cast<char*>(data) + s0 * stride * 4;
    a.nseries = ns;
    a.nsamples = nsamples;
    a.stride = stride;
    a.input_is_u32 = input_is_u32 ? 1 : 0;
    a.detrend = detrend_ ? 1 : 0;
    a.scrunch_factor = f;
    a.median_window = median_window;
    a.normalise = normalise_ ? 1 : 0;
    a.reject_constant = reject_constant_ ? 1 : 0;
    a.workspace = workspace_.data();
    a.workspace_bytes = workspace_.bytes();
    a.status = status_.data();
    const int rc = gaffa_hip_detrend_normalise_batch(&a, exec_.stream);
    if (rc == GAFFA_HIP_RUNTIME) poisoned_ = true;
    hip_ok(rc, "detrend_normalise_batch");
    pending_series_ = ns;
    if (s0 + tile < nseries) synchronize();  // the status buffer holds one tile
  }
}

void GpuPreprocessProgram::synchronize() {
  const int rc = gaffa_hip_stream_synchronize(exec_.stream);
  if (rc != GAFFA_HIP_OK) {
    poisoned_ = true;
    hip_ok(rc, "hipStreamSynchronize");
  }
  if (pending_series_ == 0) return;
  std::vector<std::int32_t> st(pending_series_);
  copy_to_host(st.data(), status_.data(), st.size() * sizeof(std::int32_t), nullptr);
  pending_series_ = 0;
  hip_ok(gaffa_hip_memset(status_.data(), 0, status_.bytes(), exec_.stream), "hipMemset");
  for (std::int32_t s : st) {
    if (s == 1) throw std::invalid_argument("running median detrend input samples must be finite");
    if (s == 2) {
      throw std::invalid_argument("normalise input standard deviation must be finite and > 0");
    }
    if (s == 3) throw std::invalid_argument("time series variance must be finite and >= 0");
  }
}

void preprocess_time_series_batch_inplace_gpu(GpuPreprocessProgram& program,
                                              MutableGpuTimeSeriesBatchView batch) {
  program.run(batch.data, batch.nseries, batch.nsamples, batch.stride, false);
}

}  // namespace gaffa
