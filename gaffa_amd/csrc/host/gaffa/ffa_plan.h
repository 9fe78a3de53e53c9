// FFA search plans (include/gaffa/ffa_plan.h:8-52), FFA peaks
// (ffa_peak.h:8-19), boxcar detection plans (ffa_detection.h:20-32) and the
// host-side arithmetic they need (time_series.h downsampled_size/variance).
#pragma once

#include <cstddef>
#include <span>
#include <vector>

namespace gaffa {

struct FfaSearchTask {
  double downsample_factor = 1.0;
  double effective_tsamp = 0.0;
  std::size_t input_nsamples = 0;
  std::size_t prepared_nsamples = 0;
  std::size_t bins = 0;
  std::size_t rows = 0;
  std::size_t rows_eval = 0;
  double period_begin = 0.0;
  double period_end = 0.0;
};

struct FfaSearchPlan {
  std::vector<FfaSearchTask> tasks;
  std::vector<std::size_t> width_trials;
};

struct RiptideFfaPlanOptions {
  double period_min = 0.0;
  double period_max = 0.0;
  std::size_t bins_min = 0;
  std::size_t bins_max = 0;
  std::size_t min_periods = 1;
  double duty_cycle_max = 0.20;
  double width_trial_spacing = 1.5;
  std::size_t max_tasks = 1'000'000;
};

FfaSearchPlan make_riptide_ffa_plan(std::size_t nsamples, double tsamp,
                                    const RiptideFfaPlanOptions& options);

// floor(nsamples / factor), 1 < factor <= nsamples (time_series.cpp:118-122).
std::size_t downsampled_size(std::size_t nsamples, double factor);
// riptide's noise-variance model of the weighted-sum downsampler
// (time_series.cpp:124-138).
double downsampled_variance(std::size_t nsamples, double factor);

struct FfaPeak {
  double period = 0.0;
  double frequency = 0.0;
  std::size_t width = 0;
  double duty_cycle = 0.0;
  std::size_t width_index = 0;
  std::size_t period_index = 0;
  std::size_t phase = 0;
  std::size_t shift = 0;
  std::size_t bins = 0;
  float snr = 0.0F;
};

bool is_better_ffa_peak(const FfaPeak& lhs, const FfaPeak& rhs);
void sort_ffa_peaks(std::vector<FfaPeak>& peaks);

struct FfaBoxcarTrial {
  std::size_t width = 0;
  std::size_t width_index = 0;
  float height = 0.0F;
  float baseline = 0.0F;
  double duty_cycle = 0.0;
};

struct FfaDetectionPlan {
  std::size_t bins = 0;
  std::vector<FfaBoxcarTrial> boxcar_trials;
  std::size_t max_width = 0;
};

FfaDetectionPlan make_ffa_detection_plan(std::span<const std::size_t> width_trials,
                                         std::size_t bins);
float ffa_task_stdnoise(const FfaSearchTask& task);
FfaPeak make_ffa_peak(const FfaSearchTask& task, const FfaBoxcarTrial& trial, std::size_t shift,
                      std::size_t phase, float snr);
// The trial period make_ffa_peak reports for (task, shift).
double ffa_trial_period(const FfaSearchTask& task, std::size_t shift);
// Every evaluated trial period of the plan, task by task (rows_eval each):
// the row order of the dense periodogram.  Ascending for riptide plans.
std::vector<double> ffa_plan_trial_periods(const FfaSearchPlan& plan);

struct FfaSearchOptions {
  float snr_threshold = 6.0F;
  // 0 means unbounded; per series raw peak guard (ffa_search.h:12-17).
  std::size_t max_peaks = 0;
};

struct FfaSearchResult {
  std::vector<FfaPeak> peaks;
};

}  // namespace gaffa
