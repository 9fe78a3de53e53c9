// Host-side FFA planning and peak bookkeeping: the riptide plan generator
// (ffa_plan.cpp:47-171 in gaffa), boxcar trials and noise normalisation
// (ffa_detection.cpp:60-124, 256-284), peak ordering (ffa_peak.cpp:16-45)
// and the downsampling size / variance model (time_series.cpp:118-138).
#include "gaffa/ffa_plan.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace gaffa {
namespace {

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

void validate_factor(std::size_t n, double factor) {
  check_arg(n != 0, "time series must not be empty");
  check_arg(std::isfinite(factor) && factor > 1.0 && factor <= static_cast<double>(n),
            "downsample factor must satisfy 1 < factor <= nsamples");
}

std::vector<std::size_t> width_trials(std::size_t bins_min, double duty, double spacing) {
  std::vector<std::size_t> widths;
  const auto width_max = std::max<std::size_t>(
      1, static_cast<std::size_t>(std::floor(duty * static_cast<double>(bins_min))));
  std::size_t w = 1;
  while (w <= width_max) {
    widths.push_back(w);
    w = std::max(w + 1, static_cast<std::size_t>(std::floor(spacing * static_cast<double>(w))));
  }
  return widths;
}

std::size_t rows_to_evaluate(std::size_t rows, std::size_t bins, double period_ceil_samples) {
  const double shift = static_cast<double>(bins) * (static_cast<double>(rows) - 1.0) *
                       (1.0 - static_cast<double>(bins) / period_ceil_samples);
  return std::min(rows, static_cast<std::size_t>(std::ceil(shift)));
}

double last_evaluated_period(double teff, std::size_t bins, std::size_t rows,
                             std::size_t rows_eval) {
  if (rows == 1 || rows_eval == 1) return teff * static_cast<double>(bins);
  const double b = static_cast<double>(bins);
  const double shift = static_cast<double>(rows_eval - 1) / static_cast<double>(rows - 1);
  return teff * b * b / (b - shift);
}

bool finite_peak(const FfaPeak& p) {
  return std::isfinite(p.snr) && std::isfinite(p.period) && std::isfinite(p.frequency);
}

}  // namespace

std::size_t downsampled_size(std::size_t nsamples, double factor) {
  validate_factor(nsamples, factor);
  return static_cast<std::size_t>(std::floor(static_cast<double>(nsamples) / factor));
}

double downsampled_variance(std::size_t nsamples, double factor) {
  validate_factor(nsamples, factor);
  const double whole = std::floor(factor);
  const double span = static_cast<double>(downsampled_size(nsamples, factor)) * (factor - whole);
  if (span > 1.0) return factor - 1.0 / 3.0;
  return std::pow(whole - 1.0, 2.0) + (2.0 / 3.0) * std::pow(span, 2.0) - span + 1.0;
}

FfaSearchPlan make_riptide_ffa_plan(std::size_t nsamples, double tsamp,
                                    const RiptideFfaPlanOptions& o) {
  check_arg(nsamples > 0, "FFA plan nsamples must be > 0");
  check_arg(tsamp > 0.0, "FFA plan tsamp must be > 0");
  check_arg(o.period_min > 0.0, "FFA plan period_min must be > 0");
  check_arg(o.period_max > o.period_min, "FFA plan period_max must be > period_min");
  check_arg(o.bins_min > 1, "FFA plan bins_min must be > 1");
  check_arg(o.bins_max >= o.bins_min, "FFA plan bins_max must be >= bins_min");
  check_arg(o.min_periods > 0, "FFA plan min_periods must be > 0");
  check_arg(o.duty_cycle_max > 0.0 && o.duty_cycle_max < 1.0,
            "FFA plan duty_cycle_max must be in (0, 1)");
  check_arg(o.width_trial_spacing > 1.0, "FFA plan width_trial_spacing must be > 1");
  check_arg(o.max_tasks > 0, "FFA plan max_tasks must be > 0");
  check_arg(o.period_min >= tsamp * static_cast<double>(o.bins_min),
            "FFA plan period_min must be >= tsamp * bins_min");

  const double tobs = static_cast<double>(nsamples) * tsamp;
  const double pmax = std::min(o.period_max, tobs / static_cast<double>(o.min_periods));
  if (pmax <= o.period_min) {
    throw std::invalid_argument("FFA search range is empty after min_periods cap");
  }
  FfaSearchPlan plan;
  plan.width_trials = width_trials(o.bins_min, o.duty_cycle_max, o.width_trial_spacing);
  const double f0 = o.period_min / (tsamp * static_cast<double>(o.bins_min));
  const double growth = (static_cast<double>(o.bins_max) + 1.0) / static_cast<double>(o.bins_min);
  const auto cycles =
      static_cast<std::size_t>(std::ceil(std::log(pmax / o.period_min) / std::log(growth)));
  for (std::size_t c = 0; c < cycles; ++c) {
    const double factor = f0 * std::pow(growth, static_cast<double>(c));
    const double teff = factor * tsamp;
    const std::size_t prepared = factor == 1.0 ? nsamples : downsampled_size(nsamples, factor);
    const double pmax_samples = pmax / teff;
    const std::size_t stop = std::min(
        {o.bins_max, prepared, static_cast<std::size_t>(std::floor(pmax_samples))});
    if (stop < o.bins_min) continue;
    for (std::size_t bins = o.bins_min; bins <= stop; ++bins) {
      const std::size_t rows = prepared / bins;
      if (rows < o.min_periods) continue;
      const double ceil_samples = std::min(pmax_samples, static_cast<double>(bins) + 1.0);
      const std::size_t rows_eval = rows_to_evaluate(rows, bins, ceil_samples);
      if (rows_eval == 0) continue;
      if (plan.tasks.size() >= o.max_tasks) {
        throw std::runtime_error("FFA search plan exceeds max_tasks");
      }
      FfaSearchTask t;
      t.downsample_factor = factor;
      t.effective_tsamp = teff;
      t.input_nsamples = nsamples;
      t.prepared_nsamples = prepared;
      t.bins = bins;
      t.rows = rows;
      t.rows_eval = rows_eval;
      t.period_begin = teff * static_cast<double>(bins);
      t.period_end = last_evaluated_period(teff, bins, rows, rows_eval);
      plan.tasks.push_back(t);
    }
  }
  return plan;
}

bool is_better_ffa_peak(const FfaPeak& a, const FfaPeak& b) {
  const bool fa = finite_peak(a), fb = finite_peak(b);
  if (fa != fb) return fa;
  if (a.snr != b.snr) return a.snr > b.snr;
  if (a.period != b.period) return a.period < b.period;
  if (a.frequency != b.frequency) return a.frequency < b.frequency;
  if (a.width != b.width) return a.width < b.width;
  if (a.phase != b.phase) return a.phase < b.phase;
  if (a.shift != b.shift) return a.shift < b.shift;
  return a.bins < b.bins;
}

void sort_ffa_peaks(std::vector<FfaPeak>& peaks) {
  std::sort(peaks.begin(), peaks.end(), is_better_ffa_peak);
}

FfaDetectionPlan make_ffa_detection_plan(std::span<const std::size_t> widths, std::size_t bins) {
  check_arg(!widths.empty(), "FFA detection width trials must not be empty");
  FfaDetectionPlan plan;
  plan.bins = bins;
  const float b = static_cast<float>(bins);
  for (std::size_t i = 0; i < widths.size(); ++i) {
    check_arg(widths[i] != 0 && widths[i] < bins,
              "FFA detection width trials must satisfy 0 < width < bins");
    const float w = static_cast<float>(widths[i]);
    FfaBoxcarTrial t;
    t.width = widths[i];
    t.width_index = i;
    t.height = std::sqrt((b - w) / (b * w));
    t.baseline = w / (b - w) * t.height;
    t.duty_cycle = static_cast<double>(widths[i]) / static_cast<double>(bins);
    plan.boxcar_trials.push_back(t);
    plan.max_width = std::max(plan.max_width, widths[i]);
  }
  return plan;
}

float ffa_task_stdnoise(const FfaSearchTask& task) {
  double variance = 1.0;
  if (task.downsample_factor != 1.0) {
    variance = downsampled_variance(task.input_nsamples, task.downsample_factor);
  }
  return static_cast<float>(std::sqrt(static_cast<double>(task.rows) * variance));
}

double ffa_trial_period(const FfaSearchTask& task, std::size_t shift) {
  if (task.rows <= 1) return task.effective_tsamp * static_cast<double>(task.bins);
  const double b = static_cast<double>(task.bins);
  return task.effective_tsamp * b * b /
         (b - static_cast<double>(shift) / static_cast<double>(task.rows - 1));
}

std::vector<double> ffa_plan_trial_periods(const FfaSearchPlan& plan) {
  std::vector<double> v;
  for (const FfaSearchTask& t : plan.tasks)
    for (std::size_t s = 0; s < t.rows_eval; ++s) v.push_back(ffa_trial_period(t, s));
  return v;
}

FfaPeak make_ffa_peak(const FfaSearchTask& task, const FfaBoxcarTrial& trial, std::size_t shift,
                      std::size_t phase, float snr) {
  const double period = ffa_trial_period(task, shift);
  FfaPeak p;
  p.period = period;
  p.frequency = 1.0 / period;
  p.width = trial.width;
  p.duty_cycle = trial.duty_cycle;
  p.width_index = trial.width_index;
  p.period_index = shift;
  p.phase = phase;
  p.shift = shift;
  p.bins = task.bins;
  p.snr = snr;
  return p;
}

}  // namespace gaffa
