// TEST INFRASTRUCTURE ONLY - the oracle C ABI implemented by calling gaffa's
// own CPU functions.  oracle/Makefile compiles this file together with the
// reference's CPU sources where they lie (the reference tree is read, never
// copied) into oracle/_ref/libgaffa_ref.so.  Tests use it to pin the
// restatement in gaffa_oracle.cpp; bench.py may time it as the CPU baseline.

#include "oracle_abi.h"

#include "gaffa/candidate.h"
#include "gaffa/dedispersion.h"
#include "gaffa/harmonic.h"
#include "gaffa/dm_search.h"
#include "gaffa/filterbank.h"
#include "gaffa/ffa.h"
#include "gaffa/folding.h"
#include "gaffa/ffa_detection.h"
#include "gaffa/ffa_executor.h"
#include "gaffa/ffa_peak.h"
#include "gaffa/ffa_plan.h"
#include "gaffa/ffa_search.h"
#include "gaffa/internal/dedispersion_delay.h"
#include "gaffa/preprocessing.h"
#include "gaffa/time_series.h"

#include <cstdio>
#include <cstring>
#include <variant>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;
struct too_small : std::exception {};

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return ORC_OK;
  } catch (const too_small&) {
    g_error = "output capacity too small";
    return ORC_TOO_SMALL;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return ORC_INVALID_ARGUMENT;
  } catch (const std::overflow_error& e) {
    g_error = e.what();
    return ORC_OVERFLOW;
  } catch (const std::out_of_range& e) {
    g_error = e.what();
    return ORC_OUT_OF_RANGE;
  } catch (const std::logic_error& e) {
    g_error = e.what();
    return ORC_LOGIC;
  } catch (const std::runtime_error& e) {
    g_error = e.what();
    return ORC_RUNTIME;
  } catch (const std::exception& e) {
    g_error = e.what();
    return ORC_OTHER;
  }
}

template <typename T>
void emit(const std::vector<T>& v, T* out, uint64_t cap, uint64_t* count) {
  *count = v.size();
  if (out == nullptr) return;
  if (cap < v.size()) throw too_small();
  std::copy(v.begin(), v.end(), out);
}

gaffa::FfaSearchTask to_task(const orc_task& t) {
  gaffa::FfaSearchTask r;
  r.downsample_factor = t.downsample_factor;
  r.effective_tsamp = t.effective_tsamp;
  r.input_nsamples = t.input_nsamples;
  r.prepared_nsamples = t.prepared_nsamples;
  r.bins = t.bins;
  r.rows = t.rows;
  r.rows_eval = t.rows_eval;
  r.period_begin = t.period_begin;
  r.period_end = t.period_end;
  return r;
}

orc_task from_task(const gaffa::FfaSearchTask& t) {
  return orc_task{t.downsample_factor, t.effective_tsamp, t.input_nsamples,
                  t.prepared_nsamples, t.bins, t.rows, t.rows_eval,
                  t.period_begin, t.period_end};
}

orc_peak from_peak(const gaffa::FfaPeak& p) {
  return orc_peak{p.period, p.frequency, p.width, p.duty_cycle, p.width_index,
                  p.period_index, p.phase, p.shift, p.bins, p.snr};
}

gaffa::FfaPeak to_peak(const orc_peak& p) {
  gaffa::FfaPeak r;
  r.period = p.period;
  r.frequency = p.frequency;
  r.width = p.width;
  r.duty_cycle = p.duty_cycle;
  r.width_index = p.width_index;
  r.period_index = p.period_index;
  r.phase = p.phase;
  r.shift = p.shift;
  r.bins = p.bins;
  r.snr = p.snr;
  return r;
}

gaffa::FfaSearchPlan to_plan(const orc_task* tasks, uint64_t ntasks,
                             const uint64_t* widths, uint64_t nwidths) {
  gaffa::FfaSearchPlan plan;
  for (uint64_t i = 0; i < ntasks; ++i) plan.tasks.push_back(to_task(tasks[i]));
  plan.width_trials.assign(widths, widths + nwidths);
  return plan;
}

gaffa::RiptideFfaPlanOptions to_options(const orc_plan_options& o) {
  gaffa::RiptideFfaPlanOptions r;
  r.period_min = o.period_min;
  r.period_max = o.period_max;
  r.bins_min = o.bins_min;
  r.bins_max = o.bins_max;
  r.min_periods = o.min_periods;
  r.duty_cycle_max = o.duty_cycle_max;
  r.width_trial_spacing = o.width_trial_spacing;
  r.max_tasks = o.max_tasks;
  return r;
}

gaffa::PreprocessPlan to_preprocess(const orc_preprocess_plan& p) {
  gaffa::PreprocessPlan plan;
  if (p.detrend) {
    gaffa::PreprocessStep s;
    s.kind = gaffa::PreprocessStepKind::DetrendRunningMedian;
    s.detrend_running_median.window_samples = p.window_samples;
    s.detrend_running_median.min_points = p.min_points;
    plan.steps.push_back(s);
  }
  if (p.normalise) {
    gaffa::PreprocessStep s;
    s.kind = gaffa::PreprocessStepKind::Normalise;
    s.normalise.reject_constant = p.reject_constant != 0;
    plan.steps.push_back(s);
  }
  return plan;
}

template <typename T>
gaffa::HostSampleView<T> view(const void* samples, uint64_t nsamples,
                              uint64_t nchans) {
  const gaffa::SampleShape shape{nsamples, 1, nchans};
  const T* data = static_cast<const T*>(samples);
  const std::size_t count =
      data == nullptr ? 0 : gaffa::sample_element_count(shape);
  return gaffa::HostSampleView<T>{std::span<const T>(data, count), shape};
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }
const char* orc_implementation(void) { return "reference"; }

int orc_delay_bins(double dm, double f, double ref, double tsamp, int32_t* out) {
  return guarded(
      [&] { *out = gaffa::internal::dedispersion_delay_bins(dm, f, ref, tsamp); });
}

int orc_multi_dm_delay_table(const double* freq, uint64_t nchans, double dm_low,
                             double dm_step, uint64_t ndm, double ref,
                             double tsamp, uint64_t cb, uint64_t ce,
                             int32_t* out, uint64_t cap, uint64_t* count) {
  return guarded([&] {
    gaffa::MultiDmDedispersionPlan p{dm_low, dm_step, ndm, ref, tsamp, cb, ce};
    emit(gaffa::internal::make_multi_dm_delay_table(
             std::span<const double>(freq, nchans), p),
         out, cap, count);
  });
}

int orc_dedisperse(int32_t kind, int32_t dtype, const void* samples,
                   uint64_t nsamples, uint64_t nchans, const double* freq,
                   double dm_low, double dm_step, uint64_t ndm, double ref,
                   double tsamp, uint64_t cb, uint64_t ce, uint64_t sbc,
                   uint64_t npn, void* out, uint64_t cap,
                   uint64_t* out_nsamples) {
  return guarded([&] {
    const gaffa::MultiDmDedispersionPlan p{dm_low, dm_step, ndm, ref,
                                           tsamp, cb, ce};
    const gaffa::SubbandDedispersionOptions so{sbc, npn};
    const std::span<const double> f(freq, nchans);
    auto finish = [&](auto&& r) {
      *out_nsamples = r.shape.nsamples;
      if (out == nullptr) return;
      if (cap < r.data.size()) throw too_small();
      std::memcpy(out, r.data.data(), r.data.size() * sizeof(r.data[0]));
    };
    auto run = [&](auto tag) {
      using T = decltype(tag);
      const auto v = view<T>(samples, nsamples, nchans);
      if (kind == 0) finish(gaffa::dedisperse_multi_dm_cpu(v, f, p));
      else finish(gaffa::dedisperse_subband_cpu(v, f, p, so));
    };
    if (dtype == 0) run(uint8_t{});
    else if (dtype == 1) run(uint16_t{});
    else if (dtype == 2) run(float{});
    else throw std::invalid_argument("unsupported sample dtype");
  });
}

int orc_dedisperse_single(int32_t dtype, const void* samples, uint64_t nsamples,
                          uint64_t nchans, const double* freq, double dm,
                          double ref, double tsamp, uint64_t cb, uint64_t ce,
                          void* out, uint64_t cap, uint64_t* out_nsamples) {
  return guarded([&] {
    const gaffa::SingleDmDedispersionPlan p{dm, ref, tsamp, cb, ce};
    const std::span<const double> f(freq, nchans);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      auto r = gaffa::dedisperse_single_dm_cpu(view<T>(samples, nsamples, nchans),
                                               f, p);
      *out_nsamples = r.shape.nsamples;
      if (out == nullptr) return;
      if (cap < r.data.size()) throw too_small();
      std::memcpy(out, r.data.data(), r.data.size() * sizeof(r.data[0]));
    };
    if (dtype == 0) run(uint8_t{});
    else if (dtype == 1) run(uint16_t{});
    else if (dtype == 2) run(float{});
    else throw std::invalid_argument("unsupported sample dtype");
  });
}

int orc_dedisperse_spectrum(int32_t dtype, const void* samples, uint64_t nsamples,
                            uint64_t nchans, const double* freq, double dm, double ref,
                            double tsamp, uint64_t cb, uint64_t ce, void* out, uint64_t cap,
                            uint64_t* out_nsamples) {
  return guarded([&] {
    const gaffa::SingleDmDedispersionPlan p{dm, ref, tsamp, cb, ce};
    const std::span<const double> f(freq, nchans);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      auto r = gaffa::dedisperse_spectrum_cpu(view<T>(samples, nsamples, nchans), f, p);
      *out_nsamples = r.shape.nsamples;
      if (out == nullptr) return;
      if (cap < r.data.size()) throw too_small();
      std::memcpy(out, r.data.data(), r.data.size() * sizeof(r.data[0]));
    };
    if (dtype == 0) run(uint8_t{});
    else if (dtype == 1) run(uint16_t{});
    else if (dtype == 2) run(float{});
    else throw std::invalid_argument("unsupported sample dtype");
  });
}

int orc_time_series_stats(const float* x, uint64_t n, double* mean,
                          double* variance, double* stddev) {
  return guarded([&] {
    const auto s = gaffa::time_series_stats_cpu(std::span<const float>(x, n));
    *mean = s.mean;
    *variance = s.variance;
    *stddev = s.stddev;
  });
}

int orc_normalise(const float* x, uint64_t n, int32_t reject, float* out) {
  return guarded([&] {
    gaffa::NormaliseOptions o;
    o.reject_constant = reject != 0;
    gaffa::normalise_cpu(std::span<const float>(x, n), std::span<float>(out, n), o);
  });
}

int orc_downsampled_size(uint64_t n, double f, uint64_t* out) {
  return guarded([&] { *out = gaffa::downsampled_size(n, f); });
}

int orc_downsampled_variance(uint64_t n, double f, double* out) {
  return guarded([&] { *out = gaffa::downsampled_variance(n, f); });
}

int orc_downsample(const float* x, uint64_t n, double f, float* out,
                   uint64_t cap, uint64_t* count) {
  return guarded([&] {
    const uint64_t m = gaffa::downsampled_size(n, f);
    *count = m;
    if (out == nullptr) return;
    if (cap < m) throw too_small();
    gaffa::downsample_weighted_sum_cpu(std::span<const float>(x, n), f,
                                       std::span<float>(out, m));
  });
}

int orc_running_median_window(double seconds, double tsamp, uint64_t* w) {
  return guarded([&] {
    *w = gaffa::running_median_window_seconds(seconds, tsamp).window_samples;
  });
}

int orc_detrend(const float* x, uint64_t n, uint64_t window, uint64_t min_points,
                float* out) {
  return guarded([&] {
    gaffa::DetrendRunningMedianOptions o;
    o.window_samples = window;
    o.min_points = min_points;
    gaffa::detrend_running_median_cpu(std::span<const float>(x, n), o,
                                      std::span<float>(out, n));
  });
}

int orc_preprocess(const float* x, uint64_t n, double tsamp,
                   const orc_preprocess_plan* plan, float* out) {
  return guarded([&] {
    gaffa::TimeSeries ts;
    ts.data.assign(x, x + n);
    ts.tsamp = tsamp;
    const auto r = gaffa::preprocess_time_series_cpu(ts, to_preprocess(*plan));
    std::copy(r.data.begin(), r.data.end(), out);
  });
}

int orc_make_riptide_plan(uint64_t n, double tsamp, const orc_plan_options* o,
                          orc_task* tasks, uint64_t tcap, uint64_t* tcount,
                          uint64_t* widths, uint64_t wcap, uint64_t* wcount) {
  return guarded([&] {
    const auto plan = gaffa::make_riptide_ffa_plan(n, tsamp, to_options(*o));
    *tcount = plan.tasks.size();
    *wcount = plan.width_trials.size();
    if (tasks == nullptr && widths == nullptr) return;
    if (tcap < plan.tasks.size() || wcap < plan.width_trials.size())
      throw too_small();
    for (std::size_t i = 0; i < plan.tasks.size(); ++i)
      tasks[i] = from_task(plan.tasks[i]);
    std::copy(plan.width_trials.begin(), plan.width_trials.end(), widths);
  });
}

int orc_ffa_transform(const float* in, uint64_t rows, uint64_t bins, float* out) {
  return guarded([&] {
    std::vector<float> scratch(rows * bins);
    gaffa::ffa_transform_block_cpu(std::span<const float>(in, rows * bins),
                                   gaffa::FfaTransformShape{rows, bins},
                                   scratch, std::span<float>(out, rows * bins));
  });
}

int orc_ffa_rows(const float* x, uint64_t n, const orc_task* tasks,
                 uint64_t ntasks, float* out, uint64_t cap, uint64_t* count) {
  return guarded([&] {
    gaffa::FfaSearchPlan plan;
    uint64_t total = 0;
    for (uint64_t i = 0; i < ntasks; ++i) {
      plan.tasks.push_back(to_task(tasks[i]));
      total += tasks[i].rows_eval * tasks[i].bins;
    }
    *count = total;
    if (out == nullptr) return;
    if (cap < total) throw too_small();
    uint64_t off = 0;
    gaffa::for_each_ffa_row_cpu(std::span<const float>(x, n), plan,
                                [&](const gaffa::FfaRowView& row) {
                                  std::copy(row.profile.begin(),
                                            row.profile.end(), out + off);
                                  off += row.profile.size();
                                });
  });
}

int orc_task_stdnoise(const orc_task* t, float* out) {
  return guarded([&] { *out = gaffa::ffa_task_stdnoise(to_task(*t)); });
}

int orc_find_ffa_peaks(const float* tr, uint64_t rows, uint64_t bins,
                       const orc_task* task, const uint64_t* widths,
                       uint64_t nw, float stdnoise, float thr,
                       uint64_t max_peaks, orc_peak* out, uint64_t cap,
                       uint64_t* count) {
  return guarded([&] {
    gaffa::FfaDetectionOptions o;
    o.snr_threshold = thr;
    o.max_peaks = max_peaks;
    const auto peaks = gaffa::find_ffa_peaks_cpu(
        std::span<const float>(tr, rows * bins),
        gaffa::FfaTransformShape{rows, bins}, to_task(*task),
        std::span<const std::size_t>(widths, nw), stdnoise, o);
    std::vector<orc_peak> r;
    for (const auto& p : peaks) r.push_back(from_peak(p));
    emit(r, out, cap, count);
  });
}

int orc_search_ffa(const float* x, uint64_t n, const orc_task* tasks,
                   uint64_t ntasks, const uint64_t* widths, uint64_t nw,
                   float thr, uint64_t max_peaks, orc_peak* out, uint64_t cap,
                   uint64_t* count) {
  return guarded([&] {
    gaffa::FfaSearchOptions o;
    o.snr_threshold = thr;
    o.max_peaks = max_peaks;
    const auto res = gaffa::search_ffa_cpu(std::span<const float>(x, n),
                                           to_plan(tasks, ntasks, widths, nw), o);
    std::vector<orc_peak> r;
    for (const auto& p : res.peaks) r.push_back(from_peak(p));
    emit(r, out, cap, count);
  });
}

int orc_sort_peaks(orc_peak* peaks, uint64_t count) {
  return guarded([&] {
    std::vector<gaffa::FfaPeak> v;
    for (uint64_t i = 0; i < count; ++i) v.push_back(to_peak(peaks[i]));
    gaffa::sort_ffa_peaks(v);
    for (uint64_t i = 0; i < count; ++i) peaks[i] = from_peak(v[i]);
  });
}

int orc_find_dm_peaks(const uint32_t* plane, uint64_t ndm, uint64_t ns,
                      const double* dms, double tsamp,
                      const orc_plan_options* po,
                      const orc_preprocess_plan* pp, float thr,
                      uint64_t max_peaks, orc_dm_peak* out, uint64_t cap,
                      uint64_t* count) {
  return guarded([&] {
    gaffa::DedispersedResult<uint32_t> input;
    input.data.assign(plane, plane + ndm * ns);
    input.shape = gaffa::DedispersedShape{ndm, ns};
    gaffa::DmSearchOptions o;
    o.plan = to_options(*po);
    if (pp->detrend || pp->normalise) o.preprocess = to_preprocess(*pp);
    o.snr_threshold = thr;
    o.max_peaks = max_peaks;
    const auto res = gaffa::find_dm_peaks_cpu(
        input, std::span<const double>(dms, ndm), tsamp, o);
    std::vector<orc_dm_peak> r;
    for (const auto& p : res.peaks)
      r.push_back(orc_dm_peak{p.dm, p.dm_index, from_peak(p.peak)});
    emit(r, out, cap, count);
  });
}

int orc_ffa_transform_f64(const double*, uint64_t, uint64_t, double*) {
  g_error = "no reference counterpart";
  return ORC_OTHER;
}

int orc_synth_filterbank(uint64_t, uint64_t, double, double, double, uint64_t,
                         const orc_pulsar*, uint64_t, const uint64_t*, uint64_t,
                         int32_t, uint64_t, uint64_t, uint8_t*) {
  g_error = "no reference counterpart";
  return ORC_OTHER;
}

namespace {

gaffa::DmPeak to_dm_peak(const orc_dm_peak& p) {
  gaffa::DmPeak d;
  d.dm = p.dm;
  d.dm_index = p.dm_index;
  d.peak = gaffa::FfaPeak{p.peak.period,       p.peak.frequency,    p.peak.width,
                          p.peak.duty_cycle,   p.peak.width_index,  p.peak.period_index,
                          p.peak.phase,        p.peak.shift,        p.peak.bins,
                          p.peak.snr};
  return d;
}

orc_candidate from_candidate(const gaffa::Candidate& c) {
  return orc_candidate{c.dm, c.dm_index, c.period, c.frequency, c.width, c.duty_cycle, c.snr,
                       c.peak_count, c.dm_index_min, c.dm_index_max, c.frequency_min,
                       c.frequency_max,
                       orc_dm_peak{c.best_peak.dm, c.best_peak.dm_index,
                                   from_peak(c.best_peak.peak)}};
}

gaffa::Candidate to_candidate(const orc_candidate& c) {
  gaffa::Candidate d;
  d.dm = c.dm;
  d.dm_index = c.dm_index;
  d.period = c.period;
  d.frequency = c.frequency;
  d.width = c.width;
  d.duty_cycle = c.duty_cycle;
  d.snr = c.snr;
  d.peak_count = c.peak_count;
  d.dm_index_min = c.dm_index_min;
  d.dm_index_max = c.dm_index_max;
  d.frequency_min = c.frequency_min;
  d.frequency_max = c.frequency_max;
  d.best_peak = to_dm_peak(c.best_peak);
  return d;
}

}  // namespace

int orc_fold_time_series(const void* data, int32_t dtype, uint64_t n, double period,
                         double tsamp, uint64_t nbin, uint64_t start, float* profile,
                         double* exposure) {
  return guarded([&] {
    const gaffa::FoldOptions o{period, tsamp, nbin, start};
    gaffa::FoldedProfile r;
    if (dtype == 2) {
      r = gaffa::fold_time_series_cpu(
          std::span<const uint32_t>(static_cast<const uint32_t*>(data), n), o);
    } else if (dtype == 3) {
      r = gaffa::fold_time_series_cpu(std::span<const float>(static_cast<const float*>(data), n),
                                      o);
    } else {
      throw std::invalid_argument("fold time series dtype must be u32 or f32");
    }
    std::copy(r.profile.begin(), r.profile.end(), profile);
    std::copy(r.exposure.begin(), r.exposure.end(), exposure);
  });
}

int orc_fold_spectrum(const void* data, int32_t dtype, uint64_t ns, uint64_t nc, double period,
                      double tsamp, uint64_t nbin, uint64_t sub, uint64_t start, uint64_t outc,
                      float* cube, double* exposure, uint64_t cap, uint64_t* nsubint) {
  return guarded([&] {
    const gaffa::FoldSpectrumOptions o{period, tsamp, nbin, sub, start, outc};
    const gaffa::SampleShape shape{ns, 1, nc};
    gaffa::FoldedCube r;
    if (dtype == 0) {
      r = gaffa::fold_spectrum_cpu(gaffa::HostSampleView<uint8_t>{
          std::span<const uint8_t>(static_cast<const uint8_t*>(data), ns * nc), shape}, o);
    } else if (dtype == 1) {
      r = gaffa::fold_spectrum_cpu(gaffa::HostSampleView<uint16_t>{
          std::span<const uint16_t>(static_cast<const uint16_t*>(data), ns * nc), shape}, o);
    } else if (dtype == 3) {
      r = gaffa::fold_spectrum_cpu(gaffa::HostSampleView<float>{
          std::span<const float>(static_cast<const float*>(data), ns * nc), shape}, o);
    } else {
      throw std::invalid_argument("fold spectrum dtype must be u8, u16 or f32");
    }
    *nsubint = r.nsubint;
    if (r.data.size() > cap) throw too_small{};
    std::copy(r.data.begin(), r.data.end(), cube);
    std::copy(r.exposure.begin(), r.exposure.end(), exposure);
  });
}

int orc_read_filterbank(const char* path, int32_t preserve, orc_fb_header* h, double* freq,
                        uint64_t freq_cap, void* data, uint64_t data_cap, uint64_t* data_bytes) {
  return guarded([&] {
    gaffa::FilterbankReadOptions o;
    o.channel_order = preserve != 0 ? gaffa::ChannelOrderPolicy::PreserveFileOrder
                                    : gaffa::ChannelOrderPolicy::FrequencyAscending;
    gaffa::FilterbankData d = gaffa::read_filterbank(path, o);
    const gaffa::FilterbankHeader& g = d.header;
    h->header_size = g.header_size;
    h->nsamples = g.nsamples;
    h->telescope_id = g.telescope_id;
    h->machine_id = g.machine_id;
    h->data_type = g.data_type;
    h->barycentric = g.barycentric;
    h->pulsarcentric = g.pulsarcentric;
    h->ibeam = g.ibeam;
    h->nbeams = g.nbeams;
    h->npuls = g.npuls;
    h->nbins = g.nbins;
    h->nbits = g.nbits;
    h->nifs = g.nifs;
    h->nchans = g.nchans;
    h->az_start = g.az_start;
    h->za_start = g.za_start;
    h->src_raj = g.src_raj;
    h->src_dej = g.src_dej;
    h->tstart = g.tstart;
    h->tsamp = g.tsamp;
    h->fch1 = g.fch1;
    h->foff = g.foff;
    h->refdm = g.refdm;
    h->period = g.period;
    h->uses_frequency_table = g.uses_frequency_table ? 1 : 0;
    std::snprintf(h->rawdatafile, sizeof(h->rawdatafile), "%s", g.rawdatafile.c_str());
    std::snprintf(h->source_name, sizeof(h->source_name), "%s", g.source_name.c_str());
    if (g.frequency_table.size() > freq_cap) throw too_small{};
    std::copy(g.frequency_table.begin(), g.frequency_table.end(), freq);
    std::visit(
        [&](const auto& v) {
          const uint64_t nbytes = v.size() * sizeof(v[0]);
          *data_bytes = nbytes;
          if (data == nullptr) return;
          if (nbytes > data_cap) throw too_small{};
          std::memcpy(data, v.data(), nbytes);
        },
        d.samples);
  });
}

int orc_select_candidates(const orc_dm_peak* peaks, uint64_t n, double obs, double radius,
                          uint64_t dm_radius, int32_t across, uint64_t max_candidates,
                          orc_candidate* out, uint64_t cap, uint64_t* count) {
  return guarded([&] {
    std::vector<gaffa::DmPeak> v;
    for (uint64_t i = 0; i < n; ++i) v.push_back(to_dm_peak(peaks[i]));
    gaffa::CandidateSelectionOptions o;
    o.frequency_cluster_radius = radius;
    o.dm_cluster_radius = dm_radius;
    o.cluster_across_widths = across != 0;
    o.max_candidates = max_candidates;
    std::vector<orc_candidate> r;
    for (const auto& c : gaffa::select_candidates_cpu(v, obs, o)) r.push_back(from_candidate(c));
    emit(r, out, cap, count);
  });
}

int orc_flag_harmonics(const orc_candidate* cands, uint64_t n, double obs, double flo,
                       double fhi, const orc_harmonic_options* ho, orc_candidate* out,
                       orc_harmonic* rel, uint64_t cap, uint64_t* count) {
  return guarded([&] {
    std::vector<gaffa::Candidate> v;
    for (uint64_t i = 0; i < n; ++i) v.push_back(to_candidate(cands[i]));
    gaffa::HarmonicOptions o;
    o.max_harmonic = ho->max_harmonic;
    o.denominator_max = ho->denominator_max;
    o.frequency_tolerance_bins = ho->frequency_tolerance_bins;
    o.phase_distance_max = ho->phase_distance_max;
    o.dm_distance_max = ho->dm_distance_max;
    o.use_snr_consistency = ho->use_snr_consistency != 0;
    o.snr_distance_max = ho->snr_distance_max;
    const auto flagged = gaffa::flag_harmonics_cpu(v, gaffa::HarmonicContext{obs, flo, fhi}, o);
    *count = flagged.size();
    if (flagged.size() > cap) throw too_small{};
    for (std::size_t i = 0; i < flagged.size(); ++i) {
      out[i] = from_candidate(flagged[i].candidate);
      const auto& h = flagged[i].harmonic;
      rel[i] = orc_harmonic{h.is_harmonic ? 1 : 0, h.parent_index, h.numerator, h.denominator,
                            h.frequency_error_bins, h.phase_distance, h.dm_distance,
                            h.expected_snr, h.snr_distance};
    }
  });
}

}  // extern "C"
