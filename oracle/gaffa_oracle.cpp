// TEST INFRASTRUCTURE ONLY - the CPU oracle for the gaffa search hot path.
//
// An independent restatement of gaffa's CPU path (reference files cited per
// function).  It is loaded only by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg, always as the checker; the product library
// (gaffa_amd/) never links or calls it.  It is built with -ffp-contract=off so
// its FP32/FP64 arithmetic rounds exactly as the reference's no-FMA x86-64 build.
//
// Pinning (what fixes each function besides this text):
//  * delay model, dedispersion, subband: golden vectors from the reference's
//    own tests, basetest.fil fixtures made by the reference build, the
//    subband(1,1) == brute-force identity, DM 0 == channel sum.
//  * downsample / variance: reference known answers (test_time_series.cpp).
//  * running median / detrend / normalise: reference known answers
//    (test_preprocessing.cpp) + reference-build fixtures at the config windows.
//  * FFA plan: reference known plan (test_ffa_plan.cpp) + reference-build
//    fixtures at the config shapes.
//  * FFA transform: the riptide 8x8 and 5x4 golden transforms, phase-rotation
//    invariance, and brute-force agreement with the FP64 transform.
//  * detection / search / DM search: reference S/N closed form tests and
//    reference-build peak-list fixtures.
//  * synthetic generator (no reference counterpart): closed forms (noise mean
//    and variance, DM-0 channel sums, dedispersing at the true DM re-aligns
//    the injected pulse) and a committed SHA-256 of a small buffer.

#include "oracle_abi.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
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

void require(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

template <typename T>
void emit(const std::vector<T>& values, T* out, uint64_t capacity,
          uint64_t* count) {
  *count = values.size();
  if (out == nullptr) return;
  if (capacity < values.size()) throw too_small();
  std::copy(values.begin(), values.end(), out);
}

// ---------------------------------------------------------------- delays ----

// include/gaffa/internal/dedispersion_delay.h:37-50: round half away from zero
// of 4148.808*DM*(f^-2 - ref^-2)/tsamp, all in FP64, range-checked to int32.
int32_t delay_bins(double dm, double f, double ref, double tsamp) {
  const double inv = 1.0 / (f * f) - 1.0 / (ref * ref);
  const double seconds = 4148.808 * dm * inv;
  const double bins = std::round(seconds / tsamp);
  if (bins < static_cast<double>(std::numeric_limits<int32_t>::min()) ||
      bins > static_cast<double>(std::numeric_limits<int32_t>::max())) {
    throw std::overflow_error("dedispersion delay exceeds int32 range");
  }
  return static_cast<int32_t>(bins);
}

struct DmPlan {
  double dm_low = 0, dm_step = 1;
  uint64_t ndm = 0;
  double ref = 0, tsamp = 0;
  uint64_t chan_begin = 0, chan_end = 0;
};

// dedispersion.cpp:53-139 validation order and messages.
void validate_common(const void* data, uint64_t nsamples, uint64_t nchans,
                     const double* freq, double ref, double tsamp,
                     uint64_t cb, uint64_t ce) {
  require(data != nullptr, "dedispersion samples data must not be null");
  require(nsamples != 0, "dedispersion requires at least one sample");
  require(nchans != 0, "dedispersion requires at least one channel");
  for (uint64_t c = 0; c < nchans; ++c) {
    require(std::isfinite(freq[c]) && freq[c] > 0.0,
            "frequency values must be finite and positive");
  }
  require(std::isfinite(ref) && ref > 0.0,
          "reference frequency must be positive");
  require(std::isfinite(tsamp) && tsamp > 0.0, "tsamp must be positive");
  require(cb < ce, "channel range must be non-empty");
  require(ce <= nchans, "channel range exceeds sample channel count");
  for (uint64_t c = cb; c < ce; ++c) {
    require(!(freq[c] > ref),
            "dedispersion requires selected frequencies <= reference "
            "frequency");
  }
}

void validate_multi(const DmPlan& p) {
  require(std::isfinite(p.dm_low) && p.dm_low >= 0.0,
          "dm_low must be finite and non-negative");
  require(p.ndm != 0, "dedispersion requires at least one DM");
  require(std::isfinite(p.dm_step) && p.dm_step > 0.0,
          "dm_step must be positive");
}

double min_freq(const double* freq, uint64_t cb, uint64_t ce) {
  double m = freq[cb];
  for (uint64_t c = cb + 1; c < ce; ++c) m = std::min(m, freq[c]);
  return m;
}

// dedispersion_delay.h:95-105: n_out = n - max delay, non-empty.
uint64_t valid_nout(uint64_t n, int32_t max_delay) {
  require(max_delay >= 0, "dedispersion requires non-negative delays");
  require(static_cast<uint64_t>(max_delay) < n,
          "dedispersion valid output range is empty");
  return n - static_cast<uint64_t>(max_delay);
}

// dedispersion_delay.h:130-139: the largest DM at the lowest selected
// frequency fixes the valid output length.
uint64_t multi_nout(const double* freq, const DmPlan& p, uint64_t n) {
  const double dm_max = p.dm_low + static_cast<double>(p.ndm - 1) * p.dm_step;
  return valid_nout(
      n, delay_bins(dm_max, min_freq(freq, p.chan_begin, p.chan_end), p.ref,
                    p.tsamp));
}

template <typename T>
struct Acc {
  using type = uint32_t;
};
template <>
struct Acc<float> {
  using type = float;
};

// dedispersion.cpp:197-215 / 291-329: out[dm][t] = sum over the selected
// channels, in channel order, of x[t + d(dm,c)][c]; samples past the end of
// the input contribute nothing.
template <typename T>
std::vector<typename Acc<T>::type> brute_force(const T* x, uint64_t n,
                                               uint64_t nchans,
                                               const double* freq,
                                               const DmPlan& p,
                                               uint64_t* nout) {
  using A = typename Acc<T>::type;
  const uint64_t nc = p.chan_end - p.chan_begin;
  std::vector<int32_t> d(p.ndm * nc);
  for (uint64_t i = 0; i < p.ndm; ++i) {
    const double dm = p.dm_low + static_cast<double>(i) * p.dm_step;
    for (uint64_t c = 0; c < nc; ++c)
      d[i * nc + c] = delay_bins(dm, freq[p.chan_begin + c], p.ref, p.tsamp);
  }
  *nout = multi_nout(freq, p, n);
  const uint64_t no = *nout;
  std::vector<A> out(p.ndm * no);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(p.ndm); ++i) {
    for (int64_t t = 0; t < static_cast<int64_t>(no); ++t) {
      const int32_t* di = d.data() + i * nc;
      A s = 0;
      for (uint64_t c = 0; c < nc; ++c) {
        const int64_t tt = t + di[c];
        if (tt >= 0 && tt < static_cast<int64_t>(n))
          s += static_cast<A>(x[tt * nchans + p.chan_begin + c]);
      }
      out[i * no + t] = s;
    }
  }
  return out;
}

// dedispersion.cpp:141-189 and 351-453: stage 1 sums each subband of
// subband_channels channels at the coarse delays of its nominal DM
// (dm_low + k*ndm_per_nominal*dm_step) over the whole input length; stage 2
// adds the subbands at the residual delay of (DM - nominal DM) evaluated at
// the subband's lower-middle channel frequency f[(begin+end-1)/2].  Both
// stages drop samples past the end of their input.
template <typename T>
std::vector<typename Acc<T>::type> subband(const T* x, uint64_t n,
                                           uint64_t nchans,
                                           const double* freq,
                                           const DmPlan& p, uint64_t sbc,
                                           uint64_t npn, uint64_t* nout) {
  using A = typename Acc<T>::type;
  require(sbc != 0, "subband_channels must be positive");
  require(npn != 0, "ndm_per_nominal must be positive");
  const uint64_t nc = p.chan_end - p.chan_begin;
  const uint64_t nsub = (nc + sbc - 1) / sbc;
  const uint64_t nnom = (p.ndm + npn - 1) / npn;
  std::vector<uint64_t> sb_begin(nsub), sb_end(nsub);
  std::vector<double> sb_freq(nsub);
  for (uint64_t s = 0; s < nsub; ++s) {
    sb_begin[s] = p.chan_begin + s * sbc;
    sb_end[s] = std::min(sb_begin[s] + sbc, p.chan_end);
    sb_freq[s] = freq[(sb_begin[s] + sb_end[s] - 1) / 2];
  }
  std::vector<int32_t> coarse(nnom * nc), resid(p.ndm * nsub);
  for (uint64_t k = 0; k < nnom; ++k) {
    const double nominal =
        p.dm_low + static_cast<double>(k * npn) * p.dm_step;
    for (uint64_t c = 0; c < nc; ++c)
      coarse[k * nc + c] =
          delay_bins(nominal, freq[p.chan_begin + c], p.ref, p.tsamp);
  }
  for (uint64_t i = 0; i < p.ndm; ++i) {
    const uint64_t k = i / npn;
    const double nominal =
        p.dm_low + static_cast<double>(k * npn) * p.dm_step;
    const double dm = p.dm_low + static_cast<double>(i) * p.dm_step;
    for (uint64_t s = 0; s < nsub; ++s)
      resid[i * nsub + s] = delay_bins(dm - nominal, sb_freq[s], p.ref, p.tsamp);
  }
  *nout = multi_nout(freq, p, n);
  const uint64_t no = *nout;
  std::vector<A> inter(nnom * nsub * n);
#pragma omp parallel for collapse(3) schedule(static)
  for (int64_t k = 0; k < static_cast<int64_t>(nnom); ++k) {
    for (int64_t s = 0; s < static_cast<int64_t>(nsub); ++s) {
      for (int64_t t = 0; t < static_cast<int64_t>(n); ++t) {
        A acc = 0;
        for (uint64_t c = sb_begin[s]; c < sb_end[s]; ++c) {
          const int64_t tt = t + coarse[k * nc + (c - p.chan_begin)];
          if (tt >= 0 && tt < static_cast<int64_t>(n))
            acc += static_cast<A>(x[tt * nchans + c]);
        }
        inter[(k * nsub + s) * n + t] = acc;
      }
    }
  }
  std::vector<A> out(p.ndm * no);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(p.ndm); ++i) {
    for (int64_t t = 0; t < static_cast<int64_t>(no); ++t) {
      const uint64_t k = static_cast<uint64_t>(i) / npn;
      A acc = 0;
      for (uint64_t s = 0; s < nsub; ++s) {
        const int64_t tt = t + resid[i * nsub + s];
        if (tt >= 0 && tt < static_cast<int64_t>(n))
          acc += inter[(k * nsub + s) * n + tt];
      }
      out[i * no + t] = acc;
    }
  }
  return out;
}

// dedispersion.cpp:217-247: single DM = one-row brute force whose valid
// length uses this DM's own maximum delay.
template <typename T>
std::vector<typename Acc<T>::type> single_dm(const T* x, uint64_t n,
                                             uint64_t nchans,
                                             const double* freq, double dm,
                                             double ref, double tsamp,
                                             uint64_t cb, uint64_t ce,
                                             uint64_t* nout) {
  using A = typename Acc<T>::type;
  require(std::isfinite(dm) && dm >= 0.0, "dm must be finite and non-negative");
  const uint64_t nc = ce - cb;
  std::vector<int32_t> d(nc);
  for (uint64_t c = 0; c < nc; ++c) d[c] = delay_bins(dm, freq[cb + c], ref, tsamp);
  *nout = valid_nout(n, delay_bins(dm, min_freq(freq, cb, ce), ref, tsamp));
  std::vector<A> out(*nout);
  for (uint64_t t = 0; t < *nout; ++t) {
    A s = 0;
    for (uint64_t c = 0; c < nc; ++c) {
      const int64_t tt = static_cast<int64_t>(t) + d[c];
      if (tt >= 0 && tt < static_cast<int64_t>(n))
        s += static_cast<A>(x[tt * nchans + cb + c]);
    }
    out[t] = s;
  }
  return out;
}

// ----------------------------------------------------------- time series ----

void validate_finite(const float* x, uint64_t n, const char* message) {
  for (uint64_t i = 0; i < n; ++i) require(std::isfinite(x[i]), message);
}

struct Stats {
  double mean, variance, stddev;
};

// time_series.cpp:50-81: population statistics, serial FP64 sums.
Stats stats(const float* x, uint64_t n) {
  require(n != 0, "time series stats input must not be empty");
  validate_finite(x, n, "time series samples must be finite");
  double sum = 0.0, sum2 = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    const double v = static_cast<double>(x[i]);
    sum += v;
    sum2 += v * v;
  }
  const double count = static_cast<double>(n);
  const double mean = sum / count;
  double var = sum2 / count - mean * mean;
  if (var < 0.0 && var > -1.0e-12) var = 0.0;
  require(var >= 0.0 && std::isfinite(var),
          "time series variance must be finite and >= 0");
  return Stats{mean, var, std::sqrt(var)};
}

// time_series.cpp:83-104: (x - mean) * (1/stddev) in FP64, rounded to float.
void normalise(const float* x, uint64_t n, bool reject_constant, float* out) {
  require(n != 0, "normalise input must not be empty");
  const Stats s = stats(x, n);
  if (!(s.stddev > 0.0) || !std::isfinite(s.stddev)) {
    require(!reject_constant,
            "normalise input standard deviation must be finite and > 0");
    std::fill(out, out + n, 0.0F);
    return;
  }
  const double scale = 1.0 / s.stddev;
  for (uint64_t i = 0; i < n; ++i)
    out[i] = static_cast<float>((static_cast<double>(x[i]) - s.mean) * scale);
}

void validate_factor(uint64_t n, double factor) {
  require(n != 0, "time series must not be empty");
  require(std::isfinite(factor) && factor > 1.0 &&
              factor <= static_cast<double>(n),
          "downsample factor must satisfy 1 < factor <= nsamples");
}

// time_series.cpp:118-122.
uint64_t downsampled_size(uint64_t n, double factor) {
  validate_factor(n, factor);
  return static_cast<uint64_t>(std::floor(static_cast<double>(n) / factor));
}

// time_series.cpp:124-138: riptide's noise-variance model of the
// fractional weighted-sum downsampler.
double downsampled_variance(uint64_t n, double factor) {
  validate_factor(n, factor);
  const double whole = std::floor(factor);
  const double frac_span =
      static_cast<double>(downsampled_size(n, factor)) * (factor - whole);
  if (frac_span > 1.0) return factor - 1.0 / 3.0;
  return std::pow(whole - 1.0, 2.0) + (2.0 / 3.0) * std::pow(frac_span, 2.0) -
         frac_span + 1.0;
}

// time_series.cpp:140-174: output k integrates [k*f, (k+1)*f) - first and
// last samples weighted by their fractional overlap (weights formed in FP64,
// rounded to float), interior samples by 1, summed serially in FP32.
void downsample(const float* x, uint64_t n, double factor, float* out,
                uint64_t nout) {
  const double last = static_cast<double>(n - 1);
  for (uint64_t k = 0; k < nout; ++k) {
    const double start = static_cast<double>(k) * factor;
    const double end = start + factor;
    const uint64_t first = static_cast<uint64_t>(std::floor(start));
    const uint64_t lastidx =
        static_cast<uint64_t>(std::min(std::floor(end), last));
    const float w0 = static_cast<float>(static_cast<double>(first + 1) - start);
    const float w1 = static_cast<float>(end - static_cast<double>(lastidx));
    float s = w0 * x[first];
    for (uint64_t i = first + 1; i < lastidx; ++i) s += x[i];
    s += w1 * x[lastidx];
    out[k] = s;
  }
}

// ------------------------------------------------------------ preprocess ----

// preprocessing.cpp:53-76: centred median of an edge-value padded window.
void exact_running_median(const float* x, uint64_t n, uint64_t window,
                          float* out) {
  require(window < n,
          "running median detrend window_samples must be < input size");
  const uint64_t half = window / 2;
  std::vector<float> buf(window);
  for (uint64_t i = 0; i < n; ++i) {
    for (uint64_t j = 0; j < window; ++j) {
      const int64_t src = static_cast<int64_t>(i + j) - static_cast<int64_t>(half);
      const int64_t clamped =
          std::min<int64_t>(std::max<int64_t>(src, 0), static_cast<int64_t>(n) - 1);
      buf[j] = x[clamped];
    }
    std::nth_element(buf.begin(), buf.begin() + half, buf.end());
    out[i] = buf[half];
  }
}

// preprocessing.cpp:124-145 (+ 78-92 mean scrunch in FP64, 94-122 FP64
// linear interpolation back to full resolution; positions before the first
// low-resolution centre take the first value, from the last centre onward
// the last value).
void fast_running_median(const float* x, uint64_t n, uint64_t window,
                         uint64_t min_points, float* out) {
  const uint64_t f = std::max<uint64_t>(
      1, static_cast<uint64_t>(static_cast<double>(window) /
                               static_cast<double>(min_points)));
  if (f == 1) {
    exact_running_median(x, n, window, out);
    return;
  }
  const uint64_t m = n / f;
  std::vector<float> low(m);
  for (uint64_t k = 0; k < m; ++k) {
    double s = 0.0;
    for (uint64_t j = 0; j < f; ++j) s += static_cast<double>(x[k * f + j]);
    low[k] = static_cast<float>(s / static_cast<double>(f));
  }
  require(m > min_points,
          "running median detrend low-resolution sample count must be > "
          "min_points");
  std::vector<float> med(m);
  exact_running_median(low.data(), m, min_points, med.data());
  require(!med.empty(),
          "running median detrend low-resolution median must not be empty");
  const double offset = 0.5 * static_cast<double>(f - 1);
  for (uint64_t i = 0; i < n; ++i) {
    const double pos = (static_cast<double>(i) - offset) / static_cast<double>(f);
    if (pos <= 0.0) {
      out[i] = med.front();
      continue;
    }
    const uint64_t lo = static_cast<uint64_t>(std::floor(pos));
    if (lo >= m - 1) {
      out[i] = med.back();
      continue;
    }
    const double fr = pos - static_cast<double>(lo);
    out[i] = static_cast<float>((1.0 - fr) * static_cast<double>(med[lo]) +
                                fr * static_cast<double>(med[lo + 1]));
  }
}

// preprocessing.cpp:8-20 + 212-231.
void detrend(const float* x, uint64_t n, uint64_t window, uint64_t min_points,
             float* out) {
  require(window != 0 && window % 2 == 1,
          "running median detrend window_samples must be odd and > 0");
  require(min_points != 0 && min_points % 2 == 1,
          "running median detrend min_points must be odd and > 0");
  require(n != 0, "running median detrend input must not be empty");
  validate_finite(x, n, "running median detrend input samples must be finite");
  std::vector<float> med(n);
  fast_running_median(x, n, window, min_points, med.data());
  for (uint64_t i = 0; i < n; ++i) out[i] = x[i] - med[i];
}

// preprocessing.cpp:164-188.
uint64_t running_median_window(double seconds, double tsamp) {
  require(seconds > 0.0 && std::isfinite(seconds),
          "running median window_seconds must be finite and > 0");
  require(tsamp > 0.0 && std::isfinite(tsamp),
          "running median tsamp must be finite and > 0");
  const double r = std::round(seconds / tsamp);
  const double max_size =
      static_cast<double>(std::numeric_limits<std::size_t>::max());
  require(std::isfinite(r) && !(r > max_size - 1.0),
          "running median window_samples is too large");
  uint64_t w = static_cast<uint64_t>(std::max(1.0, r));
  if (w % 2 == 0) ++w;
  return w;
}

// preprocessing.cpp:241-257: steps applied in plan order.
void preprocess(const float* x, uint64_t n, double tsamp,
                const orc_preprocess_plan& plan, float* out) {
  require(n != 0, "time series data must not be empty");
  require(tsamp > 0.0 && std::isfinite(tsamp),
          "time series tsamp must be finite and > 0");
  require(plan.detrend || plan.normalise,
          "preprocess plan must contain at least one step");
  std::vector<float> cur(x, x + n), next(n);
  if (plan.detrend) {
    detrend(cur.data(), n, plan.window_samples, plan.min_points, next.data());
    cur.swap(next);
  }
  if (plan.normalise) {
    normalise(cur.data(), n, plan.reject_constant != 0, next.data());
    cur.swap(next);
  }
  std::copy(cur.begin(), cur.end(), out);
}

// -------------------------------------------------------------- FFA plan ----

// ffa_plan.cpp:47-54.
uint64_t ceil_shift(uint64_t rows, uint64_t bins, double pmax_samples) {
  const double v = static_cast<double>(bins) *
                   (static_cast<double>(rows) - 1.0) *
                   (1.0 - static_cast<double>(bins) / pmax_samples);
  return static_cast<uint64_t>(std::ceil(v));
}

// ffa_plan.cpp:56-72: w <- max(w+1, floor(spacing*w)) while w <= max width.
std::vector<uint64_t> width_trials(uint64_t bins_min, double duty,
                                   double spacing) {
  std::vector<uint64_t> w;
  const uint64_t wmax = std::max<uint64_t>(
      1, static_cast<uint64_t>(std::floor(duty * static_cast<double>(bins_min))));
  for (uint64_t x = 1; x <= wmax;) {
    w.push_back(x);
    x = std::max(x + 1, static_cast<uint64_t>(
                            std::floor(spacing * static_cast<double>(x))));
  }
  return w;
}

// ffa_plan.cpp:74-88.
double period_end(double teff, uint64_t bins, uint64_t rows, uint64_t rows_eval) {
  if (rows == 1 || rows_eval == 1) return teff * static_cast<double>(bins);
  const double b = static_cast<double>(bins);
  const double shift =
      static_cast<double>(rows_eval - 1) / static_cast<double>(rows - 1);
  return teff * b * b / (b - shift);
}

// ffa_plan.cpp:90-171: geometric downsampling cycles f0*g^c, one task per
// bins value in [bins_min, min(bins_max, prepared, floor(pmax/teff))].
void make_plan(uint64_t n, double tsamp, const orc_plan_options& o,
               std::vector<orc_task>& tasks, std::vector<uint64_t>& widths) {
  require(n > 0, "FFA plan nsamples must be > 0");
  require(tsamp > 0.0, "FFA plan tsamp must be > 0");
  require(o.period_min > 0.0, "FFA plan period_min must be > 0");
  require(o.period_max > o.period_min, "FFA plan period_max must be > period_min");
  require(o.bins_min > 1, "FFA plan bins_min must be > 1");
  require(o.bins_max >= o.bins_min, "FFA plan bins_max must be >= bins_min");
  require(o.min_periods > 0, "FFA plan min_periods must be > 0");
  require(o.duty_cycle_max > 0.0 && o.duty_cycle_max < 1.0,
          "FFA plan duty_cycle_max must be in (0, 1)");
  require(o.width_trial_spacing > 1.0, "FFA plan width_trial_spacing must be > 1");
  require(o.max_tasks > 0, "FFA plan max_tasks must be > 0");
  require(o.period_min >= tsamp * static_cast<double>(o.bins_min),
          "FFA plan period_min must be >= tsamp * bins_min");
  const double tobs = static_cast<double>(n) * tsamp;
  const double pmax = std::min(o.period_max, tobs / static_cast<double>(o.min_periods));
  if (pmax <= o.period_min)
    throw std::invalid_argument("FFA search range is empty after min_periods cap");
  widths = width_trials(o.bins_min, o.duty_cycle_max, o.width_trial_spacing);
  tasks.clear();
  const double f0 = o.period_min / (tsamp * static_cast<double>(o.bins_min));
  const double g = (static_cast<double>(o.bins_max) + 1.0) /
                   static_cast<double>(o.bins_min);
  const uint64_t cycles = static_cast<uint64_t>(
      std::ceil(std::log(pmax / o.period_min) / std::log(g)));
  for (uint64_t c = 0; c < cycles; ++c) {
    const double f = f0 * std::pow(g, static_cast<double>(c));
    const double teff = f * tsamp;
    const uint64_t prepared = f == 1.0 ? n : downsampled_size(n, f);
    const double pmax_samples = pmax / teff;
    const uint64_t stop = std::min(
        {o.bins_max, prepared, static_cast<uint64_t>(std::floor(pmax_samples))});
    if (stop < o.bins_min) continue;
    for (uint64_t b = o.bins_min; b <= stop; ++b) {
      const uint64_t rows = prepared / b;
      if (rows < o.min_periods) continue;
      const double pceil = std::min(pmax_samples, static_cast<double>(b) + 1.0);
      const uint64_t re = std::min(rows, ceil_shift(rows, b, pceil));
      if (re == 0) continue;
      if (tasks.size() >= o.max_tasks)
        throw std::runtime_error("FFA search plan exceeds max_tasks");
      tasks.push_back(orc_task{f, teff, n, prepared, b, rows, re,
                               teff * static_cast<double>(b),
                               period_end(teff, b, rows, re)});
    }
  }
}

// --------------------------------------------------------- FFA transform ----

// ffa.cpp:133-143: the head/tail row picked for output row s is
// (size_t)(scale*s + 0.5f) with scale = (child_rows-1)/(rows-1), all FP32.
inline uint64_t pick(uint64_t child_rows, uint64_t rows, uint64_t s) {
  const float scale =
      static_cast<float>(child_rows - 1) / static_cast<float>(rows - 1);
  return static_cast<uint64_t>(scale * static_cast<float>(s) + 0.5F);
}

// ffa.cpp:110-150: out[s][i] = head[h][i] + tail[t][(i + s - t) mod bins].
template <typename T>
void merge_row(const T* head, uint64_t hrows, const T* tail, uint64_t trows,
               uint64_t rows, uint64_t bins, uint64_t s, T* out) {
  const uint64_t h = pick(hrows, rows, s);
  const uint64_t t = pick(trows, rows, s);
  const uint64_t roll = (h + (s - (h + t))) % bins;
  const T* hr = head + h * bins;
  const T* tr = tail + t * bins;
  for (uint64_t i = 0; i < bins; ++i) {
    uint64_t j = i + roll;
    if (j >= bins) j -= bins;
    out[i] = hr[i] + tr[j];
  }
}

// ffa.cpp:173-195: head = first floor(R/2) rows, tail = the rest; R == 1 is a
// copy and R == 2 a direct sum / rolled-by-one sum.
template <typename T>
void transform(const T* in, uint64_t rows, uint64_t bins, T* tmp, T* out) {
  if (rows == 1) {
    std::memcpy(out, in, bins * sizeof(T));
    return;
  }
  if (rows == 2) {
    merge_row(in, 1, in + bins, 1, 2, bins, 0, out);
    merge_row(in, 1, in + bins, 1, 2, bins, 1, out + bins);
    return;
  }
  const uint64_t h = rows / 2;
  const uint64_t t = rows - h;
  transform(in, h, bins, out, tmp);
  transform(in + h * bins, t, bins, out + h * bins, tmp + h * bins);
  for (uint64_t s = 0; s < rows; ++s)
    merge_row(tmp, h, tmp + h * bins, t, rows, bins, s, out + s * bins);
}

// ffa.cpp:203-251: only the first rows_eval final rows, each merged from
// the two materialised half transforms.
void emit_rows(const float* in, uint64_t rows, uint64_t bins,
               uint64_t rows_eval, float* out) {
  if (rows <= 2) {
    std::vector<float> full(rows * bins), tmp(rows * bins);
    transform(in, rows, bins, tmp.data(), full.data());
    std::copy(full.begin(), full.begin() + rows_eval * bins, out);
    return;
  }
  const uint64_t h = rows / 2;
  const uint64_t t = rows - h;
  std::vector<float> work(rows * bins), scratch(rows * bins);
  transform(in, h, bins, scratch.data(), work.data());
  transform(in + h * bins, t, bins, scratch.data() + h * bins,
            work.data() + h * bins);
  for (uint64_t s = 0; s < rows_eval; ++s)
    merge_row(work.data(), h, work.data() + h * bins, t, rows, bins, s,
              out + s * bins);
}

// ---------------------------------------------------------- detection ----

struct Trial {
  uint64_t width, index;
  float height, baseline;
  double duty;
};

// ffa_detection.cpp:60-80.
std::vector<Trial> boxcar_trials(const uint64_t* widths, uint64_t nw,
                                 uint64_t bins) {
  require(nw != 0, "FFA detection width trials must not be empty");
  std::vector<Trial> trials;
  const float b = static_cast<float>(bins);
  for (uint64_t i = 0; i < nw; ++i) {
    require(widths[i] != 0 && widths[i] < bins,
            "FFA detection width trials must satisfy 0 < width < bins");
    const float w = static_cast<float>(widths[i]);
    const float h = std::sqrt((b - w) / (b * w));
    trials.push_back(Trial{widths[i], i, h, w / (b - w) * h,
                           static_cast<double>(widths[i]) /
                               static_cast<double>(bins)});
  }
  return trials;
}

// ffa_detection.cpp:256-264.
float task_stdnoise(const orc_task& task) {
  double var = 1.0;
  if (task.downsample_factor != 1.0)
    var = downsampled_variance(task.input_nsamples, task.downsample_factor);
  return static_cast<float>(std::sqrt(static_cast<double>(task.rows) * var));
}

// ffa_detection.cpp:115-124 + 266-284.
orc_peak make_peak(const orc_task& task, const Trial& tr, uint64_t shift,
                   uint64_t phase, float snr) {
  double period;
  if (task.rows <= 1) {
    period = task.effective_tsamp * static_cast<double>(task.bins);
  } else {
    const double b = static_cast<double>(task.bins);
    period = task.effective_tsamp * b * b /
             (b - static_cast<double>(shift) / static_cast<double>(task.rows - 1));
  }
  return orc_peak{period, 1.0 / period, tr.width, tr.duty, tr.index, shift,
                  phase, shift, task.bins, snr};
}

struct Collector {
  std::vector<orc_peak>* peaks;
  uint64_t max_peaks;
  void add(const orc_peak& p) {
    if (max_peaks != 0 && peaks->size() >= max_peaks)
      throw std::runtime_error(
          "FFA detection peak count exceeded max_peaks safety guard");
    peaks->push_back(p);
  }
};

// ffa_detection.cpp:126-239 + 315-372: serial FP32 circular prefix over
// bins + max_width entries; per width the maximum boxcar sum and the first
// phase reaching it; snr = ((h+b)*max - b*sum) * (1/stdnoise).
void detect_row(const float* prof, uint64_t bins, uint64_t shift,
                const orc_task& task, const std::vector<Trial>& trials,
                uint64_t maxw, float stdnoise, float thr,
                std::vector<float>& pre, Collector& col) {
  pre.assign(bins + maxw + 1, 0.0F);
  pre[0] = 0.0F;
  for (uint64_t i = 0; i < bins; ++i) pre[i + 1] = pre[i] + prof[i];
  for (uint64_t i = 0; i < maxw; ++i) pre[bins + i + 1] = pre[bins + i] + prof[i];
  const float total = pre[bins];
  const float inv = 1.0F / stdnoise;
  for (const Trial& tr : trials) {
    float best = -std::numeric_limits<float>::infinity();
    uint64_t phase = 0;
    for (uint64_t p = 0; p < bins; ++p) {
      const float v = pre[p + tr.width] - pre[p];
      if (v > best) {
        best = v;
        phase = p;
      }
    }
    const float snr = ((tr.height + tr.baseline) * best - tr.baseline * total) * inv;
    if (snr >= thr) col.add(make_peak(task, tr, shift, phase, snr));
  }
}

bool finite_peak(const orc_peak& p) {
  return std::isfinite(p.snr) && std::isfinite(p.period) &&
         std::isfinite(p.frequency);
}

// ffa_peak.cpp:16-41.
bool better(const orc_peak& a, const orc_peak& b) {
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

void validate_task(const orc_task& t, uint64_t n) {
  require(t.input_nsamples == n,
          "FFA executor task input_nsamples must match time series size");
  require(std::isfinite(t.downsample_factor) && t.downsample_factor >= 1.0,
          "FFA executor task downsample_factor must be finite and >= 1");
  require(t.effective_tsamp > 0.0 && std::isfinite(t.effective_tsamp),
          "FFA executor task effective_tsamp must be finite and > 0");
  require(t.bins > 1, "FFA executor task bins must be > 1");
  require(t.rows != 0, "FFA executor task rows must be > 0");
  require(t.rows_eval != 0 && t.rows_eval <= t.rows,
          "FFA executor task rows_eval must satisfy 0 < rows_eval <= rows");
  require(t.prepared_nsamples != 0,
          "FFA executor task prepared_nsamples must be > 0");
  if (t.bins != 0 && t.rows > std::numeric_limits<uint64_t>::max() / t.bins)
    throw std::overflow_error("FFA executor shape size overflow");
  require(t.rows * t.bins <= t.prepared_nsamples,
          "FFA executor task rows * bins must be <= prepared_nsamples");
  if (t.downsample_factor == 1.0) {
    require(t.prepared_nsamples == n,
            "FFA executor no-downsample task prepared_nsamples must match "
            "input size");
  } else {
    require(t.prepared_nsamples == downsampled_size(n, t.downsample_factor),
            "FFA executor task prepared_nsamples must match downsampled_size");
  }
}

// ffa_executor.cpp:99-115 + 154-189 and ffa_search.cpp:26-61: per task,
// prepare (cached across consecutive tasks with the same factor), stream the
// rows_eval final rows, detect each row, then sort all peaks.
std::vector<orc_peak> search(const float* x, uint64_t n, const orc_task* tasks,
                             uint64_t ntasks, const uint64_t* widths,
                             uint64_t nw, float thr, uint64_t max_peaks) {
  require(std::isfinite(thr), "FFA search S/N threshold must be finite");
  require(nw != 0, "FFA search plan width_trials must not be empty");
  require(n != 0, "FFA executor time series must not be empty");
  require(ntasks != 0, "FFA executor plan must contain at least one task");
  for (uint64_t i = 0; i < ntasks; ++i) validate_task(tasks[i], n);
  std::vector<orc_peak> peaks;
  Collector col{&peaks, max_peaks};
  std::vector<float> cache, rows, pre;
  double cache_factor = 0.0;
  bool cache_valid = false;
  for (uint64_t k = 0; k < ntasks; ++k) {
    const orc_task& t = tasks[k];
    const float* prepared = x;
    if (t.downsample_factor != 1.0) {
      if (!cache_valid || cache_factor != t.downsample_factor ||
          cache.size() != t.prepared_nsamples) {
        cache.resize(t.prepared_nsamples);
        downsample(x, n, t.downsample_factor, cache.data(), cache.size());
        cache_factor = t.downsample_factor;
        cache_valid = true;
      }
      prepared = cache.data();
    }
    rows.resize(t.rows_eval * t.bins);
    emit_rows(prepared, t.rows, t.bins, t.rows_eval, rows.data());
    const std::vector<Trial> trials = boxcar_trials(widths, nw, t.bins);
    uint64_t maxw = 0;
    for (const Trial& tr : trials) maxw = std::max(maxw, tr.width);
    const float sn = task_stdnoise(t);
    require(sn > 0.0F && std::isfinite(sn),
            "FFA row detection stdnoise must be finite and > 0");
    for (uint64_t s = 0; s < t.rows_eval; ++s)
      detect_row(rows.data() + s * t.bins, t.bins, s, t, trials, maxw, sn, thr,
                 pre, col);
  }
  std::sort(peaks.begin(), peaks.end(), better);
  return peaks;
}

// ------------------------------------------------------- synthetic data ----

// Seeded counter-based generator (no reference counterpart).  Sample (t, c)
// is pop - 384 + floor((dither + signal_q + rfi_q) / 4096) clipped to
// [0, 255], where pop is the popcount of 1024 hashed bits (Binomial(1024,1/2):
// mean 512, sigma 16), dither is 12 hashed bits (uniform, so the sub-unit
// signal survives rounding on average), and signal_q is a 1/4096-unit
// integer pulse-profile table entry picked by the exact integer phase
// ((t - d_c) * 2^20 - phase0_q) mod P_q of each pulsar.  Everything past the
// host-built tables is integer arithmetic, so CPU and GPU emit equal bytes.
inline uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

constexpr int kPhaseBins = 4096;
constexpr int64_t kPeriodFrac = 1LL << 20;

struct PulsarTables {
  std::vector<int32_t> delay;   // [p][c]
  std::vector<int32_t> profile; // [p][kPhaseBins]
  std::vector<int64_t> period_q, phase0_q;
};

PulsarTables pulsar_tables(const orc_pulsar* ps, uint64_t np,
                           const std::vector<double>& freq, double tsamp) {
  PulsarTables tb;
  const uint64_t nc = freq.size();
  const double ref = *std::max_element(freq.begin(), freq.end());
  tb.delay.resize(np * nc);
  tb.profile.resize(np * kPhaseBins);
  for (uint64_t p = 0; p < np; ++p) {
    require(ps[p].period_s > 0.0 && ps[p].duty > 0.0 && ps[p].duty < 1.0,
            "pulsar period and duty must be positive, duty < 1");
    for (uint64_t c = 0; c < nc; ++c)
      tb.delay[p * nc + c] = delay_bins(ps[p].dm, freq[c], ref, tsamp);
    const double sigma = ps[p].duty / (2.0 * std::sqrt(2.0 * std::log(2.0)));
    for (int b = 0; b < kPhaseBins; ++b) {
      double x = (static_cast<double>(b) + 0.5) / kPhaseBins;
      if (x > 0.5) x = 1.0 - x;
      const double g = std::exp(-0.5 * (x / sigma) * (x / sigma));
      tb.profile[p * kPhaseBins + b] =
          static_cast<int32_t>(std::llround(4096.0 * 16.0 * ps[p].amp_sigma * g));
    }
    const int64_t pq = std::llround(ps[p].period_s / tsamp * kPeriodFrac);
    require(pq > 0, "pulsar period too short for the sample time");
    tb.period_q.push_back(pq);
    double ph = ps[p].phase0 - std::floor(ps[p].phase0);
    tb.phase0_q.push_back(std::llround(ph * static_cast<double>(pq)) % pq);
  }
  return tb;
}

}  // namespace

// ================================================================ C ABI ====

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }
const char* orc_implementation(void) { return "port"; }

int orc_delay_bins(double dm, double f, double ref, double tsamp, int32_t* out) {
  return guarded([&] { *out = delay_bins(dm, f, ref, tsamp); });
}

int orc_multi_dm_delay_table(const double* freq, uint64_t nchans, double dm_low,
                             double dm_step, uint64_t ndm, double ref,
                             double tsamp, uint64_t cb, uint64_t ce,
                             int32_t* out, uint64_t capacity, uint64_t* count) {
  return guarded([&] {
    require(cb < ce && ce <= nchans, "channel range exceeds sample channel count");
    std::vector<int32_t> d;
    for (uint64_t i = 0; i < ndm; ++i) {
      const double dm = dm_low + static_cast<double>(i) * dm_step;
      for (uint64_t c = cb; c < ce; ++c) d.push_back(delay_bins(dm, freq[c], ref, tsamp));
    }
    emit(d, out, capacity, count);
  });
}

int orc_dedisperse(int32_t kind, int32_t dtype, const void* samples,
                   uint64_t nsamples, uint64_t nchans, const double* freq,
                   double dm_low, double dm_step, uint64_t ndm, double ref,
                   double tsamp, uint64_t cb, uint64_t ce, uint64_t sbc,
                   uint64_t npn, void* out, uint64_t capacity,
                   uint64_t* out_nsamples) {
  return guarded([&] {
    validate_common(samples, nsamples, nchans, freq, ref, tsamp, cb, ce);
    const DmPlan p{dm_low, dm_step, ndm, ref, tsamp, cb, ce};
    validate_multi(p);
    auto run = [&](auto* x) {
      uint64_t no = 0;
      auto r = kind == 0 ? brute_force(x, nsamples, nchans, freq, p, &no)
                         : subband(x, nsamples, nchans, freq, p, sbc, npn, &no);
      *out_nsamples = no;
      if (out == nullptr) return;
      if (capacity < r.size()) throw too_small();
      std::memcpy(out, r.data(), r.size() * sizeof(r[0]));
    };
    if (dtype == 0) run(static_cast<const uint8_t*>(samples));
    else if (dtype == 1) run(static_cast<const uint16_t*>(samples));
    else if (dtype == 2) run(static_cast<const float*>(samples));
    else throw std::invalid_argument("unsupported sample dtype");
  });
}

int orc_dedisperse_single(int32_t dtype, const void* samples, uint64_t nsamples,
                          uint64_t nchans, const double* freq, double dm,
                          double ref, double tsamp, uint64_t cb, uint64_t ce,
                          void* out, uint64_t capacity, uint64_t* out_nsamples) {
  return guarded([&] {
    validate_common(samples, nsamples, nchans, freq, ref, tsamp, cb, ce);
    auto run = [&](auto* x) {
      uint64_t no = 0;
      auto r = single_dm(x, nsamples, nchans, freq, dm, ref, tsamp, cb, ce, &no);
      *out_nsamples = no;
      if (out == nullptr) return;
      if (capacity < r.size()) throw too_small();
      std::memcpy(out, r.data(), r.size() * sizeof(r[0]));
    };
    if (dtype == 0) run(static_cast<const uint8_t*>(samples));
    else if (dtype == 1) run(static_cast<const uint16_t*>(samples));
    else if (dtype == 2) run(static_cast<const float*>(samples));
    else throw std::invalid_argument("unsupported sample dtype");
  });
}

// dedispersion.cpp:249-289: each channel read at its own delay, valid range
// only (every read is in range there).
int orc_dedisperse_spectrum(int32_t dtype, const void* samples, uint64_t nsamples,
                            uint64_t nchans, const double* freq, double dm, double ref,
                            double tsamp, uint64_t cb, uint64_t ce, void* out,
                            uint64_t capacity, uint64_t* out_nsamples) {
  return guarded([&] {
    validate_common(samples, nsamples, nchans, freq, ref, tsamp, cb, ce);
    require(std::isfinite(dm) && dm >= 0.0, "dm must be finite and non-negative");
    const uint64_t nc = ce - cb;
    std::vector<int32_t> d(nc);
    for (uint64_t c = 0; c < nc; ++c) d[c] = delay_bins(dm, freq[cb + c], ref, tsamp);
    const uint64_t no = valid_nout(nsamples, delay_bins(dm, min_freq(freq, cb, ce), ref, tsamp));
    *out_nsamples = no;
    if (out == nullptr) return;
    if (capacity < no * nc) throw too_small();
    auto run = [&](const auto* x) {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(x)>>;
      T* o = static_cast<T*>(out);
      for (uint64_t t = 0; t < no; ++t)
        for (uint64_t c = 0; c < nc; ++c) o[t * nc + c] = x[(t + d[c]) * nchans + cb + c];
    };
    if (dtype == 0) run(static_cast<const uint8_t*>(samples));
    else if (dtype == 1) run(static_cast<const uint16_t*>(samples));
    else if (dtype == 2) run(static_cast<const float*>(samples));
    else throw std::invalid_argument("unsupported sample dtype");
  });
}

int orc_time_series_stats(const float* x, uint64_t n, double* mean,
                          double* variance, double* stddev) {
  return guarded([&] {
    const Stats s = stats(x, n);
    *mean = s.mean;
    *variance = s.variance;
    *stddev = s.stddev;
  });
}

int orc_normalise(const float* x, uint64_t n, int32_t reject, float* out) {
  return guarded([&] { normalise(x, n, reject != 0, out); });
}

int orc_downsampled_size(uint64_t n, double f, uint64_t* out) {
  return guarded([&] { *out = downsampled_size(n, f); });
}

int orc_downsampled_variance(uint64_t n, double f, double* out) {
  return guarded([&] { *out = downsampled_variance(n, f); });
}

int orc_downsample(const float* x, uint64_t n, double f, float* out,
                   uint64_t capacity, uint64_t* count) {
  return guarded([&] {
    const uint64_t m = downsampled_size(n, f);
    *count = m;
    if (out == nullptr) return;
    if (capacity < m) throw too_small();
    downsample(x, n, f, out, m);
  });
}

int orc_running_median_window(double seconds, double tsamp, uint64_t* w) {
  return guarded([&] { *w = running_median_window(seconds, tsamp); });
}

int orc_detrend(const float* x, uint64_t n, uint64_t window, uint64_t min_points,
                float* out) {
  return guarded([&] { detrend(x, n, window, min_points, out); });
}

int orc_preprocess(const float* x, uint64_t n, double tsamp,
                   const orc_preprocess_plan* plan, float* out) {
  return guarded([&] { preprocess(x, n, tsamp, *plan, out); });
}

int orc_make_riptide_plan(uint64_t n, double tsamp, const orc_plan_options* o,
                          orc_task* tasks, uint64_t tcap, uint64_t* tcount,
                          uint64_t* widths, uint64_t wcap, uint64_t* wcount) {
  return guarded([&] {
    std::vector<orc_task> t;
    std::vector<uint64_t> w;
    make_plan(n, tsamp, *o, t, w);
    *tcount = t.size();
    *wcount = w.size();
    if (tasks == nullptr && widths == nullptr) return;
    if (tcap < t.size() || wcap < w.size()) throw too_small();
    std::copy(t.begin(), t.end(), tasks);
    std::copy(w.begin(), w.end(), widths);
  });
}

int orc_ffa_transform(const float* in, uint64_t rows, uint64_t bins, float* out) {
  return guarded([&] {
    require(rows != 0, "FFA transform rows must be > 0");
    require(bins > 1, "FFA transform bins must be > 1");
    std::vector<float> tmp(rows * bins);
    transform(in, rows, bins, tmp.data(), out);
  });
}

int orc_ffa_transform_f64(const double* in, uint64_t rows, uint64_t bins,
                          double* out) {
  return guarded([&] {
    require(rows != 0, "FFA transform rows must be > 0");
    require(bins > 1, "FFA transform bins must be > 1");
    std::vector<double> tmp(rows * bins);
    transform(in, rows, bins, tmp.data(), out);
  });
}

int orc_ffa_rows(const float* x, uint64_t n, const orc_task* tasks,
                 uint64_t ntasks, float* out, uint64_t capacity,
                 uint64_t* count) {
  return guarded([&] {
    uint64_t total = 0;
    for (uint64_t k = 0; k < ntasks; ++k) {
      validate_task(tasks[k], n);
      total += tasks[k].rows_eval * tasks[k].bins;
    }
    *count = total;
    if (out == nullptr) return;
    if (capacity < total) throw too_small();
    std::vector<float> cache;
    double cf = 0.0;
    uint64_t off = 0;
    for (uint64_t k = 0; k < ntasks; ++k) {
      const orc_task& t = tasks[k];
      const float* prep = x;
      if (t.downsample_factor != 1.0) {
        if (cf != t.downsample_factor || cache.size() != t.prepared_nsamples) {
          cache.resize(t.prepared_nsamples);
          downsample(x, n, t.downsample_factor, cache.data(), cache.size());
          cf = t.downsample_factor;
        }
        prep = cache.data();
      }
      emit_rows(prep, t.rows, t.bins, t.rows_eval, out + off);
      off += t.rows_eval * t.bins;
    }
  });
}

int orc_task_stdnoise(const orc_task* t, float* out) {
  return guarded([&] { *out = task_stdnoise(*t); });
}

int orc_find_ffa_peaks(const float* tr, uint64_t rows, uint64_t bins,
                       const orc_task* task, const uint64_t* widths,
                       uint64_t nw, float stdnoise, float thr,
                       uint64_t max_peaks, orc_peak* out, uint64_t capacity,
                       uint64_t* count) {
  return guarded([&] {
    require(rows != 0, "FFA detection rows must be > 0");
    require(bins > 1, "FFA detection bins must be > 1");
    require(task->bins == bins, "FFA detection task bins must match shape");
    require(task->rows_eval == rows, "FFA detection task rows_eval must match shape");
    require(task->rows >= task->rows_eval && task->rows != 0,
            "FFA detection task rows must be >= rows_eval and > 0");
    require(task->effective_tsamp > 0.0 && std::isfinite(task->effective_tsamp),
            "FFA detection task effective_tsamp must be finite and > 0");
    require(stdnoise > 0.0F && std::isfinite(stdnoise),
            "FFA detection stdnoise must be finite and > 0");
    require(std::isfinite(thr), "FFA detection S/N threshold must be finite");
    const std::vector<Trial> trials = boxcar_trials(widths, nw, bins);
    uint64_t maxw = 0;
    for (const Trial& t : trials) maxw = std::max(maxw, t.width);
    std::vector<orc_peak> peaks;
    Collector col{&peaks, max_peaks};
    std::vector<float> pre;
    for (uint64_t s = 0; s < rows; ++s)
      detect_row(tr + s * bins, bins, s, *task, trials, maxw, stdnoise, thr,
                 pre, col);
    std::sort(peaks.begin(), peaks.end(), better);
    emit(peaks, out, capacity, count);
  });
}

int orc_search_ffa(const float* x, uint64_t n, const orc_task* tasks,
                   uint64_t ntasks, const uint64_t* widths, uint64_t nw,
                   float thr, uint64_t max_peaks, orc_peak* out,
                   uint64_t capacity, uint64_t* count) {
  return guarded([&] {
    emit(search(x, n, tasks, ntasks, widths, nw, thr, max_peaks), out, capacity,
         count);
  });
}

int orc_sort_peaks(orc_peak* peaks, uint64_t count) {
  return guarded([&] { std::sort(peaks, peaks + count, better); });
}

// dm_search.cpp:119-199: per DM (OpenMP dynamic when ndm > 4): uint32 ->
// float, preprocess, FFA search, tag; then a global sort by peak order,
// dm_index, dm.
int orc_find_dm_peaks(const uint32_t* plane, uint64_t ndm, uint64_t ns,
                      const double* dms, double tsamp,
                      const orc_plan_options* po,
                      const orc_preprocess_plan* pp, float thr,
                      uint64_t max_peaks, orc_dm_peak* out, uint64_t capacity,
                      uint64_t* count) {
  return guarded([&] {
    require(ndm != 0, "DM search dedispersed ndm must be > 0");
    require(ns != 0, "DM search dedispersed nsamples must be > 0");
    require(tsamp > 0.0 && std::isfinite(tsamp),
            "DM search tsamp must be finite and > 0");
    for (uint64_t i = 0; i < ndm; ++i)
      require(std::isfinite(dms[i]), "DM search dms must be finite");
    require(std::isfinite(thr), "DM search S/N threshold must be finite");
    std::vector<orc_task> tasks;
    std::vector<uint64_t> widths;
    make_plan(ns, tsamp, *po, tasks, widths);
    std::vector<orc_dm_peak> all;
    std::exception_ptr err;
    std::atomic<bool> failed{false};
#pragma omp parallel if (ndm > 4)
    {
      std::vector<orc_dm_peak> local;
      std::vector<float> ts(ns), pre(ns);
#pragma omp for schedule(dynamic, 1)
      for (int64_t i = 0; i < static_cast<int64_t>(ndm); ++i) {
        if (failed.load()) continue;
        try {
          for (uint64_t j = 0; j < ns; ++j)
            ts[j] = static_cast<float>(plane[i * ns + j]);
          const float* series = ts.data();
          if (pp->detrend || pp->normalise) {
            preprocess(ts.data(), ns, tsamp, *pp, pre.data());
            series = pre.data();
          }
          for (const orc_peak& p :
               search(series, ns, tasks.data(), tasks.size(), widths.data(),
                      widths.size(), thr, max_peaks))
            local.push_back(orc_dm_peak{dms[i], static_cast<uint64_t>(i), p});
        } catch (...) {
          bool expected = false;
          if (failed.compare_exchange_strong(expected, true)) {
#pragma omp critical(orc_dm_error)
            err = std::current_exception();
          }
        }
      }
#pragma omp critical(orc_dm_peaks)
      all.insert(all.end(), local.begin(), local.end());
    }
    if (err) std::rethrow_exception(err);
    std::sort(all.begin(), all.end(), [](const orc_dm_peak& a, const orc_dm_peak& b) {
      if (better(a.peak, b.peak)) return true;
      if (better(b.peak, a.peak)) return false;
      if (a.dm_index != b.dm_index) return a.dm_index < b.dm_index;
      return a.dm < b.dm;
    });
    emit(all, out, capacity, count);
  });
}

int orc_synth_filterbank(uint64_t nsamples, uint64_t nchans, double fch1,
                         double foff, double tsamp, uint64_t seed,
                         const orc_pulsar* ps, uint64_t np,
                         const uint64_t* rfi, uint64_t nrfi, int32_t rfi_offset,
                         uint64_t t_begin, uint64_t t_count, uint8_t* out) {
  return guarded([&] {
    require(nchans != 0 && nsamples != 0, "empty filterbank");
    require(t_begin + t_count <= nsamples, "time range exceeds nsamples");
    std::vector<double> freq(nchans);
    for (uint64_t c = 0; c < nchans; ++c)
      freq[c] = foff < 0.0 ? fch1 + static_cast<double>(nchans - 1 - c) * foff
                           : fch1 + static_cast<double>(c) * foff;
    const PulsarTables tb = pulsar_tables(ps, np, freq, tsamp);
    std::vector<int64_t> rfi_q(nchans, 0);
    for (uint64_t i = 0; i < nrfi; ++i) {
      require(rfi[i] < nchans, "rfi channel out of range");
      rfi_q[rfi[i]] = static_cast<int64_t>(rfi_offset) * 4096;
    }
    const uint64_t key = splitmix(seed);
#pragma omp parallel for schedule(static)
    for (int64_t tt = 0; tt < static_cast<int64_t>(t_count); ++tt) {
      const uint64_t t = t_begin + static_cast<uint64_t>(tt);
      for (uint64_t c = 0; c < nchans; ++c) {
        const uint64_t k = (t * nchans + c) * 17;
        int64_t pop = 0;
        for (uint64_t j = 0; j < 16; ++j)
          pop += __builtin_popcountll(splitmix(key + k + j));
        int64_t acc = static_cast<int64_t>(splitmix(key + k + 16) & 4095u);
        acc += rfi_q[c];
        for (uint64_t p = 0; p < np; ++p) {
          const int64_t tau = static_cast<int64_t>(t) - tb.delay[p * nchans + c];
          const int64_t pq = tb.period_q[p];
          int64_t ph = (tau * kPeriodFrac - tb.phase0_q[p]) % pq;
          if (ph < 0) ph += pq;
          const int64_t bin = ph * kPhaseBins / pq;
          acc += tb.profile[p * kPhaseBins + bin];
        }
        // floor division by 4096 (acc may be negative with strong RFI cuts)
        const int64_t q = (acc >= 0) ? acc / 4096 : -((-acc + 4095) / 4096);
        const int64_t v = pop - 384 + q;
        out[static_cast<uint64_t>(tt) * nchans + c] =
            static_cast<uint8_t>(std::min<int64_t>(255, std::max<int64_t>(0, v)));
      }
    }
  });
}

}  // extern "C"

// Candidate selection and harmonic flagging are host code in the product
// itself; they are checked against the reference build, not restated here.
int orc_select_candidates(const orc_dm_peak*, uint64_t, double, double, uint64_t, int32_t,
                          uint64_t, orc_candidate*, uint64_t, uint64_t*) {
  g_error = "candidate selection is checked against the reference build";
  return ORC_OTHER;
}

int orc_flag_harmonics(const orc_candidate*, uint64_t, double, double, double,
                       const orc_harmonic_options*, orc_candidate*, orc_harmonic*, uint64_t,
                       uint64_t*) {
  g_error = "harmonic flagging is checked against the reference build";
  return ORC_OTHER;
}

// Folding on device data is checked against the reference build.
int orc_fold_time_series(const void*, int32_t, uint64_t, double, double, uint64_t, uint64_t,
                         float*, double*) {
  g_error = "folding is checked against the reference build";
  return ORC_OTHER;
}

int orc_fold_spectrum(const void*, int32_t, uint64_t, uint64_t, double, double, uint64_t,
                      uint64_t, uint64_t, uint64_t, float*, double*, uint64_t, uint64_t*) {
  g_error = "folding is checked against the reference build";
  return ORC_OTHER;
}
