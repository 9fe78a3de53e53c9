// Host side of GPU dedispersion: validation (same checks and messages as
// dedispersion.cpp:53-139 in gaffa), host-built delay tables, and the calls
// into the C ABI.
#include "gaffa/dedispersion.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace gaffa {
namespace {

constexpr std::size_t kRowAlign = 64;  // device rows padded to 256 bytes

std::size_t padded_stride(std::size_t n) { return (n + kRowAlign - 1) / kRowAlign * kRowAlign; }

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

void validate_samples(const SampleShape& shape, bool has_data,
                      std::span<const double> frequency_mhz) {
  check_arg(has_data, "dedispersion samples data must not be null");
  check_arg(shape.nifs == 1, "dedispersion requires nifs == 1");
  check_arg(shape.nsamples != 0, "dedispersion requires at least one sample");
  check_arg(shape.nchans != 0, "dedispersion requires at least one channel");
  check_arg(frequency_mhz.size() == shape.nchans,
            "frequency table length must match sample channel count");
  for (double f : frequency_mhz) {
    check_arg(std::isfinite(f) && f > 0.0, "frequency values must be finite and positive");
  }
}

void validate_plan_values(double ref, double tsamp) {
  check_arg(std::isfinite(ref) && ref > 0.0, "reference frequency must be positive");
  check_arg(std::isfinite(tsamp) && tsamp > 0.0, "tsamp must be positive");
}

void validate_channels(std::span<const double> f, double ref, std::size_t cb, std::size_t ce,
                       std::size_t nchans) {
  check_arg(cb < ce, "channel range must be non-empty");
  check_arg(ce <= nchans, "channel range exceeds sample channel count");
  for (std::size_t c = cb; c < ce; ++c) {
    check_arg(!(f[c] > ref),
              "dedispersion requires selected frequencies <= reference frequency");
  }
}

void validate_multi(const SampleShape& shape, bool has_data, std::span<const double> f,
                    const MultiDmDedispersionPlan& p) {
  validate_samples(shape, has_data, f);
  validate_plan_values(p.ref_frequency_mhz, p.tsamp);
  check_arg(std::isfinite(p.dm_low) && p.dm_low >= 0.0,
            "dm_low must be finite and non-negative");
  check_arg(p.ndm != 0, "dedispersion requires at least one DM");
  check_arg(std::isfinite(p.dm_step) && p.dm_step > 0.0, "dm_step must be positive");
  validate_channels(f, p.ref_frequency_mhz, p.chan_begin, p.chan_end, shape.nchans);
}

double min_selected(std::span<const double> f, std::size_t cb, std::size_t ce) {
  return *std::min_element(f.begin() + static_cast<std::ptrdiff_t>(cb),
                           f.begin() + static_cast<std::ptrdiff_t>(ce));
}

std::size_t valid_output(std::size_t n, std::int32_t max_delay) {
  check_arg(max_delay >= 0, "dedispersion requires non-negative delays");
  check_arg(static_cast<std::size_t>(max_delay) < n,
            "dedispersion valid output range is empty");
  return n - static_cast<std::size_t>(max_delay);
}

// max - min over every (16 DM x 32 channel) block of a [ndm][nc] table: the
// LDS slab height the brute-force kernel needs.
std::int32_t block_spread(const std::vector<std::int32_t>& d, std::size_t ndm, std::size_t nc) {
  std::int32_t spread = 0;
  for (std::size_t i0 = 0; i0 < ndm; i0 += 16) {
    for (std::size_t c0 = 0; c0 < nc; c0 += 32) {
      std::int32_t lo = std::numeric_limits<std::int32_t>::max(), hi = 0;
      for (std::size_t i = i0; i < std::min(ndm, i0 + 16); ++i)
        for (std::size_t c = c0; c < std::min(nc, c0 + 32); ++c) {
          lo = std::min(lo, d[i * nc + c]);
          hi = std::max(hi, d[i * nc + c]);
        }
      spread = std::max(spread, hi - lo);
    }
  }
  return spread;
}

}  // namespace

std::int32_t dedispersion_delay_bins(double dm, double frequency_mhz, double ref_frequency_mhz,
                                     double tsamp) {
  const double f2 = frequency_mhz * frequency_mhz;
  const double r2 = ref_frequency_mhz * ref_frequency_mhz;
  const double delay = 4148.808 * dm * (1.0 / f2 - 1.0 / r2);
  const double bins = std::round(delay / tsamp);
  if (bins < static_cast<double>(std::numeric_limits<std::int32_t>::min()) ||
      bins > static_cast<double>(std::numeric_limits<std::int32_t>::max())) {
    throw std::overflow_error("dedispersion delay exceeds int32 range");
  }
  return static_cast<std::int32_t>(bins);
}

std::vector<std::int32_t> make_multi_dm_delay_table(std::span<const double> f,
                                                    const MultiDmDedispersionPlan& p) {
  const std::size_t nc = p.chan_end - p.chan_begin;
  if (p.ndm != 0 && nc > std::numeric_limits<std::size_t>::max() / p.ndm) {
    throw std::overflow_error("dedispersion delay table size overflow");
  }
  std::vector<std::int32_t> d(p.ndm * nc);
  for (std::size_t i = 0; i < p.ndm; ++i) {
    const double dm = p.dm_low + static_cast<double>(i) * p.dm_step;
    for (std::size_t c = 0; c < nc; ++c)
      d[i * nc + c] = dedispersion_delay_bins(dm, f[p.chan_begin + c], p.ref_frequency_mhz, p.tsamp);
  }
  return d;
}

std::size_t multi_dm_output_nsamples(std::size_t n, std::span<const double> f,
                                     const MultiDmDedispersionPlan& p) {
  const double dm_max = p.dm_low + static_cast<double>(p.ndm - 1) * p.dm_step;
  return valid_output(n, dedispersion_delay_bins(dm_max, min_selected(f, p.chan_begin, p.chan_end),
                                                 p.ref_frequency_mhz, p.tsamp));
}

SubbandTables make_subband_tables(std::span<const double> f, const MultiDmDedispersionPlan& p,
                                  const SubbandDedispersionOptions& o) {
  check_arg(o.subband_channels != 0, "subband_channels must be positive");
  check_arg(o.ndm_per_nominal != 0, "ndm_per_nominal must be positive");
  SubbandTables t;
  const std::size_t nc = p.chan_end - p.chan_begin;
  t.nsub = (nc + o.subband_channels - 1) / o.subband_channels;
  t.n_nominal = (p.ndm + o.ndm_per_nominal - 1) / o.ndm_per_nominal;
  std::vector<double> sub_freq(t.nsub);
  for (std::size_t s = 0; s < t.nsub; ++s) {
    const std::size_t b = p.chan_begin + s * o.subband_channels;
    const std::size_t e = std::min(b + o.subband_channels, p.chan_end);
    sub_freq[s] = f[(b + e - 1) / 2];  // lower-middle channel (dedispersion.cpp:376)
  }
  t.coarse.resize(t.n_nominal * nc);
  for (std::size_t k = 0; k < t.n_nominal; ++k) {
    const double nominal =
        p.dm_low + static_cast<double>(k * o.ndm_per_nominal) * p.dm_step;
    for (std::size_t c = 0; c < nc; ++c)
      t.coarse[k * nc + c] =
          dedispersion_delay_bins(nominal, f[p.chan_begin + c], p.ref_frequency_mhz, p.tsamp);
    for (std::size_t c = 0; c < nc; ++c) t.max_coarse = std::max(t.max_coarse, t.coarse[k * nc + c]);
  }
  t.residual.resize(p.ndm * t.nsub);
  for (std::size_t i = 0; i < p.ndm; ++i) {
    const std::size_t k = i / o.ndm_per_nominal;
    const double nominal = p.dm_low + static_cast<double>(k * o.ndm_per_nominal) * p.dm_step;
    const double dm = p.dm_low + static_cast<double>(i) * p.dm_step;
    for (std::size_t s = 0; s < t.nsub; ++s) {
      const std::int32_t r =
          dedispersion_delay_bins(dm - nominal, sub_freq[s], p.ref_frequency_mhz, p.tsamp);
      check_arg(r >= 0, "dedispersion requires non-negative delays");
      t.residual[i * t.nsub + s] = r;
      t.max_residual = std::max(t.max_residual, r);
    }
  }
  for (std::size_t k = 0; k < t.n_nominal; ++k) {
    const std::size_t i1 = std::min(p.ndm, (k + 1) * o.ndm_per_nominal);
    std::int64_t window = 0;
    for (std::size_t s = 0; s < t.nsub; ++s) {
      std::int32_t h = 0;
      for (std::size_t i = k * o.ndm_per_nominal; i < i1; ++i)
        h = std::max(h, t.residual[i * t.nsub + s]);
      window += (static_cast<std::int64_t>(h) + 7) / 8 * 8;
    }
    t.residual_window = std::max(t.residual_window, window);
  }
  return t;
}

void dedisperse_multi_dm_gpu_into(GpuSampleView samples, std::span<const double> f,
                                  const MultiDmDedispersionPlan& p, std::uint32_t* out,
                                  std::size_t out_stride, const GpuDedispersionOptions& o,
    std::size_t out_nsamples) {
  validate_multi(samples.shape, samples.data != nullptr, f, p);
  std::size_t n_out = multi_dm_output_nsamples(samples.shape.nsamples, f, p);
  if (out_nsamples != 0) {
    check_arg(out_nsamples <= n_out, "requested output length exceeds the valid range");
    n_out = out_nsamples;
  }
  check_arg(out_stride >= n_out, "output stride must be >= out_nsamples");
  const std::vector<std::int32_t> delays = make_multi_dm_delay_table(f, p);
  for (std::int32_t d : delays) check_arg(d >= 0, "dedispersion requires non-negative delays");
  GpuDeviceScope scope(o.device_id);
  GpuDeviceBuffer<std::int32_t> d_delays = upload(delays.data(), delays.size());
  gaffa_hip_dedisperse_multi_args a{};
  a.samples = samples.data;
  a.nsamples = samples.shape.nsamples;
  a.nchans = samples.shape.nchans;
  a.chan_begin = p.chan_begin;
  a.chan_end = p.chan_end;
  a.delays = d_delays.data();
  a.ndm = p.ndm;
  a.out = out;
  a.out_nsamples = n_out;
  a.out_stride = out_stride;
  a.max_block_spread = block_spread(delays, p.ndm, p.chan_end - p.chan_begin);
  hip_ok(gaffa_hip_dedisperse_multi_u8(&a, o.stream), "dedisperse_multi_u8");
  hip_ok(gaffa_hip_dedisperse_check(o.stream), "dedisperse_multi_u8");
}

GpuSegmentTables make_gpu_segment_tables(std::span<const double> f,
                                         const MultiDmDedispersionPlan& p, bool subband,
                                         const SubbandDedispersionOptions& so) {
  GpuSegmentTables t;
  t.subband = subband;
  t.ndm = p.ndm;
  t.chan_begin = p.chan_begin;
  t.chan_end = p.chan_end;
  t.subband_channels = so.subband_channels;
  t.ndm_per_nominal = so.ndm_per_nominal;
  if (subband) {
    const SubbandTables st = make_subband_tables(f, p, so);
    const std::size_t nsel = p.chan_end - p.chan_begin;
    if (gaffa_hip_subband_time_tile(nsel, so.subband_channels,
                                    std::min(so.ndm_per_nominal, p.ndm), st.max_residual,
                                    st.residual_window) == 0)
      throw std::overflow_error("subband dedispersion delays too large for LDS");
    for (std::int32_t d : st.coarse) check_arg(d >= 0, "dedispersion requires non-negative delays");
    t.first = upload(st.coarse.data(), st.coarse.size());
    t.second = upload(st.residual.data(), st.residual.size());
    t.host_first = st.coarse;
    t.host_second = st.residual;
    t.n_nominal = st.n_nominal;
    t.bound_first = st.max_coarse;
    t.bound_second = st.max_residual;
    t.residual_window = st.residual_window;
    // stage 1 reads up to one chunk (<= 512) + residual + pad past the
    // coarse delay
    t.lookahead = static_cast<std::size_t>(st.max_coarse) + st.max_residual + 2048 + 64;
  } else {
    const std::vector<std::int32_t> delays = make_multi_dm_delay_table(f, p);
    for (std::int32_t d : delays) check_arg(d >= 0, "dedispersion requires non-negative delays");
    t.first = upload(delays.data(), delays.size());
    t.bound_first = block_spread(delays, p.ndm, p.chan_end - p.chan_begin);
    std::int32_t dmax = 0;
    for (std::int32_t d : delays) dmax = std::max(dmax, d);
    t.lookahead = static_cast<std::size_t>(dmax) + 512;  // slab of 256 + spread + pad
  }
  return t;
}

void enqueue_segment_range(const GpuSegmentTables& t, const std::uint8_t* samples,
                           const GpuChannelMajorView& cm, std::size_t nsamples,
                           std::size_t nchans, std::uint32_t* out, std::size_t out_stride,
                           std::size_t begin, std::size_t end, gaffa_hip_stream stream) {
  check_arg(begin < end && end <= nsamples, "dedispersion output range is empty");
  if (t.subband) {
    check_arg(begin % 4 == 0, "subband output ranges must start on a multiple of 4 samples");
    gaffa_hip_dedisperse_subband_args a{};
    a.channels = cm.data + (t.chan_begin - cm.chan_begin) * cm.stride + begin;
    a.channel_stride = cm.stride;
    a.channel_length = cm.stride - begin;
    a.nsamples = nsamples - begin;
    a.nsel = t.chan_end - t.chan_begin;
    a.subband_channels = t.subband_channels;
    a.ndm_per_nominal = t.ndm_per_nominal;
    a.ndm = t.ndm;
    a.coarse_delays = t.first.data();
    a.residual_delays = t.second.data();
    a.out = out + begin;
    a.out_nsamples = end - begin;
    a.out_stride = out_stride;
    a.max_residual = t.bound_second;
    a.max_coarse = t.bound_first;
    a.residual_window = t.residual_window;
    hip_ok(gaffa_hip_dedisperse_subband_u8(&a, stream), "dedisperse_subband_u8");
  } else {
    gaffa_hip_dedisperse_multi_args a{};
    a.samples = samples + begin * nchans;
    a.nsamples = nsamples - begin;
    a.nchans = nchans;
    a.chan_begin = t.chan_begin;
    a.chan_end = t.chan_end;
    a.delays = t.first.data();
    a.ndm = t.ndm;
    a.out = out + begin;
    a.out_nsamples = end - begin;
    a.out_stride = out_stride;
    a.max_block_spread = t.bound_first;
    hip_ok(gaffa_hip_dedisperse_multi_u8(&a, stream), "dedisperse_multi_u8");
  }
}

GpuSubbandBatch make_gpu_subband_batch(const std::vector<GpuSegmentTables>& segments) {
  check_arg(!segments.empty(), "subband batch requires at least one segment");
  GpuSubbandBatch b;
  const GpuSegmentTables& s0 = segments.front();
  b.chan_begin = s0.chan_begin;
  b.chan_end = s0.chan_end;
  b.subband_channels = s0.subband_channels;
  const std::size_t nsel = s0.chan_end - s0.chan_begin;
  std::vector<std::int32_t> coarse, residual;
  std::vector<gaffa_hip_subband_group> groups;
  for (const GpuSegmentTables& t : segments) {
    check_arg(t.subband && t.chan_begin == b.chan_begin && t.chan_end == b.chan_end &&
                  t.subband_channels == b.subband_channels,
              "subband batch segments must share the channel range and subband width");
    for (std::size_t k = 0; k < t.n_nominal; ++k) {
      const std::size_t first = k * t.ndm_per_nominal;
      gaffa_hip_subband_group g{};
      g.coarse_row = static_cast<std::uint32_t>(b.n_coarse_rows + k);
      g.first_dm = static_cast<std::uint32_t>(b.ndm + first);
      g.ndm = static_cast<std::uint32_t>(std::min(t.ndm_per_nominal, t.ndm - first));
      b.max_group = std::max<std::size_t>(b.max_group, g.ndm);
      groups.push_back(g);
    }
    coarse.insert(coarse.end(), t.host_first.begin(), t.host_first.end());
    residual.insert(residual.end(), t.host_second.begin(), t.host_second.end());
    b.n_coarse_rows += t.n_nominal;
    b.ndm += t.ndm;
    b.max_coarse = std::max(b.max_coarse, t.bound_first);
    b.max_residual = std::max(b.max_residual, t.bound_second);
    b.residual_window = std::max(b.residual_window, t.residual_window);
  }
  check_arg(coarse.size() == b.n_coarse_rows * nsel, "subband batch coarse table size mismatch");
  b.ngroups = groups.size();
  b.coarse = upload(coarse.data(), coarse.size());
  b.residual = upload(residual.data(), residual.size());
  b.groups = upload(groups.data(), groups.size());
  return b;
}

void enqueue_subband_batch_range(const GpuSubbandBatch& b, const GpuChannelMajorView& cm,
                                 std::size_t nsamples, std::uint32_t* out, std::size_t out_stride,
                                 std::size_t begin, std::size_t end, gaffa_hip_stream stream) {
  check_arg(begin < end && end <= nsamples, "dedispersion output range is empty");
  check_arg(begin % 4 == 0, "subband output ranges must start on a multiple of 4 samples");
  gaffa_hip_dedisperse_subband_args a{};
  a.channels = cm.data + (b.chan_begin - cm.chan_begin) * cm.stride + begin;
  a.channel_stride = cm.stride;
  a.channel_length = cm.stride - begin;
  a.nsamples = nsamples - begin;
  a.nsel = b.chan_end - b.chan_begin;
  a.subband_channels = b.subband_channels;
  a.ndm = b.ndm;
  a.coarse_delays = b.coarse.data();
  a.residual_delays = b.residual.data();
  a.out = out + begin;
  a.out_nsamples = end - begin;
  a.out_stride = out_stride;
  a.max_residual = b.max_residual;
  a.max_coarse = b.max_coarse;
  a.groups = b.groups.data();
  a.ngroups = b.ngroups;
  a.n_coarse_rows = b.n_coarse_rows;
  a.max_group_ndm = b.max_group;
  a.residual_window = b.residual_window;
  hip_ok(gaffa_hip_dedisperse_subband_u8(&a, stream), "dedisperse_subband_u8");
}

std::size_t spectrum_output_nsamples(std::size_t nsamples, std::span<const double> f,
                                     const SingleDmDedispersionPlan& p) {
  const MultiDmDedispersionPlan m{p.dm, 1.0, 1, p.ref_frequency_mhz, p.tsamp, p.chan_begin,
                                  p.chan_end};
  return multi_dm_output_nsamples(nsamples, f, m);
}

void dedisperse_spectrum_gpu_raw(const void* samples, std::size_t esize, const SampleShape& shape,
                                 std::span<const double> f, const SingleDmDedispersionPlan& p,
                                 void* out, std::size_t out_nsamples,
                                 const GpuDedispersionOptions& o) {
  const MultiDmDedispersionPlan m{p.dm, 1.0, 1, p.ref_frequency_mhz, p.tsamp, p.chan_begin,
                                  p.chan_end};
  validate_multi(shape, samples != nullptr, f, m);
  check_arg(shape.nifs == 1, "GPU dedispersion supports nifs == 1");
  check_arg(out_nsamples == multi_dm_output_nsamples(shape.nsamples, f, m),
            "spectrum output length must be the valid range");
  const std::vector<std::int32_t> delays = make_multi_dm_delay_table(f, m);
  for (std::int32_t d : delays) check_arg(d >= 0, "dedispersion requires non-negative delays");
  // largest channel tile whose delay spread keeps a 64-sample tile in LDS
  const std::size_t nc = delays.size();
  std::uint32_t tile = 256;
  std::int32_t spread = 0;
  for (; tile >= 8; tile /= 2) {
    spread = 0;
    for (std::size_t c0 = 0; c0 < nc; c0 += tile) {
      const auto b = delays.begin() + static_cast<std::ptrdiff_t>(c0);
      const auto e = delays.begin() + static_cast<std::ptrdiff_t>(std::min(nc, c0 + tile));
      spread = std::max(spread, *std::max_element(b, e) - *std::min_element(b, e));
    }
    if ((64 + static_cast<std::size_t>(spread)) * tile * esize <= 96 * 1024) break;
  }
  check_arg(tile >= 8, "spectrum delay spread too large for the GPU tile");
  GpuDeviceScope scope(o.device_id);
  GpuDeviceBuffer<std::int32_t> d_delays = upload(delays.data(), delays.size());
  gaffa_hip_spectrum_args a{};
  a.samples = samples;
  a.nsamples = shape.nsamples;
  a.nchans = shape.nchans;
  a.chan_begin = p.chan_begin;
  a.chan_end = p.chan_end;
  a.element_size = static_cast<std::uint32_t>(esize);
  a.delays = d_delays.data();
  a.out = out;
  a.out_nsamples = out_nsamples;
  a.channel_tile = tile;
  a.max_tile_spread = spread;
  hip_ok(gaffa_hip_dedisperse_spectrum(&a, o.stream), "dedisperse_spectrum");
  hip_ok(gaffa_hip_dedisperse_check(o.stream), "dedisperse_spectrum");
}

std::size_t subband_channel_stride(std::size_t nsamples, std::span<const double> f,
                                   const MultiDmDedispersionPlan& p,
                                   const SubbandDedispersionOptions& so) {
  const SubbandTables t = make_subband_tables(f, p, so);
  return static_cast<std::size_t>(
      gaffa_hip_subband_channel_stride(nsamples, t.max_coarse, t.max_residual));
}

void transpose_filterbank_gpu_into(GpuSampleView samples, std::size_t chan_begin,
                                   std::size_t chan_end, std::uint8_t* out, std::size_t stride,
                                   const GpuDedispersionOptions& o) {
  check_arg(samples.data != nullptr && out != nullptr, "transpose pointers must not be null");
  check_arg(samples.shape.nifs == 1, "GPU dedispersion supports nifs == 1");
  GpuDeviceScope scope(o.device_id);
  gaffa_hip_transpose_args a{};
  a.samples = samples.data;
  a.nsamples = samples.shape.nsamples;
  a.nchans = samples.shape.nchans;
  a.chan_begin = chan_begin;
  a.chan_end = chan_end;
  a.out = out;
  a.out_stride = stride;
  hip_ok(gaffa_hip_filterbank_transpose_u8(&a, o.stream), "filterbank_transpose_u8");
}

void dedisperse_subband_gpu_cm_into(const GpuChannelMajorView& cm, std::span<const double> f,
                                    const MultiDmDedispersionPlan& p,
                                    const SubbandDedispersionOptions& so, std::uint32_t* out,
                                    std::size_t out_stride, const GpuDedispersionOptions& o,
                                    std::size_t out_nsamples) {
  validate_multi(SampleShape{cm.nsamples, 1, f.size()}, cm.data != nullptr, f, p);
  check_arg(p.chan_begin >= cm.chan_begin && p.chan_end <= cm.chan_begin + cm.nchans,
            "plan channel range lies outside the channel-major copy");
  check_arg(so.subband_channels != 0, "subband_channels must be positive");
  check_arg(so.ndm_per_nominal != 0, "ndm_per_nominal must be positive");
  std::size_t n_out = multi_dm_output_nsamples(cm.nsamples, f, p);
  if (out_nsamples != 0) {
    check_arg(out_nsamples <= n_out, "requested output length exceeds the valid range");
    n_out = out_nsamples;
  }
  check_arg(out_stride >= n_out, "output stride must be >= out_nsamples");
  const SubbandTables t = make_subband_tables(f, p, so);
  for (std::int32_t d : t.coarse) check_arg(d >= 0, "dedispersion requires non-negative delays");
  GpuDeviceScope scope(o.device_id);
  GpuDeviceBuffer<std::int32_t> d_coarse = upload(t.coarse.data(), t.coarse.size());
  GpuDeviceBuffer<std::int32_t> d_resid = upload(t.residual.data(), t.residual.size());
  gaffa_hip_dedisperse_subband_args a{};
  a.channels = cm.data + (p.chan_begin - cm.chan_begin) * cm.stride;
  a.channel_stride = cm.stride;
  a.channel_length = cm.length;
  a.nsamples = cm.nsamples;
  a.nsel = p.chan_end - p.chan_begin;
  a.subband_channels = so.subband_channels;
  a.ndm_per_nominal = so.ndm_per_nominal;
  a.ndm = p.ndm;
  a.coarse_delays = d_coarse.data();
  a.residual_delays = d_resid.data();
  a.out = out;
  a.out_nsamples = n_out;
  a.out_stride = out_stride;
  a.max_residual = t.max_residual;
  a.max_coarse = t.max_coarse;
  a.residual_window = t.residual_window;
  hip_ok(gaffa_hip_dedisperse_subband_u8(&a, o.stream), "dedisperse_subband_u8");
  hip_ok(gaffa_hip_dedisperse_check(o.stream), "dedisperse_subband_u8");
}

void dedisperse_subband_gpu_into(GpuSampleView samples, std::span<const double> f,
                                 const MultiDmDedispersionPlan& p,
                                 const SubbandDedispersionOptions& so, std::uint32_t* out,
                                 std::size_t out_stride, const GpuDedispersionOptions& o,
                                 std::size_t out_nsamples) {
  validate_multi(samples.shape, samples.data != nullptr, f, p);
  const std::size_t stride = subband_channel_stride(samples.shape.nsamples, f, p, so);
  GpuDeviceScope scope(o.device_id);
  GpuDeviceBuffer<std::uint8_t> cm((p.chan_end - p.chan_begin) * stride);
  transpose_filterbank_gpu_into(samples, p.chan_begin, p.chan_end, cm.data(), stride, o);
  const GpuChannelMajorView v{cm.data(), samples.shape.nsamples, p.chan_begin,
                              p.chan_end - p.chan_begin, stride, o.device_id};
  // synchronises the stream (the delay-bound check), so `cm` may be freed after
  dedisperse_subband_gpu_cm_into(v, f, p, so, out, out_stride, o, out_nsamples);
}

GpuDedispersedResult<std::uint32_t> dedisperse_multi_dm_gpu_device(
    GpuSampleView samples, std::span<const double> f, const MultiDmDedispersionPlan& p,
    const GpuDedispersionOptions& o) {
  validate_multi(samples.shape, samples.data != nullptr, f, p);
  const std::size_t n_out = multi_dm_output_nsamples(samples.shape.nsamples, f, p);
  GpuDeviceScope scope(o.device_id);
  GpuDedispersedResult<std::uint32_t> r;
  r.shape = DedispersedShape{p.ndm, n_out};
  r.stride = padded_stride(n_out);
  r.device_id = o.device_id;
  r.data = GpuDeviceBuffer<std::uint32_t>(p.ndm * r.stride);
  dedisperse_multi_dm_gpu_into(samples, f, p, r.data.data(), r.stride, o);
  return r;
}

GpuDedispersedResult<std::uint32_t> dedisperse_subband_gpu_device(
    GpuSampleView samples, std::span<const double> f, const MultiDmDedispersionPlan& p,
    const SubbandDedispersionOptions& so, const GpuDedispersionOptions& o) {
  validate_multi(samples.shape, samples.data != nullptr, f, p);
  const std::size_t n_out = multi_dm_output_nsamples(samples.shape.nsamples, f, p);
  GpuDeviceScope scope(o.device_id);
  GpuDedispersedResult<std::uint32_t> r;
  r.shape = DedispersedShape{p.ndm, n_out};
  r.stride = padded_stride(n_out);
  r.device_id = o.device_id;
  r.data = GpuDeviceBuffer<std::uint32_t>(p.ndm * r.stride);
  dedisperse_subband_gpu_into(samples, f, p, so, r.data.data(), r.stride, o);
  return r;
}

GpuDedispersedResult<std::uint32_t> dedisperse_single_dm_gpu_device(
    GpuSampleView samples, std::span<const double> f, const SingleDmDedispersionPlan& sp,
    const GpuDedispersionOptions& o) {
  validate_samples(samples.shape, samples.data != nullptr, f);
  validate_plan_values(sp.ref_frequency_mhz, sp.tsamp);
  validate_channels(f, sp.ref_frequency_mhz, sp.chan_begin, sp.chan_end, samples.shape.nchans);
  check_arg(std::isfinite(sp.dm) && sp.dm >= 0.0, "dm must be finite and non-negative");
  // a single DM is a one-row multi-DM plan whose output length uses this DM
  const MultiDmDedispersionPlan p{sp.dm, 1.0, 1, sp.ref_frequency_mhz, sp.tsamp, sp.chan_begin,
                                  sp.chan_end};
  return dedisperse_multi_dm_gpu_device(samples, f, p, o);
}

namespace {

template <typename Run>
DedispersedResult<std::uint32_t> via_device(HostSampleView<std::uint8_t> samples,
                                            const GpuDedispersionOptions& o, Run&& run) {
  check_arg(!samples.data.empty(), "dedispersion samples data must not be null");
  check_arg(samples.data.size() == samples.size(),
            "dedispersion sample data size does not match shape");
  GpuDeviceScope scope(o.device_id);
  GpuDeviceBuffer<std::uint8_t> d = upload(samples.data.data(), samples.data.size());
  GpuSampleView v{d.data(), samples.shape, o.device_id};
  return copy_to_host(run(v));
}

}  // namespace

DedispersedResult<std::uint32_t> dedisperse_single_dm_gpu(HostSampleView<std::uint8_t> samples,
                                                          std::span<const double> f,
                                                          const SingleDmDedispersionPlan& p,
                                                          const GpuDedispersionOptions& o) {
  validate_samples(samples.shape, !samples.data.empty(), f);
  return via_device(samples, o, [&](GpuSampleView v) {
    return dedisperse_single_dm_gpu_device(v, f, p, o);
  });
}

DedispersedResult<std::uint32_t> dedisperse_multi_dm_gpu(HostSampleView<std::uint8_t> samples,
                                                         std::span<const double> f,
                                                         const MultiDmDedispersionPlan& p,
                                                         const GpuDedispersionOptions& o) {
  validate_multi(samples.shape, !samples.data.empty(), f, p);
  return via_device(samples, o, [&](GpuSampleView v) {
    return dedisperse_multi_dm_gpu_device(v, f, p, o);
  });
}

DedispersedResult<std::uint32_t> dedisperse_subband_gpu(HostSampleView<std::uint8_t> samples,
                                                        std::span<const double> f,
                                                        const MultiDmDedispersionPlan& p,
                                                        const SubbandDedispersionOptions& so,
                                                        const GpuDedispersionOptions& o) {
  validate_multi(samples.shape, !samples.data.empty(), f, p);
  check_arg(so.subband_channels != 0, "subband_channels must be positive");
  check_arg(so.ndm_per_nominal != 0, "ndm_per_nominal must be positive");
  return via_device(samples, o, [&](GpuSampleView v) {
    return dedisperse_subband_gpu_device(v, f, p, so, o);
  });
}

DedispersedResult<std::uint32_t> copy_to_host(const GpuDedispersedResult<std::uint32_t>& r) {
  DedispersedResult<std::uint32_t> h;
  h.shape = r.shape;
  h.data.resize(dedispersed_element_count(r.shape));
  if (r.shape.ndm == 0) return h;
  GpuDeviceScope scope(r.device_id);
  if (r.stride == r.shape.nsamples) {
    copy_to_host(h.data.data(), r.data.data(), h.data.size() * 4, nullptr);
  } else {
    std::vector<std::uint32_t> tmp(r.shape.ndm * r.stride);
    copy_to_host(tmp.data(), r.data.data(), tmp.size() * 4, nullptr);
    for (std::size_t i = 0; i < r.shape.ndm; ++i)
      std::copy_n(tmp.data() + i * r.stride, r.shape.nsamples, h.data.data() + i * r.shape.nsamples);
  }
  return h;
}

}  // namespace gaffa
