// Host tables of the seeded synthetic filterbank generator and its launch.
#include "gaffa/synth.h"

#include "gaffa/dedispersion.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace gaffa {
namespace {

constexpr int kPhaseBins = 4096;
constexpr std::int64_t kPeriodFrac = 1LL << 20;

std::uint64_t splitmix64(std::uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

std::vector<double> synth_frequencies(const SynthFilterbankSpec& s) {
  std::vector<double> f(s.nchans);
  for (std::size_t c = 0; c < s.nchans; ++c) {
    f[c] = s.foff < 0.0 ? s.fch1 + static_cast<double>(s.nchans - 1 - c) * s.foff
                        : s.fch1 + static_cast<double>(c) * s.foff;
  }
  return f;
}

SynthTables make_synth_tables(const SynthFilterbankSpec& s) {
  if (s.nchans == 0 || s.nsamples == 0) throw std::invalid_argument("empty filterbank");
  if (!(s.tsamp > 0.0)) throw std::invalid_argument("tsamp must be positive");
  const std::vector<double> f = synth_frequencies(s);
  const double ref = *std::max_element(f.begin(), f.end());
  SynthTables t;
  t.key = splitmix64(s.seed);
  const std::size_t np = s.pulsars.size();
  t.delays.resize(np * s.nchans);
  t.profiles.resize(np * kPhaseBins);
  for (std::size_t p = 0; p < np; ++p) {
    const SynthPulsar& ps = s.pulsars[p];
    if (!(ps.period_s > 0.0) || !(ps.duty > 0.0) || !(ps.duty < 1.0)) {
      throw std::invalid_argument("pulsar period and duty must be positive, duty < 1");
    }
    for (std::size_t c = 0; c < s.nchans; ++c)
      t.delays[p * s.nchans + c] = dedispersion_delay_bins(ps.dm, f[c], ref, s.tsamp);
    // Gaussian of FWHM = duty (in phase), centred on phase 0, in 1/4096 units
    const double sigma = ps.duty / (2.0 * std::sqrt(2.0 * std::log(2.0)));
    for (int b = 0; b < kPhaseBins; ++b) {
      double x = (static_cast<double>(b) + 0.5) / kPhaseBins;
      if (x > 0.5) x = 1.0 - x;
      const double g = std::exp(-0.5 * (x / sigma) * (x / sigma));
      t.profiles[p * kPhaseBins + b] =
          static_cast<std::int32_t>(std::llround(4096.0 * 16.0 * ps.amp_sigma * g));
    }
    const std::int64_t pq = std::llround(ps.period_s / s.tsamp * kPeriodFrac);
    if (pq <= 0) throw std::invalid_argument("pulsar period too short for the sample time");
    t.period_q.push_back(pq);
    const double ph = ps.phase0 - std::floor(ps.phase0);
    t.phase0_q.push_back(std::llround(ph * static_cast<double>(pq)) % pq);
  }
  t.rfi_q.assign(s.nchans, 0);
  for (std::size_t c : s.rfi_channels) {
    if (c >= s.nchans) throw std::invalid_argument("rfi channel out of range");
    t.rfi_q[c] = s.rfi_offset * 4096;
  }
  return t;
}

void synth_filterbank_gpu(const SynthFilterbankSpec& s, std::uint8_t* out, std::size_t t_begin,
                          std::size_t t_count, gaffa_hip_stream stream) {
  if (t_begin + t_count > s.nsamples) throw std::invalid_argument("time range exceeds nsamples");
  const SynthTables t = make_synth_tables(s);
  GpuDeviceBuffer<std::int32_t> d_delays = upload(t.delays.data(), t.delays.size());
  GpuDeviceBuffer<std::int32_t> d_prof = upload(t.profiles.data(), t.profiles.size());
  GpuDeviceBuffer<std::int64_t> d_pq = upload(t.period_q.data(), t.period_q.size());
  GpuDeviceBuffer<std::int64_t> d_ph = upload(t.phase0_q.data(), t.phase0_q.size());
  GpuDeviceBuffer<std::int32_t> d_rfi = upload(t.rfi_q.data(), t.rfi_q.size());
  gaffa_hip_synth_args a{};
  a.out = out;
  a.nchans = s.nchans;
  a.t_begin = t_begin;
  a.t_count = t_count;
  a.key = t.key;
  a.npulsars = static_cast<std::uint32_t>(s.pulsars.size());
  a.delays = d_delays.data();
  a.profiles = d_prof.data();
  a.period_q = d_pq.data();
  a.phase0_q = d_ph.data();
  a.rfi_q = d_rfi.data();
  hip_ok(gaffa_hip_synth_filterbank(&a, stream), "synth_filterbank");
  hip_ok(gaffa_hip_stream_synchronize(stream), "hipStreamSynchronize");  // tables freed on return
}

}  // namespace gaffa
