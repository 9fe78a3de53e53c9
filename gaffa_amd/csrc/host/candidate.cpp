// Candidate selection and harmonic flagging over the search's peak list.
#include "gaffa/candidate.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <optional>
#include <parallel/algorithm>
#include <stdexcept>

namespace gaffa {
namespace {

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

bool peak_before(const DmPeak& a, const DmPeak& b) { return is_better_dm_peak(a, b); }

bool candidate_before(const Candidate& a, const Candidate& b) {
  if (peak_before(a.best_peak, b.best_peak)) return true;
  if (peak_before(b.best_peak, a.best_peak)) return false;
  if (a.peak_count != b.peak_count) return a.peak_count > b.peak_count;
  if (a.frequency != b.frequency) return a.frequency < b.frequency;
  return a.dm_index < b.dm_index;
}

// A peak's clustering key: frequency cluster order is set first, then DM
// clusters need (width when widths stay apart, DM index, width, frequency)
// order.  Peaks equal in all four keep no defined order, which changes
// nothing: they split nowhere between them and the best of a cluster is
// chosen by the full comparator.  The S/N rides along so the best peak of a
// cluster needs the full records only for peaks at its highest S/N.
struct DmKey {
  std::size_t width_split, dm_index, width;
  double frequency;
  float snr;
  std::uint32_t i;
  bool operator<(const DmKey& o) const {
    if (width_split != o.width_split) return width_split < o.width_split;
    if (dm_index != o.dm_index) return dm_index < o.dm_index;
    if (width != o.width) return width < o.width;
    return frequency < o.frequency;
  }
};

constexpr std::size_t kParallelClusterPeaks = std::size_t{1} << 18;

// One DM cluster (keys [begin, end)) -> its candidate; every peak is finite.
Candidate make_candidate(std::span<const DmPeak> peaks, const DmKey* begin, const DmKey* end) {
  Candidate c;
  c.dm_index_min = std::numeric_limits<std::size_t>::max();
  c.frequency_min = std::numeric_limits<double>::infinity();
  c.frequency_max = -std::numeric_limits<double>::infinity();
  float top = -std::numeric_limits<float>::infinity();
  for (const DmKey* k = begin; k != end; ++k) {
    c.dm_index_min = std::min(c.dm_index_min, k->dm_index);
    c.dm_index_max = std::max(c.dm_index_max, k->dm_index);
    c.frequency_min = std::min(c.frequency_min, k->frequency);
    c.frequency_max = std::max(c.frequency_max, k->frequency);
    top = std::max(top, k->snr);
  }
  // is_better_dm_peak orders finite peaks by S/N first: the best has the top S/N
  const DmPeak* best = nullptr;
  for (const DmKey* k = begin; k != end; ++k) {
    if (k->snr != top) continue;
    const DmPeak& p = peaks[k->i];
    if (best == nullptr || peak_before(p, *best)) best = &p;  // strict: the first of equals
  }
  c.dm = best->dm;
  c.dm_index = best->dm_index;
  c.period = best->peak.period;
  c.frequency = best->peak.frequency;
  c.width = best->peak.width;
  c.duty_cycle = best->peak.duty_cycle;
  c.snr = best->peak.snr;
  c.peak_count = static_cast<std::size_t>(end - begin);
  c.best_peak = *best;
  return c;
}

struct Ratio {
  std::int32_t n, d;
  double value;
};

// Reduced fractions n/d != 1 with n <= max_harmonic, d <= denominator_max,
// ascending by value (ties by n, d), each once.
std::vector<Ratio> harmonic_ratios(const HarmonicOptions& o) {
  std::vector<Ratio> r;
  for (std::size_t n = 1; n <= o.max_harmonic; ++n) {
    for (std::size_t d = 1; d <= o.denominator_max; ++d) {
      const std::size_t g = std::gcd(n, d);
      if (n / g == d / g) continue;
      r.push_back(Ratio{static_cast<std::int32_t>(n / g), static_cast<std::int32_t>(d / g),
                        static_cast<double>(n / g) / static_cast<double>(d / g)});
    }
  }
  std::sort(r.begin(), r.end(), [](const Ratio& a, const Ratio& b) {
    if (a.value != b.value) return a.value < b.value;
    if (a.n != b.n) return a.n < b.n;
    return a.d < b.d;
  });
  r.erase(std::unique(r.begin(), r.end(),
                      [](const Ratio& a, const Ratio& b) { return a.n == b.n && a.d == b.d; }),
          r.end());
  return r;
}

// The ratio nearest to x: of the first ratio >= x and the one before it, the
// strictly closer (the first on a tie).
std::optional<Ratio> nearest_ratio(double x, const std::vector<Ratio>& r) {
  if (!(x > 0.0) || r.empty()) return std::nullopt;
  const auto hi = std::lower_bound(r.begin(), r.end(), x,
                                   [](const Ratio& a, double v) { return a.value < v; });
  std::optional<Ratio> best;
  if (hi != r.end()) best = *hi;
  if (hi != r.begin()) {
    const Ratio& lo = *std::prev(hi);
    if (!best || std::abs(lo.value - x) < std::abs(best->value - x)) best = lo;
  }
  return best;
}

constexpr double kDispersionMs = 4.148808e3;  // ms MHz^2 per pc cm^-3

std::optional<HarmonicRelation> relation(const Candidate& parent, const Candidate& child,
                                         std::size_t parent_index, const HarmonicContext& ctx,
                                         const HarmonicOptions& o, const std::vector<Ratio>& r) {
  const auto ratio = nearest_ratio(child.frequency / parent.frequency, r);
  if (!ratio) return std::nullopt;
  HarmonicRelation h;
  h.is_harmonic = true;
  h.parent_index = parent_index;
  h.numerator = ratio->n;
  h.denominator = ratio->d;
  const double expected = parent.frequency * ratio->value;
  h.frequency_error_bins = std::abs(child.frequency - expected) * ctx.observation_seconds;
  if (h.frequency_error_bins > o.frequency_tolerance_bins) return std::nullopt;
  const double fast_duty =
      child.frequency >= parent.frequency ? child.duty_cycle : parent.duty_cycle;
  h.phase_distance = h.frequency_error_bins / fast_duty;
  if (h.phase_distance > o.phase_distance_max) return std::nullopt;
  const double lo = std::min(ctx.frequency_low_mhz, ctx.frequency_high_mhz);
  const double hi = std::max(ctx.frequency_low_mhz, ctx.frequency_high_mhz);
  const double delay_ms =
      std::abs(parent.dm - child.dm) * kDispersionMs * std::abs(1.0 / (lo * lo) - 1.0 / (hi * hi));
  const double width_s =
      std::min(parent.period * parent.duty_cycle, child.period * child.duty_cycle);
  h.dm_distance = (delay_ms / 1000.0) / width_s;
  if (h.dm_distance > o.dm_distance_max) return std::nullopt;
  h.expected_snr = static_cast<double>(parent.snr) /
                   std::sqrt(static_cast<double>(ratio->n) * static_cast<double>(ratio->d));
  h.snr_distance = std::abs(static_cast<double>(child.snr) - h.expected_snr);
  if (o.use_snr_consistency && h.snr_distance > o.snr_distance_max) return std::nullopt;
  return h;
}

}  // namespace

std::vector<Candidate> select_candidates_cpu(std::span<const DmPeak> peaks,
                                             double observation_seconds,
                                             const CandidateSelectionOptions& o) {
  check_arg(observation_seconds > 0.0 && std::isfinite(observation_seconds),
            "Candidate selection observation_seconds must be finite and > 0");
  check_arg(o.frequency_cluster_radius >= 0.0 && std::isfinite(o.frequency_cluster_radius),
            "Candidate selection frequency_cluster_radius must be finite and >= 0");
  if (peaks.empty()) return {};
  check_arg(peaks.size() < (std::size_t{1} << 32), "Candidate selection peak list too long");
  const std::int64_t np = static_cast<std::int64_t>(peaks.size());
  // frequency order of (frequency, index) pairs, then the gaps that split
  struct F {
    double f;
    std::uint32_t i;
  };
  std::vector<F> by_f(peaks.size());
  bool finite = true;
#pragma omp parallel for reduction(&& : finite) schedule(static)
  for (std::int64_t i = 0; i < np; ++i) {
    const DmPeak& p = peaks[static_cast<std::size_t>(i)];
    finite = finite && std::isfinite(p.dm) && std::isfinite(p.peak.period) &&
             std::isfinite(p.peak.frequency) && std::isfinite(p.peak.duty_cycle) &&
             std::isfinite(p.peak.snr);
    by_f[static_cast<std::size_t>(i)] = F{p.peak.frequency, static_cast<std::uint32_t>(i)};
  }
  check_arg(finite, "Candidate selection peaks must contain finite scientific values");
  __gnu_parallel::sort(by_f.begin(), by_f.end(), [](const F& a, const F& b) { return a.f < b.f; });
  const double radius_hz = o.frequency_cluster_radius / observation_seconds;
  std::vector<std::size_t> starts{0};
  for (std::size_t k = 1; k < by_f.size(); ++k)
    if (by_f[k].f - by_f[k - 1].f > radius_hz) starts.push_back(k);
  starts.push_back(by_f.size());
  // every peak's DM-cluster key, in frequency-cluster order
  std::vector<DmKey> keys(peaks.size());
#pragma omp parallel for schedule(static)
  for (std::int64_t k = 0; k < np; ++k) {
    const std::uint32_t i = by_f[static_cast<std::size_t>(k)].i;
    const DmPeak& p = peaks[i];
    keys[static_cast<std::size_t>(k)] =
        DmKey{o.cluster_across_widths ? std::size_t{0} : p.peak.width, p.dm_index, p.peak.width,
              p.peak.frequency, p.peak.snr, i};
  }
  by_f = std::vector<F>();
  const std::int64_t nclusters = static_cast<std::int64_t>(starts.size()) - 1;
  // sort each frequency cluster: large ones one at a time with the parallel
  // sort, the rest spread over the threads
  for (std::int64_t c = 0; c < nclusters; ++c)
    if (starts[c + 1] - starts[c] >= kParallelClusterPeaks)
      __gnu_parallel::sort(keys.begin() + static_cast<std::ptrdiff_t>(starts[c]),
                           keys.begin() + static_cast<std::ptrdiff_t>(starts[c + 1]));
#pragma omp parallel for schedule(dynamic, 16)
  for (std::int64_t c = 0; c < nclusters; ++c)
    if (starts[c + 1] - starts[c] < kParallelClusterPeaks)
      std::sort(keys.begin() + static_cast<std::ptrdiff_t>(starts[c]),
                keys.begin() + static_cast<std::ptrdiff_t>(starts[c + 1]));
  // DM-cluster boundaries: a frequency-cluster start, or a width change (when
  // widths stay apart) or a DM gap beyond the radius after the previous peak
  std::vector<std::uint8_t> split(peaks.size(), 0);
  for (std::int64_t c = 0; c < nclusters; ++c) split[starts[c]] = 1;
#pragma omp parallel for schedule(static)
  for (std::int64_t k = 1; k < np; ++k) {
    const DmKey& prev = keys[static_cast<std::size_t>(k - 1)];
    const DmKey& cur = keys[static_cast<std::size_t>(k)];
    const bool same_width = o.cluster_across_widths || cur.width == prev.width;
    const bool same_dm = cur.dm_index <= prev.dm_index ||
                         cur.dm_index - prev.dm_index <= o.dm_cluster_radius;
    if (!same_width || !same_dm) split[static_cast<std::size_t>(k)] = 1;
  }
  std::vector<std::size_t> seg;
  for (std::size_t k = 0; k < split.size(); ++k)
    if (split[k]) seg.push_back(k);
  seg.push_back(split.size());
  const std::int64_t nseg = static_cast<std::int64_t>(seg.size()) - 1;
  std::vector<Candidate> out(static_cast<std::size_t>(nseg));
#pragma omp parallel for schedule(dynamic, 1)
  for (std::int64_t j = 0; j < nseg; ++j)
    out[static_cast<std::size_t>(j)] =
        make_candidate(peaks, keys.data() + seg[j], keys.data() + seg[j + 1]);
  __gnu_parallel::sort(out.begin(), out.end(), candidate_before);
  if (o.max_candidates != 0 && out.size() > o.max_candidates) out.resize(o.max_candidates);
  return out;
}

std::vector<HarmonicCandidate> flag_harmonics_cpu(std::span<const Candidate> candidates,
                                                  const HarmonicContext& ctx,
                                                  const HarmonicOptions& o) {
  check_arg(ctx.observation_seconds > 0.0 && std::isfinite(ctx.observation_seconds),
            "Harmonic context observation_seconds must be finite and > 0");
  check_arg(ctx.frequency_low_mhz > 0.0 && ctx.frequency_high_mhz > 0.0 &&
                std::isfinite(ctx.frequency_low_mhz) && std::isfinite(ctx.frequency_high_mhz),
            "Harmonic context frequencies must be finite and > 0 MHz");
  check_arg(o.max_harmonic >= 2, "Harmonic max_harmonic must be >= 2");
  check_arg(o.denominator_max != 0, "Harmonic denominator_max must be > 0");
  check_arg(o.frequency_tolerance_bins >= 0.0 && std::isfinite(o.frequency_tolerance_bins),
            "Harmonic frequency_tolerance_bins must be finite and >= 0");
  check_arg(o.phase_distance_max >= 0.0 && std::isfinite(o.phase_distance_max),
            "Harmonic phase_distance_max must be finite and >= 0");
  check_arg(o.dm_distance_max >= 0.0 && std::isfinite(o.dm_distance_max),
            "Harmonic dm_distance_max must be finite and >= 0");
  check_arg(o.snr_distance_max >= 0.0 && std::isfinite(o.snr_distance_max),
            "Harmonic snr_distance_max must be finite and >= 0");
  if (candidates.empty()) return {};
  std::vector<HarmonicCandidate> out;
  out.reserve(candidates.size());
  for (const Candidate& c : candidates) {
    check_arg(std::isfinite(c.dm) && std::isfinite(c.period) && std::isfinite(c.frequency) &&
                  std::isfinite(c.duty_cycle) && std::isfinite(c.snr),
              "Harmonic candidates must contain finite scientific values");
    check_arg(c.frequency > 0.0 && c.period > 0.0 && c.duty_cycle > 0.0 && c.snr >= 0.0F,
              "Harmonic candidate frequency, period, duty_cycle, and snr are invalid");
    out.push_back(HarmonicCandidate{c, {}});
  }
  std::sort(out.begin(), out.end(), [](const HarmonicCandidate& x, const HarmonicCandidate& y) {
    const Candidate& a = x.candidate;
    const Candidate& b = y.candidate;
    if (a.snr != b.snr) return a.snr > b.snr;
    if (a.peak_count != b.peak_count) return a.peak_count > b.peak_count;
    if (a.frequency != b.frequency) return a.frequency < b.frequency;
    return a.dm_index < b.dm_index;
  });
  const std::vector<Ratio> ratios = harmonic_ratios(o);
  const std::int64_t n = static_cast<std::int64_t>(out.size());
  for (std::int64_t p = 0; p < n; ++p) {
    if (out[static_cast<std::size_t>(p)].harmonic.is_harmonic) continue;
    const Candidate parent = out[static_cast<std::size_t>(p)].candidate;
    // children are independent of each other for a fixed parent
#pragma omp parallel for schedule(static) if (n - p > 4096)
    for (std::int64_t c = p + 1; c < n; ++c) {
      HarmonicCandidate& child = out[static_cast<std::size_t>(c)];
      if (child.harmonic.is_harmonic) continue;
      if (auto h = relation(parent, child.candidate, static_cast<std::size_t>(p), ctx, o, ratios))
        child.harmonic = *h;
    }
  }
  return out;
}

std::vector<Candidate> remove_harmonics_cpu(std::span<const HarmonicCandidate> candidates,
                                            std::size_t max_candidates) {
  std::vector<Candidate> out;
  for (const HarmonicCandidate& c : candidates) {
    if (c.harmonic.is_harmonic) continue;
    out.push_back(c.candidate);
    if (max_candidates != 0 && out.size() >= max_candidates) break;
  }
  return out;
}

}  // namespace gaffa
