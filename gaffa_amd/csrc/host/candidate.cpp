// Candidate selection and harmonic flagging over the search's peak list.
#include "gaffa/candidate.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <optional>
#include <omp.h>
#include <parallel/algorithm>
#include <stdexcept>

namespace gaffa {
namespace {

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

bool peak_before(const DmPeak& a, const DmPeak& b) { return is_better_dm_peak(a, b); }

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

bool candidate_before(const Candidate& a, const Candidate& b) {
  if (peak_before(a.best_peak, b.best_peak)) return true;
  if (peak_before(b.best_peak, a.best_peak)) return false;
  if (a.peak_count != b.peak_count) return a.peak_count > b.peak_count;
  if (a.frequency != b.frequency) return a.frequency < b.frequency;
  return a.dm_index < b.dm_index;
}

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

namespace {

// A peak in the ranked selection: its index, its frequency position (period
// ranks counted down, so ascending frequency), its (DM, width) bucket, its
// frequency cluster and its S/N.
struct RankedPeak {
  std::uint32_t i, rdesc, dmw, cluster;
  float snr;
};

// Stable parallel counting sort of `in` into `out` by bucket(e) < nb: each
// thread counts a contiguous chunk, the offsets run bucket-major then
// thread-major, and each thread scatters its chunk in order.
template <class Bucket>
void counting_sort(const std::vector<RankedPeak>& in, std::vector<RankedPeak>& out,
                   std::size_t nb, Bucket bucket) {
  const std::size_t n = in.size();
  out.resize(n);
  const int nt = static_cast<int>(
      std::max<std::size_t>(1, std::min<std::size_t>(omp_get_max_threads(), n / 65536 + 1)));
  std::vector<std::size_t> hist(static_cast<std::size_t>(nt) * nb, 0);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; ++t) {
    std::size_t* h = hist.data() + static_cast<std::size_t>(t) * nb;
    for (std::size_t k = n * t / nt; k < n * (t + 1) / nt; ++k) ++h[bucket(in[k])];
  }
  // bucket totals and starts, then each thread's offset inside each bucket
  std::vector<std::size_t> start(nb + 1, 0);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (std::int64_t b = 0; b < static_cast<std::int64_t>(nb); ++b) {
    std::size_t total = 0;
    for (int t = 0; t < nt; ++t) total += hist[static_cast<std::size_t>(t) * nb + b];
    start[b + 1] = total;
  }
  for (std::size_t b = 0; b < nb; ++b) start[b + 1] += start[b];
#pragma omp parallel for num_threads(nt) schedule(static)
  for (std::int64_t b = 0; b < static_cast<std::int64_t>(nb); ++b) {
    std::size_t run = start[b];
    for (int t = 0; t < nt; ++t) {
      std::size_t& h = hist[static_cast<std::size_t>(t) * nb + b];
      const std::size_t c = h;
      h = run;
      run += c;
    }
  }
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; ++t) {
    std::size_t* h = hist.data() + static_cast<std::size_t>(t) * nb;
    for (std::size_t k = n * t / nt; k < n * (t + 1) / nt; ++k) out[h[bucket(in[k])]++] = in[k];
  }
}

}  // namespace

std::vector<Candidate> select_candidates_ranked_cpu(std::span<const DmPeak> peaks,
                                                    std::span<const std::uint32_t> period_rank,
                                                    double observation_seconds,
                                                    const CandidateSelectionOptions& o) {
  check_arg(observation_seconds > 0.0 && std::isfinite(observation_seconds),
            "Candidate selection observation_seconds must be finite and > 0");
  check_arg(o.frequency_cluster_radius >= 0.0 && std::isfinite(o.frequency_cluster_radius),
            "Candidate selection frequency_cluster_radius must be finite and >= 0");
  if (peaks.empty()) return {};
  check_arg(period_rank.size() == peaks.size(), "one period rank per peak is required");
  check_arg(peaks.size() < (std::size_t{1} << 32), "Candidate selection peak list too long");
  const std::int64_t np = static_cast<std::int64_t>(peaks.size());
  bool finite = true;
  std::uint32_t max_rank = 0;
  std::size_t max_dm = 0;
#pragma omp parallel for reduction(&& : finite) reduction(max : max_rank, max_dm) schedule(static)
  for (std::int64_t i = 0; i < np; ++i) {
    const DmPeak& p = peaks[static_cast<std::size_t>(i)];
    finite = finite && std::isfinite(p.dm) && std::isfinite(p.peak.period) &&
             std::isfinite(p.peak.frequency) && std::isfinite(p.peak.duty_cycle) &&
             std::isfinite(p.peak.snr);
    max_rank = std::max(max_rank, period_rank[static_cast<std::size_t>(i)]);
    max_dm = std::max(max_dm, p.dm_index);
  }
  check_arg(finite, "Candidate selection peaks must contain finite scientific values");
  // the distinct widths (a search has a handful), in ascending order
  std::vector<std::size_t> widths;
#pragma omp parallel
  {
    std::vector<std::size_t> local;
#pragma omp for schedule(static) nowait
    for (std::int64_t i = 0; i < np; ++i) {
      const std::size_t w = peaks[static_cast<std::size_t>(i)].peak.width;
      if (local.size() <= 64 && std::find(local.begin(), local.end(), w) == local.end())
        local.push_back(w);
    }
#pragma omp critical(gaffa_ranked_widths)
    for (std::size_t w : local)
      if (std::find(widths.begin(), widths.end(), w) == widths.end()) widths.push_back(w);
  }
  const std::size_t nr = std::size_t{max_rank} + 1, ndm = max_dm + 1, nw = widths.size();
  if (nw > 64 || nr > (std::size_t{1} << 26) || ndm * nw > (std::size_t{1} << 26))
    return select_candidates_cpu(peaks, observation_seconds, o);
  std::sort(widths.begin(), widths.end());
  const auto widx = [&](std::size_t w) {
    return static_cast<std::uint32_t>(std::lower_bound(widths.begin(), widths.end(), w) -
                                      widths.begin());
  };
  // each element, and each rank's frequency: the first peak of a rank sets
  // it (one compare-and-swap on its bits), every other peak of the rank must
  // carry the same value
  constexpr std::uint64_t kUnset = ~std::uint64_t{0};  // a NaN: never a finite frequency
  std::vector<std::uint64_t> fbits(nr, kUnset);
  std::vector<RankedPeak> a(peaks.size()), b;
  bool consistent = true;
#pragma omp parallel for reduction(&& : consistent) schedule(static)
  for (std::int64_t i = 0; i < np; ++i) {
    const DmPeak& p = peaks[static_cast<std::size_t>(i)];
    const std::uint32_t rk = period_rank[static_cast<std::size_t>(i)];
    std::uint64_t bits = 0;
    std::memcpy(&bits, &p.peak.frequency, sizeof(bits));
    std::uint64_t expected = kUnset;
    if (!__atomic_compare_exchange_n(&fbits[rk], &expected, bits, false, __ATOMIC_RELAXED,
                                     __ATOMIC_RELAXED))
      consistent = consistent && expected == bits;
    const std::uint32_t w = widx(p.peak.width);
    const std::size_t dmw = o.cluster_across_widths ? p.dm_index * nw + w : w * ndm + p.dm_index;
    a[static_cast<std::size_t>(i)] = RankedPeak{static_cast<std::uint32_t>(i),
                                                static_cast<std::uint32_t>(nr - 1 - rk),
                                                static_cast<std::uint32_t>(dmw), 0, p.peak.snr};
  }
  // the frequency clusters over the ranks present, ascending frequency; the
  // ranks must order the frequencies (1 / period is monotone), else the
  // general path
  std::vector<std::uint32_t> cluster_of(nr, 0);
  std::vector<double> freq_of(nr, 0.0);
  const double radius_hz = o.frequency_cluster_radius / observation_seconds;
  std::uint32_t nclusters = 0;
  double prev = 0.0;
  bool monotone = consistent;
  for (std::size_t rd = 0; rd < nr && monotone; ++rd) {
    const std::uint64_t bits = fbits[nr - 1 - rd];
    if (bits == kUnset) continue;
    double f = 0.0;
    std::memcpy(&f, &bits, sizeof(f));
    if (nclusters == 0) {
      nclusters = 1;
    } else {
      monotone = f >= prev;
      if (f - prev > radius_hz) ++nclusters;
    }
    cluster_of[rd] = nclusters - 1;
    freq_of[rd] = f;
    prev = f;
  }
  if (!monotone) return select_candidates_cpu(peaks, observation_seconds, o);
  counting_sort(a, b, nr, [](const RankedPeak& e) { return e.rdesc; });
#pragma omp parallel for schedule(static)
  for (std::int64_t k = 0; k < np; ++k) b[static_cast<std::size_t>(k)].cluster =
      cluster_of[b[static_cast<std::size_t>(k)].rdesc];
  // (cluster, DM, width) order with frequency ascending inside: two more
  // stable passes
  counting_sort(b, a, ndm * nw, [](const RankedPeak& e) { return e.dmw; });
  counting_sort(a, b, nclusters, [](const RankedPeak& e) { return e.cluster; });
  // DM-cluster boundaries, as select_candidates_cpu draws them
  const auto dm_of = [&](const RankedPeak& e) {
    return o.cluster_across_widths ? e.dmw / nw : e.dmw % ndm;
  };
  const auto w_of = [&](const RankedPeak& e) {
    return o.cluster_across_widths ? e.dmw % nw : e.dmw / ndm;
  };
  std::vector<std::uint8_t> split(peaks.size(), 0);
  split[0] = 1;
#pragma omp parallel for schedule(static)
  for (std::int64_t k = 1; k < np; ++k) {
    const RankedPeak& prev_e = b[static_cast<std::size_t>(k - 1)];
    const RankedPeak& cur = b[static_cast<std::size_t>(k)];
    const std::size_t dp = dm_of(prev_e), dc = dm_of(cur);
    const bool same_width = o.cluster_across_widths || w_of(cur) == w_of(prev_e);
    const bool same_dm = dc <= dp || dc - dp <= o.dm_cluster_radius;
    if (cur.cluster != prev_e.cluster || !same_width || !same_dm)
      split[static_cast<std::size_t>(k)] = 1;
  }
  std::vector<std::size_t> seg;
  for (std::size_t k = 0; k < split.size(); ++k)
    if (split[k]) seg.push_back(k);
  seg.push_back(split.size());
  const std::int64_t nseg = static_cast<std::int64_t>(seg.size()) - 1;
  std::vector<Candidate> out(static_cast<std::size_t>(nseg));
#pragma omp parallel for schedule(dynamic, 1)
  for (std::int64_t j = 0; j < nseg; ++j) {
    const RankedPeak* begin = b.data() + seg[j];
    const RankedPeak* end = b.data() + seg[j + 1];
    Candidate c;
    c.dm_index_min = std::numeric_limits<std::size_t>::max();
    c.frequency_min = std::numeric_limits<double>::infinity();
    c.frequency_max = -std::numeric_limits<double>::infinity();
    float top = -std::numeric_limits<float>::infinity();
    for (const RankedPeak* e = begin; e != end; ++e) {
      const std::size_t dm = dm_of(*e);
      c.dm_index_min = std::min(c.dm_index_min, dm);
      c.dm_index_max = std::max(c.dm_index_max, dm);
      c.frequency_min = std::min(c.frequency_min, freq_of[e->rdesc]);
      c.frequency_max = std::max(c.frequency_max, freq_of[e->rdesc]);
      top = std::max(top, e->snr);
    }
    const DmPeak* best = nullptr;
    for (const RankedPeak* e = begin; e != end; ++e) {
      if (e->snr != top) continue;
      const DmPeak& p = peaks[e->i];
      if (best == nullptr || peak_before(p, *best)) best = &p;
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
    out[static_cast<std::size_t>(j)] = c;
  }
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
