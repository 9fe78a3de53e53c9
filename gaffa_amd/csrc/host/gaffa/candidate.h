// Candidate selection (clustering DM peaks in frequency and DM) and harmonic
// flagging: the consumers of the search's DmPeak list, with gaffa's types,
// defaults and ordering (include/gaffa/candidate.h, harmonic.h).  Host code:
// they run once per search over the peak list the GPU produced.
#pragma once

#include "gaffa/ffa_gpu.h"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <vector>

namespace gaffa {

struct Candidate {
  double dm = 0.0;
  std::size_t dm_index = 0;
  double period = 0.0;
  double frequency = 0.0;
  std::size_t width = 0;
  double duty_cycle = 0.0;
  float snr = 0.0F;
  std::size_t peak_count = 0;
  std::size_t dm_index_min = 0;
  std::size_t dm_index_max = 0;
  double frequency_min = 0.0;
  double frequency_max = 0.0;
  DmPeak best_peak{};
};

struct CandidateSelectionOptions {
  // frequency clustering radius in Fourier bins (radius / observation_seconds Hz)
  double frequency_cluster_radius = 0.1;
  // DM clustering radius in trial indices
  std::size_t dm_cluster_radius = 1;
  bool cluster_across_widths = true;
  std::size_t max_candidates = 0;  // 0 = all
};

// The order of the candidate list (is_better_candidate, candidate.cpp:48-62):
// best peak first by is_better_dm_peak, then more peaks, lower frequency,
// lower DM index.
bool candidate_before(const Candidate& a, const Candidate& b);

// candidate.cpp:175-212: single-linkage clusters along frequency (gaps larger
// than the radius split), then along DM index inside each frequency cluster
// (optionally per width); one candidate per cluster from its best peak,
// ordered best first.
std::vector<Candidate> select_candidates_cpu(std::span<const DmPeak> peaks,
                                             double observation_seconds,
                                             const CandidateSelectionOptions& options = {});

// The same candidates as select_candidates_cpu, for peaks whose trial
// periods carry a dense rank (period_rank[i] < period_rank[j] exactly when
// peaks[i].peak.period < peaks[j].peak.period; equal ranks, equal periods):
// the frequency order is the rank order, so the sorts are stable counting
// sorts instead of comparison sorts.  No gaffa counterpart: the search
// pipeline's fast path (its ranks come from ffa_period_ranks).
std::vector<Candidate> select_candidates_ranked_cpu(std::span<const DmPeak> peaks,
                                                    std::span<const std::uint32_t> period_rank,
                                                    double observation_seconds,
                                                    const CandidateSelectionOptions& options = {});

inline constexpr std::size_t kNoHarmonicParent = std::numeric_limits<std::size_t>::max();

struct HarmonicContext {
  double observation_seconds = 0.0;
  double frequency_low_mhz = 0.0;
  double frequency_high_mhz = 0.0;
};

struct HarmonicOptions {
  std::size_t max_harmonic = 16;
  std::size_t denominator_max = 5;
  double frequency_tolerance_bins = 1.5;
  double phase_distance_max = 1.0;
  double dm_distance_max = 3.0;
  bool use_snr_consistency = false;
  double snr_distance_max = 3.0;
};

struct HarmonicRelation {
  bool is_harmonic = false;
  std::size_t parent_index = kNoHarmonicParent;
  std::int32_t numerator = 1;
  std::int32_t denominator = 1;
  double frequency_error_bins = 0.0;
  double phase_distance = 0.0;
  double dm_distance = 0.0;
  double expected_snr = 0.0;
  double snr_distance = 0.0;
};

struct HarmonicCandidate {
  Candidate candidate;
  HarmonicRelation harmonic;
};

// harmonic.cpp:245-311: candidates ordered by S/N; each unflagged candidate,
// in that order, flags the later unflagged candidates whose frequency is a
// reduced ratio n/d (n <= max_harmonic, d <= denominator_max) of its own
// within the frequency, phase and DM-smearing tolerances.
std::vector<HarmonicCandidate> flag_harmonics_cpu(std::span<const Candidate> candidates,
                                                  const HarmonicContext& context,
                                                  const HarmonicOptions& options = {});

std::vector<Candidate> remove_harmonics_cpu(std::span<const HarmonicCandidate> candidates,
                                            std::size_t max_candidates = 0);

}  // namespace gaffa
