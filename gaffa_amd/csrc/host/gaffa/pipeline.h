// Device-resident DM search pipeline: filterbank -> (subband or brute-force)
// dedispersion over one or more uniform DM segments -> in-place convert +
// detrend + normalise -> batched FFA search -> DmPeak list.  This is the
// composition gaffa assembles by hand in benchmarks/cuda_dm_search_benchmark
// .cpp:179-294 and tests/cpp/test_cuda_search_pipeline.cpp:59-155, kept on
// one stream with no host round trip between stages.
#pragma once

#include "gaffa/candidate.h"
#include "gaffa/dedispersion.h"
#include "gaffa/ffa_gpu.h"
#include "gaffa/filterbank.h"
#include "gaffa/preprocessing.h"
#include "gaffa/synth.h"

#include <cstddef>
#include <cstdint>
#include <exception>
#include <memory>
#include <span>
#include <thread>
#include <vector>

namespace gaffa {

// One uniform run of DM trials; a plan may chain several (non-uniform steps).
struct DmSegment {
  double dm_low = 0.0;
  double dm_step = 1.0;
  std::size_t ndm = 0;
};

struct SearchPipelineConfig {
  std::size_t nsamples = 0;
  std::size_t nchans = 0;
  double tsamp = 0.0;
  std::vector<double> frequency_mhz;  // ascending, one per channel
  std::vector<DmSegment> dm_segments;
  bool subband = true;
  SubbandDedispersionOptions subband_options{};
  PreprocessPlan preprocess{};
  RiptideFfaPlanOptions ffa{};
  float snr_threshold = 6.0F;
  std::size_t ffa_workspace_bytes_limit = 32ULL << 30;
  std::size_t ffa_series_tile = 0;
  int device_id = 0;
};

// GPU time per stage from HIP events on the pipeline stream: dedispersion
// (incl. its table uploads), preprocessing, the FFA launches of every series
// tile (kernels only; the peak readback follows), and first event to last.
struct StageTimes {
  float dedisperse_ms = 0.0F;
  float preprocess_ms = 0.0F;
  float ffa_ms = 0.0F;
  float total_ms = 0.0F;
  // With kernel timing on: each FFA launch stage summed over the tiles of
  // the last pass (prepare, subtree, levels, detect), and the tile count.
  float ffa_stage_ms[FfaStageMarks::kStages] = {};
  std::uint32_t ffa_tiles = 0;
};

class GpuSearchPipeline {
 public:
  explicit GpuSearchPipeline(SearchPipelineConfig config);
  ~GpuSearchPipeline();
  GpuSearchPipeline(const GpuSearchPipeline&) = delete;
  GpuSearchPipeline& operator=(const GpuSearchPipeline&) = delete;

  void generate(const SynthFilterbankSpec& spec);
  void upload(const std::uint8_t* samples, std::size_t count);
  void download_filterbank(std::uint8_t* out) const;

  // Stages, each enqueued on the pipeline stream.
  void dedisperse();
  void preprocess();
  void search();
  StageTimes run();  // all three, timed with HIP events
  // Stream an 8-bit SIGPROC file (matching the configuration) to the device
  // and dedisperse each output range as soon as the samples it needs have
  // arrived, then preprocess and search.  stream_stats() describes the load.
  StageTimes run_file(const std::string& path, std::size_t chunk_samples = std::size_t{1} << 18);
  const FilterbankStreamStats& stream_stats() const noexcept { return stream_stats_; }
  // Record events between the FFA launch stages of every tile (off by
  // default: the marks split one enqueue into four).
  void set_kernel_timing(bool on) noexcept { kernel_timing_ = on; }

  std::size_t ndm() const noexcept { return ndm_; }
  // DM trials per nominal group each segment's subband tables use (always
  // the configured value); empty for brute force.
  std::vector<std::size_t> segment_ndm_per_nominal() const {
    std::vector<std::size_t> v;
    for (const GpuSegmentTables& t : segment_tables_)
      if (t.subband) v.push_back(t.ndm_per_nominal);
    return v;
  }
  std::size_t out_nsamples() const noexcept { return n_out_; }
  std::size_t plane_stride() const noexcept { return stride_; }
  double tsamp() const noexcept { return cfg_.tsamp; }
  int device_id() const noexcept { return cfg_.device_id; }
  std::size_t nchans() const noexcept { return cfg_.nchans; }
  const std::vector<double>& dms() const noexcept { return dms_; }
  const FfaSearchPlan& ffa_plan() const noexcept { return ffa_plan_; }
  const GpuFfaProgram& ffa_program() const noexcept { return *ffa_; }
  GpuFfaProgram& ffa_program() noexcept { return *ffa_; }
  // The last search's peaks in is_better_dm_peak order.  The search ends by
  // merging its per-tile sorted parts on a host thread, beside whatever the
  // caller enqueues next (candidate selection and folds read the arrival-order
  // records instead); peaks() and the host candidate selection wait for it.
  const std::vector<DmPeak>& peaks() const;
  std::size_t peak_count() const noexcept { return npeaks_; }
  // Waits for the ordering of the last search's peaks (rethrows its error).
  void wait_peaks() const;
  // select_candidates_cpu of the last search's peaks (observation =
  // out_nsamples * tsamp), equal candidate for candidate: on the device from
  // the compact peaks where the sort key fits (frequency clusters numbered on
  // the host from the ranks present), else through the host's ranked path;
  // both carry the trial periods as the plan's dense period ranks.
  std::vector<Candidate> select_candidates(const CandidateSelectionOptions& options);
  // The host selection alone (the fallback of select_candidates).
  std::vector<Candidate> select_candidates_host(const CandidateSelectionOptions& options) const;
  std::uint32_t* plane_u32() noexcept { return plane_.data(); }
  const std::uint8_t* filterbank() const noexcept { return filterbank_.data(); }
  gaffa_hip_stream stream() const noexcept { return stream_.get(); }
  // Copies rows [0, ndm) x [0, out_nsamples) of the plane (uint32 or float
  // bits, depending on the stage reached) into out.
  void download_plane(void* out) const;

  // Dense S/N periodogram of every DM trial, written by the following
  // searches: [ndm][trial rows][width trials] floats, rows in the order of
  // trial_periods() (0 where a task has fewer widths).  For the candidate
  // scorer's periodogram patches.
  void enable_periodogram(bool on);
  bool periodogram_enabled() const noexcept { return periodogram_.size() != 0; }
  const std::vector<double>& trial_periods() const noexcept { return trial_periods_; }
  std::size_t periodogram_widths() const noexcept { return ffa_plan_.width_trials.size(); }
  void download_periodogram(std::size_t dm_index, float* out) const;
  // Scorer inputs of candidates of the last search (preprocessed plane and
  // periodogram): cubes [n][32][64], the candidate's DM-trial series folded
  // at its period into 32 subints of 64 phase bins, and patches [n][128],
  // the 128 periodogram trials around its period at its width (edges
  // repeated).  Cubes are normalised to zero mean and unit variance,
  // patches to zero mean in S/N units.  Device outputs, enqueued on the
  // pipeline stream.
  void candidate_features(std::span<const Candidate> candidates, float* cubes, float* patches);
  // gaffa's full fold (pfold) of candidates from the resident filterbank:
  // per candidate the aligned spectrum at its DM (dedisperse_spectrum),
  // folded at its period into nsubint subints x output_channels x nbin.
  // Cubes go to `cubes` on the host when given ([n][nsubint][output
  // channels][nbin]); returns the device time in ms.
  double fold_candidates_full(std::span<const Candidate> candidates, std::size_t nbin,
                              std::size_t nsubint, std::size_t output_channels,
                              std::vector<float>* cubes);

 private:
  SearchPipelineConfig cfg_;
  GpuStream stream_;
  std::size_t ndm_ = 0, n_out_ = 0, stride_ = 0;
  std::vector<double> dms_;
  std::vector<MultiDmDedispersionPlan> segment_plans_;
  std::vector<GpuSegmentTables> segment_tables_;
  GpuSubbandBatch subband_batch_;  // every segment's nominal groups, one launch
  GpuDeviceBuffer<std::uint8_t> filterbank_;
  GpuDeviceBuffer<std::uint8_t> channels_;  // channel-major copy (subband path)
  std::size_t channel_stride_ = 0;
  GpuDeviceBuffer<std::uint32_t> plane_;
  FfaSearchPlan ffa_plan_;
  std::unique_ptr<GpuPreprocessProgram> prep_;
  std::unique_ptr<GpuFfaProgram> ffa_;
  std::vector<DmPeak> peaks_;
  std::size_t npeaks_ = 0;
  mutable std::thread merge_thread_;       // orders peaks_ after a search
  mutable std::exception_ptr merge_error_;
  std::vector<gaffa_hip_event> events_;
  bool kernel_timing_ = false;
  FilterbankStreamStats stream_stats_{};
  void dedisperse_range(std::size_t begin, std::size_t end);
  std::size_t lookahead_ = 0;  // input samples past an output range its kernels read
  std::vector<FfaStageMarks> marks_;  // one per tile, events created on demand
  std::size_t marks_used_ = 0;
  GpuDeviceBuffer<float> periodogram_;
  GpuPinnedBuffer compact_host_;  // compact peaks read back during a search
  GpuPinnedBuffer tile_counts_;   // running peak count after each FFA tile
  std::vector<gaffa_hip_event> tile_done_;  // after each tile's count copy
  GpuStream copy_stream_;         // peak downloads beside the search
  std::vector<DmPeakKey> peak_keys_;
  std::vector<DmPeakKey> merge_scratch_;
  std::vector<DmPeak> expanded_;        // peaks in arrival order, expanded per tile
  std::vector<std::size_t> run_ends_;   // end of each tile's sorted keys
  std::vector<std::uint32_t> period_ranks_;  // ffa_period_ranks(ffa_plan_)
  // device candidate selection: the rank table per evaluated (task, row),
  // each task's first row, the frequency of each descending rank, and the
  // buffers of the last selection (grown as needed)
  std::uint32_t nranks_ = 0;
  std::vector<double> freq_of_rdesc_;
  GpuDeviceBuffer<std::uint32_t> cand_rank_table_;
  GpuDeviceBuffer<std::uint64_t> cand_first_row_;
  GpuDeviceBuffer<std::uint32_t> cand_peak_rank_, cand_cluster_of_, cand_counts_;
  GpuDeviceBuffer<std::uint8_t> cand_present_, cand_workspace_;
  GpuDeviceBuffer<gaffa_hip_candidate_segment> cand_segments_;
  GpuDeviceBuffer<gaffa_hip_candidate_tie> cand_ties_;
  std::vector<Candidate> select_candidates_gpu(const CandidateSelectionOptions& options);
  std::vector<double> trial_periods_;
  std::vector<std::uint32_t> trial_order_;
};

}  // namespace gaffa
