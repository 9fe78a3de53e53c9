// Device-resident DM search pipeline (see gaffa/pipeline.h).
#include "gaffa/pipeline.h"

#include "gaffa/folding.h"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <limits>
#include <parallel/algorithm>
#include <stdexcept>

namespace gaffa {
namespace {

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

}  // namespace

GpuSearchPipeline::GpuSearchPipeline(SearchPipelineConfig config) : cfg_(std::move(config)) {
  check_arg(cfg_.nsamples > 0 && cfg_.nchans > 0, "pipeline filterbank must not be empty");
  check_arg(cfg_.frequency_mhz.size() == cfg_.nchans,
            "frequency table length must match sample channel count");
  check_arg(!cfg_.dm_segments.empty(), "pipeline requires at least one DM segment");
  GpuDeviceScope scope(cfg_.device_id);
  stream_ = GpuStream::create();
  const double ref = *std::max_element(cfg_.frequency_mhz.begin(), cfg_.frequency_mhz.end());
  n_out_ = cfg_.nsamples;
  for (const DmSegment& s : cfg_.dm_segments) {
    check_arg(s.ndm > 0, "dedispersion requires at least one DM");
    MultiDmDedispersionPlan p{s.dm_low, s.dm_step, s.ndm, ref, cfg_.tsamp, 0, cfg_.nchans};
    n_out_ = std::min(n_out_, multi_dm_output_nsamples(cfg_.nsamples, cfg_.frequency_mhz, p));
    segment_plans_.push_back(p);
    for (std::size_t i = 0; i < s.ndm; ++i)
      dms_.push_back(s.dm_low + static_cast<double>(i) * s.dm_step);
    ndm_ += s.ndm;
  }
  stride_ = (n_out_ + 63) / 64 * 64;
  for (const MultiDmDedispersionPlan& p : segment_plans_) {
    segment_tables_.push_back(
        make_gpu_segment_tables(cfg_.frequency_mhz, p, cfg_.subband, cfg_.subband_options));
    lookahead_ = std::max(lookahead_, segment_tables_.back().lookahead);
  }
  filterbank_ = GpuDeviceBuffer<std::uint8_t>(cfg_.nsamples * cfg_.nchans);
  if (cfg_.subband) {
    subband_batch_ = make_gpu_subband_batch(segment_tables_);
    for (const MultiDmDedispersionPlan& p : segment_plans_)
      channel_stride_ = std::max(channel_stride_,
                                 subband_channel_stride(cfg_.nsamples, cfg_.frequency_mhz, p,
                                                        cfg_.subband_options));
    channels_ = GpuDeviceBuffer<std::uint8_t>(cfg_.nchans * channel_stride_);
  }
  plane_ = GpuDeviceBuffer<std::uint32_t>(ndm_ * stride_);
  ffa_plan_ = make_riptide_ffa_plan(n_out_, cfg_.tsamp, cfg_.ffa);
  if (!cfg_.preprocess.steps.empty()) {
    GpuPreprocessExecutionOptions pe;
    pe.series_tile_size = std::min<std::size_t>(ndm_, 65535);  // one launch: the statistics run a lane per series
    pe.max_nsamples = n_out_;
    pe.stream = stream_.get();
    prep_ = std::make_unique<GpuPreprocessProgram>(cfg_.preprocess,
                                                   GpuPreprocessProgramOptions{cfg_.device_id}, pe);
  }
  GpuFfaExecutionOptions fe;
  fe.workspace_bytes_limit = cfg_.ffa_workspace_bytes_limit;
  fe.series_tile_size = cfg_.ffa_series_tile;
  fe.stream = stream_.get();
  // a full DM plane easily yields millions of significant peaks: start large
  // so the first run rarely repeats a tile
  fe.initial_peak_buffer_bytes = std::max<std::size_t>(fe.initial_peak_buffer_bytes, 256u << 20);
  ffa_ = std::make_unique<GpuFfaProgram>(ffa_plan_, GpuFfaProgramOptions{cfg_.device_id}, fe);
  period_ranks_ = ffa_period_ranks(ffa_plan_);
  if (!period_ranks_.empty()) {
    // device candidate selection: the rank of each evaluated (task, row),
    // each task's first row, and the frequency of each descending rank (the
    // peaks' own 1 / period, make_ffa_peak)
    const std::vector<double> periods = ffa_plan_trial_periods(ffa_plan_);
    std::vector<std::uint64_t> first_row(ffa_plan_.tasks.size() + 1, 0);
    for (std::size_t k = 0; k < ffa_plan_.tasks.size(); ++k)
      first_row[k + 1] = first_row[k] + ffa_plan_.tasks[k].rows_eval;
    std::uint32_t top = 0;
    for (std::uint32_t r : period_ranks_) top = std::max(top, r);
    nranks_ = top + 1;
    freq_of_rdesc_.assign(nranks_, 0.0);
    for (std::size_t i = 0; i < periods.size(); ++i)
      freq_of_rdesc_[nranks_ - 1 - period_ranks_[i]] = 1.0 / periods[i];
    cand_rank_table_ = gaffa::upload(period_ranks_.data(), period_ranks_.size());
    cand_first_row_ = gaffa::upload(first_row.data(), first_row.size());
  }
  events_.resize(7, nullptr);
  for (auto& e : events_) hip_ok(gaffa_hip_event_create(&e), "hipEventCreate");
}

GpuSearchPipeline::~GpuSearchPipeline() {
  if (merge_thread_.joinable()) merge_thread_.join();
  for (auto e : events_)
    if (e != nullptr) (void)gaffa_hip_event_destroy(e);
  for (auto e : tile_done_)
    if (e != nullptr) (void)gaffa_hip_event_destroy(e);
  for (auto& m : marks_)
    for (auto e : m.at)
      if (e != nullptr) (void)gaffa_hip_event_destroy(e);
}

void GpuSearchPipeline::generate(const SynthFilterbankSpec& spec) {
  check_arg(spec.nsamples == cfg_.nsamples && spec.nchans == cfg_.nchans,
            "synthetic filterbank shape must match the pipeline");
  GpuDeviceScope scope(cfg_.device_id);
  synth_filterbank_gpu(spec, filterbank_.data(), 0, spec.nsamples, stream_.get());
}

void GpuSearchPipeline::upload(const std::uint8_t* samples, std::size_t count) {
  check_arg(count == filterbank_.size(), "filterbank size does not match the pipeline");
  GpuDeviceScope scope(cfg_.device_id);
  copy_to_device(filterbank_.data(), samples, count, nullptr);
}

void GpuSearchPipeline::download_filterbank(std::uint8_t* out) const {
  GpuDeviceScope scope(cfg_.device_id);
  copy_to_host(out, filterbank_.data(), filterbank_.size(), stream_.get());
}

void GpuSearchPipeline::dedisperse_range(std::size_t begin, std::size_t end) {
  const GpuChannelMajorView cm{channels_.data(), cfg_.nsamples, 0, cfg_.nchans, channel_stride_,
                               cfg_.device_id};
  if (cfg_.subband) {
    enqueue_subband_batch_range(subband_batch_, cm, cfg_.nsamples, plane_.data(), stride_, begin,
                                end, stream_.get());
    return;
  }
  std::size_t row = 0;
  for (std::size_t k = 0; k < segment_tables_.size(); ++k) {
    enqueue_segment_range(segment_tables_[k], filterbank_.data(), cm, cfg_.nsamples, cfg_.nchans,
                          plane_.data() + row * stride_, stride_, begin, end, stream_.get());
    row += segment_plans_[k].ndm;
  }
}

void GpuSearchPipeline::dedisperse() {
  GpuDeviceScope scope(cfg_.device_id);
  if (cfg_.subband) {
    // one channel-major copy serves every DM segment
    const GpuSampleView v{filterbank_.data(), SampleShape{cfg_.nsamples, 1, cfg_.nchans},
                          cfg_.device_id};
    GpuDedispersionOptions o;
    o.device_id = cfg_.device_id;
    o.stream = stream_.get();
    transpose_filterbank_gpu_into(v, 0, cfg_.nchans, channels_.data(), channel_stride_, o);
  }
  dedisperse_range(0, n_out_);
  // one check of the kernels' delay-bound guard for the whole plane
  hip_ok(gaffa_hip_dedisperse_check(stream_.get()), "dedisperse");
}

StageTimes GpuSearchPipeline::run_file(const std::string& path, std::size_t chunk_samples) {
  GpuDeviceScope scope(cfg_.device_id);
  const FilterbankHeader h = read_filterbank_header(path);
  check_arg(h.nbits == 8 && h.nifs == 1, "pipeline files must be 8-bit with nifs == 1");
  check_arg(static_cast<std::size_t>(h.nchans) == cfg_.nchans &&
                h.nsamples == cfg_.nsamples && h.tsamp == cfg_.tsamp,
            "filterbank file shape does not match the pipeline");
  for (std::size_t c = 0; c < cfg_.nchans; ++c)
    check_arg(std::abs(h.frequency_table[c] - cfg_.frequency_mhz[c]) <= 1e-9,
              "filterbank file frequencies do not match the pipeline");
  chunk_samples = std::max<std::size_t>(256, (chunk_samples + 255) / 256 * 256);
  gaffa_hip_stream st = stream_.get();
  hip_ok(gaffa_hip_event_record(events_[0], st), "hipEventRecord");
  std::size_t transposed = 0, done = 0;
  stream_stats_ = stream_filterbank_u8(path, h, filterbank_.data(), st, chunk_samples,
                                       [&](std::size_t ready) {
    if (cfg_.subband) {
      gaffa_hip_transpose_args a{};
      a.samples = filterbank_.data();
      a.nsamples = ready;
      a.nchans = cfg_.nchans;
      a.chan_begin = 0;
      a.chan_end = cfg_.nchans;
      a.out = channels_.data();
      a.out_stride = channel_stride_;
      a.t_begin = transposed;
      // the last chunk also writes the zero padding to the end of the rows
      a.t_count = ready == cfg_.nsamples ? channel_stride_ - transposed : ready - transposed;
      hip_ok(gaffa_hip_filterbank_transpose_u8(&a, st), "filterbank_transpose_u8");
      transposed = ready;
    }
    // ranges start on 256-sample boundaries: the subband kernel reads its
    // channel rows in aligned dwords from the range start
    const std::size_t target =
        ready == cfg_.nsamples
            ? n_out_
            : (ready > lookahead_ ? std::min(n_out_, (ready - lookahead_) / 256 * 256) : 0);
    if (target > done) {
      dedisperse_range(done, target);
      done = target;
    }
  });
  hip_ok(gaffa_hip_dedisperse_check(st), "dedisperse");
  hip_ok(gaffa_hip_event_record(events_[1], st), "hipEventRecord");
  preprocess();
  hip_ok(gaffa_hip_event_record(events_[2], st), "hipEventRecord");
  if (prep_) prep_->synchronize();
  search();
  stream_.synchronize();
  StageTimes t;
  hip_ok(gaffa_hip_event_elapsed_ms(events_[0], events_[1], &t.dedisperse_ms), "elapsed");
  hip_ok(gaffa_hip_event_elapsed_ms(events_[1], events_[2], &t.preprocess_ms), "elapsed");
  hip_ok(gaffa_hip_event_elapsed_ms(events_[3], events_[4], &t.ffa_ms), "elapsed");
  hip_ok(gaffa_hip_event_elapsed_ms(events_[0], events_[4], &t.total_ms), "elapsed");
  return t;
}

void GpuSearchPipeline::preprocess() {
  GpuDeviceScope scope(cfg_.device_id);
  if (prep_) {
    prep_->run(plane_.data(), ndm_, n_out_, stride_, true);
  } else {
    gaffa_hip_preprocess_args a{};
    a.data = plane_.data();
    a.nseries = ndm_;
    a.nsamples = n_out_;
    a.stride = stride_;
    a.input_is_u32 = 1;
    a.scrunch_factor = 1;
    a.median_window = 1;
    GpuDeviceBuffer<std::uint8_t> ws(gaffa_hip_preprocess_workspace_bytes(ndm_, n_out_, 1));
    GpuDeviceBuffer<std::int32_t> st(ndm_);
    hip_ok(gaffa_hip_memset(st.data(), 0, st.bytes(), stream_.get()), "hipMemset");
    a.workspace = ws.data();
    a.workspace_bytes = ws.bytes();
    a.status = st.data();
    hip_ok(gaffa_hip_detrend_normalise_batch(&a, stream_.get()), "convert_uint32_to_float");
    stream_.synchronize();
  }
}

void GpuSearchPipeline::wait_peaks() const {
  if (merge_thread_.joinable()) merge_thread_.join();
  if (merge_error_) {
    std::exception_ptr e = merge_error_;
    merge_error_ = nullptr;
    std::rethrow_exception(e);
  }
}

const std::vector<DmPeak>& GpuSearchPipeline::peaks() const {
  wait_peaks();
  return peaks_;
}

void GpuSearchPipeline::search() {
  wait_peaks();  // the previous search's ordering still reads the buffers below
  GpuDeviceScope scope(cfg_.device_id);
  GpuFfaProgram& prog = *ffa_;
  const float* series = reinterpret_cast<const float*>(plane_.data());
  const std::size_t tile = prog.tile_capacity();
  const std::size_t ntiles = (ndm_ + tile - 1) / tile;
  if (kernel_timing_) {
    while (marks_.size() < ntiles) {
      FfaStageMarks m;
      for (auto& e : m.at) hip_ok(gaffa_hip_event_create(&e), "hipEventCreate");
      marks_.push_back(m);
    }
  }
  // per tile: the running peak count, copied back on the search stream, and
  // an event after it
  while (tile_done_.size() < ntiles) {
    gaffa_hip_event e = nullptr;
    hip_ok(gaffa_hip_event_create(&e), "hipEventCreate");
    tile_done_.push_back(e);
  }
  if (tile_counts_.bytes() < ntiles * sizeof(std::uint32_t))
    tile_counts_ = GpuPinnedBuffer(ntiles * sizeof(std::uint32_t));
  auto* counts = static_cast<std::uint32_t*>(tile_counts_.data());
  if (copy_stream_.get() == nullptr) copy_stream_ = GpuStream::create();
  for (;;) {
    const std::size_t capacity = prog.peak_capacity();
    hip_ok(gaffa_hip_event_record(events_[3], stream_.get()), "hipEventRecord");
    for (std::size_t k = 0; k < ntiles; ++k) {
      const std::size_t s0 = k * tile;
      const std::size_t ns = std::min(tile, ndm_ - s0);
      prog.enqueue_tile(series + s0 * stride_, ns, stride_, static_cast<std::uint32_t>(s0),
                        cfg_.snr_threshold, s0 == 0, kernel_timing_ ? &marks_[k] : nullptr);
      prog.enqueue_peak_count_copy(counts + k);
      hip_ok(gaffa_hip_event_record(tile_done_[k], stream_.get()), "hipEventRecord");
    }
    marks_used_ = kernel_timing_ ? ntiles : 0;
    hip_ok(gaffa_hip_event_record(events_[4], stream_.get()), "hipEventRecord");
    // While the device searches tile k + 1 and on, the host brings tile k's
    // peaks back (they follow tile k - 1's in the append-only buffer),
    // computes and sorts their keys and expands them; the end only merges.
    std::size_t done = 0;
    bool overflow = false;
    run_ends_.clear();
    for (std::size_t k = 0; k < ntiles; ++k) {
      hip_ok(gaffa_hip_event_synchronize(tile_done_[k]), "hipEventSynchronize");
      const std::size_t c = counts[k];
      if (c > capacity) {
        overflow = true;
        break;
      }
      if (c == done) continue;
      if (peak_keys_.size() < c) peak_keys_.resize(std::max(c, peak_keys_.size() * 2));
      if (expanded_.size() < c) expanded_.resize(std::max(c, expanded_.size() * 2));
      if (compact_host_.bytes() < c * sizeof(gaffa_hip_ffa_peak)) {
        // grow the pinned buffer, keeping the peaks already brought back
        GpuPinnedBuffer bigger(std::min(capacity, c + c / 2) * sizeof(gaffa_hip_ffa_peak));
        if (done != 0) std::memcpy(bigger.data(), compact_host_.data(), done * sizeof(gaffa_hip_ffa_peak));
        compact_host_ = std::move(bigger);
      }
      auto* compact = static_cast<gaffa_hip_ffa_peak*>(compact_host_.data());
      prog.enqueue_peak_copy(done, c - done, compact + done, copy_stream_.get());
      copy_stream_.synchronize();
      const std::span<const gaffa_hip_ffa_peak> part(compact + done, c - done);
      dm_peak_keys(prog.execution_plan(), part, done, dms_, period_ranks_, peak_keys_.data() + done);
      sort_dm_peak_keys(std::span<DmPeakKey>(peak_keys_.data() + done, c - done));
      expand_dm_peaks(prog.execution_plan(), part, dms_, expanded_.data() + done);
      run_ends_.push_back(c);
      done = c;
    }
    if (overflow) {
      stream_.synchronize();
      prog.grow_peaks(counts[ntiles - 1]);  // repeat the search with room for every peak
      continue;
    }
    peak_keys_.resize(done);
    if (run_ends_.empty()) run_ends_.push_back(0);
    // the ordering reads expanded_, peak_keys_ and run_ends_ and writes
    // peaks_; every member function that touches them waits for it first
    npeaks_ = done;
    merge_thread_ = std::thread([this, done] {
      try {
        merge_sorted_dm_peaks(std::span<const DmPeak>(expanded_.data(), done), run_ends_,
                              peak_keys_, merge_scratch_, peaks_);
      } catch (...) {
        merge_error_ = std::current_exception();
      }
    });
    break;
  }
}

std::vector<Candidate> GpuSearchPipeline::select_candidates(
    const CandidateSelectionOptions& options) {
  std::vector<Candidate> out = select_candidates_gpu(options);
  if (!out.empty()) return out;
  return select_candidates_host(options);
}

namespace {

std::uint32_t key_bits(std::uint64_t count) {  // bits to hold 0 .. count - 1
  std::uint32_t b = 1;
  while (b < 64 && (std::uint64_t{1} << b) < count) ++b;
  return b;
}

}  // namespace

// The device selection, or an empty list where it does not apply (no rank
// table, peaks not on the device, or a sort key wider than 64 bits); every
// peak of a search is finite except possibly its S/N, which the device checks.
std::vector<Candidate> GpuSearchPipeline::select_candidates_gpu(
    const CandidateSelectionOptions& o) {
  const double obs = static_cast<double>(n_out_) * cfg_.tsamp;
  check_arg(obs > 0.0 && std::isfinite(obs),
            "Candidate selection observation_seconds must be finite and > 0");
  check_arg(o.frequency_cluster_radius >= 0.0 && std::isfinite(o.frequency_cluster_radius),
            "Candidate selection frequency_cluster_radius must be finite and >= 0");
  const std::size_t n = npeaks_;
  const std::size_t nw = ffa_plan_.width_trials.size();
  if (n == 0 || nranks_ == 0 || n >= (std::size_t{1} << 32) || expanded_.size() < n ||
      ffa_->peak_capacity() < n || nw == 0 || nw >= (1u << 16))
    return {};
  for (double dm : dms_)
    check_arg(std::isfinite(dm), "Candidate selection peaks must contain finite scientific values");
  GpuDeviceScope scope(cfg_.device_id);
  gaffa_hip_stream st = stream_.get();
  if (cand_peak_rank_.size() < n) {
    cand_peak_rank_ = GpuDeviceBuffer<std::uint32_t>(n + n / 4);
    cand_segments_ = GpuDeviceBuffer<gaffa_hip_candidate_segment>(n + n / 4);
    cand_ties_ = GpuDeviceBuffer<gaffa_hip_candidate_tie>(n + n / 4);
    std::uint64_t bytes = 0;
    hip_ok(gaffa_hip_candidate_workspace_bytes(n + n / 4, &bytes), "candidate workspace");
    cand_workspace_ = GpuDeviceBuffer<std::uint8_t>(bytes);
  }
  if (cand_present_.size() < nranks_) {
    cand_present_ = GpuDeviceBuffer<std::uint8_t>(nranks_);
    cand_cluster_of_ = GpuDeviceBuffer<std::uint32_t>(nranks_);
  }
  if (cand_counts_.size() < 3) cand_counts_ = GpuDeviceBuffer<std::uint32_t>(3);
  std::uint64_t ws_bytes = 0;
  hip_ok(gaffa_hip_candidate_workspace_bytes(n, &ws_bytes), "candidate workspace");
  if (cand_workspace_.size() < ws_bytes) cand_workspace_ = GpuDeviceBuffer<std::uint8_t>(ws_bytes);
  gaffa_hip_candidate_args a{};
  a.peaks = ffa_->peak_data();
  a.npeaks = n;
  a.period_rank = cand_rank_table_.data();
  a.task_first_row = cand_first_row_.data();
  a.ntasks = static_cast<std::uint32_t>(ffa_plan_.tasks.size());
  a.nranks = nranks_;
  a.ndm = static_cast<std::uint32_t>(ndm_);
  a.nwidths = static_cast<std::uint32_t>(nw);
  a.peak_rank = cand_peak_rank_.data();
  a.present = cand_present_.data();
  a.status = cand_counts_.data() + 2;
  hip_ok(gaffa_hip_candidate_ranks(&a, st), "candidate ranks");
  // frequency clusters over the ranks present, ascending frequency
  std::vector<std::uint8_t> present(nranks_);
  std::uint32_t status = 0;
  copy_to_host(present.data(), cand_present_.data(), nranks_, st);
  copy_to_host(&status, cand_counts_.data() + 2, sizeof(status), st);
  check_arg((status & GAFFA_HIP_CANDIDATE_NONFINITE) == 0,
            "Candidate selection peaks must contain finite scientific values");
  if ((status & GAFFA_HIP_CANDIDATE_BAD_PEAK) != 0)
    throw std::logic_error("candidate selection: a compact peak lies outside the plan");
  const double radius_hz = o.frequency_cluster_radius / obs;
  std::vector<std::uint32_t> cluster_of(nranks_, 0);
  std::uint32_t nclusters = 0;
  double prev = 0.0;
  for (std::uint32_t rd = 0; rd < nranks_; ++rd) {
    if (!present[nranks_ - 1 - rd]) continue;
    const double f = freq_of_rdesc_[rd];
    check_arg(std::isfinite(f), "Candidate selection peaks must contain finite scientific values");
    if (nclusters == 0 || f - prev > radius_hz) ++nclusters;
    cluster_of[rd] = nclusters - 1;
    prev = f;
  }
  if (key_bits(nclusters) + key_bits(nranks_) +
          key_bits(static_cast<std::uint64_t>(ndm_) * nw) > 64)
    return {};
  copy_to_device(cand_cluster_of_.data(), cluster_of.data(), nranks_ * sizeof(std::uint32_t), st);
  a.cluster_of_rdesc = cand_cluster_of_.data();
  a.nclusters = nclusters;
  a.dm_cluster_radius = static_cast<std::uint32_t>(std::min<std::size_t>(o.dm_cluster_radius, ndm_));
  a.cluster_across_widths = o.cluster_across_widths ? 1u : 0u;
  a.workspace = cand_workspace_.data();
  a.workspace_bytes = cand_workspace_.size();
  a.segments = cand_segments_.data();
  a.nsegments = cand_counts_.data();
  a.ties = cand_ties_.data();
  a.nties = cand_counts_.data() + 1;
  hip_ok(gaffa_hip_candidate_clusters(&a, st), "candidate clusters");
  std::uint32_t counts[2] = {0, 0};
  copy_to_host(counts, cand_counts_.data(), sizeof(counts), st);
  const std::size_t nseg = counts[0], nties = counts[1];
  if (nseg == 0 || nseg > n || nties < nseg || nties > n)
    throw std::logic_error("candidate selection: inconsistent cluster counts");
  std::vector<gaffa_hip_candidate_segment> seg(nseg);
  std::vector<gaffa_hip_candidate_tie> ties(nties);
  copy_to_host(seg.data(), cand_segments_.data(), nseg * sizeof(seg[0]), st);
  copy_to_host(ties.data(), cand_ties_.data(), nties * sizeof(ties[0]), st);
  // each cluster's representative: the best of its peaks at the top S/N
  std::vector<std::uint32_t> best(nseg, std::numeric_limits<std::uint32_t>::max());
  for (const gaffa_hip_candidate_tie& t : ties) {
    if (t.segment >= nseg || t.peak >= n)
      throw std::logic_error("candidate selection: tie record out of range");
    std::uint32_t& b = best[t.segment];
    if (b == std::numeric_limits<std::uint32_t>::max() ||
        is_better_dm_peak(expanded_[t.peak], expanded_[b]))
      b = t.peak;
  }
  for (std::uint32_t b : best)
    if (b == std::numeric_limits<std::uint32_t>::max())
      throw std::logic_error("candidate selection: a cluster without its top peak");
  std::vector<Candidate> out(nseg);
#pragma omp parallel for schedule(static)
  for (std::int64_t j = 0; j < static_cast<std::int64_t>(nseg); ++j) {
    const gaffa_hip_candidate_segment& g = seg[static_cast<std::size_t>(j)];
    const DmPeak& p = expanded_[best[static_cast<std::size_t>(j)]];
    Candidate& c = out[static_cast<std::size_t>(j)];
    c.dm = p.dm;
    c.dm_index = p.dm_index;
    c.period = p.peak.period;
    c.frequency = p.peak.frequency;
    c.width = p.peak.width;
    c.duty_cycle = p.peak.duty_cycle;
    c.snr = p.peak.snr;
    c.peak_count = g.count;
    c.dm_index_min = g.dm_min;
    c.dm_index_max = g.dm_max;
    c.frequency_min = freq_of_rdesc_[g.rdesc_min];
    c.frequency_max = freq_of_rdesc_[g.rdesc_max];
    c.best_peak = p;
  }
  __gnu_parallel::sort(out.begin(), out.end(), candidate_before);
  if (o.max_candidates != 0 && out.size() > o.max_candidates) out.resize(o.max_candidates);
  return out;
}

std::vector<Candidate> GpuSearchPipeline::select_candidates_host(
    const CandidateSelectionOptions& options) const {
  wait_peaks();
  const double obs = static_cast<double>(n_out_) * cfg_.tsamp;
  const std::size_t n = peaks_.size();
  const auto& tasks = ffa_plan_.tasks;
  if (n == 0 || period_ranks_.empty() || expanded_.size() < n ||
      compact_host_.bytes() < n * sizeof(gaffa_hip_ffa_peak))
    return select_candidates_cpu(peaks_, obs, options);
  // the peaks in arrival order (the same multiset as peaks_) and the dense
  // rank of each one's trial period, from its task and row
  std::vector<std::size_t> first_row(tasks.size() + 1, 0);
  for (std::size_t k = 0; k < tasks.size(); ++k)
    first_row[k + 1] = first_row[k] + tasks[k].rows_eval;
  const auto* compact = static_cast<const gaffa_hip_ffa_peak*>(compact_host_.data());
  std::vector<std::uint32_t> rank(n);
  bool ok = period_ranks_.size() == first_row.back();
#pragma omp parallel for reduction(&& : ok) schedule(static)
  for (std::int64_t i = 0; i < static_cast<std::int64_t>(n); ++i) {
    const gaffa_hip_ffa_peak& c = compact[i];
    const bool in = c.task < tasks.size() && c.shift < tasks[c.task].rows_eval;
    ok = ok && in;
    rank[static_cast<std::size_t>(i)] = in ? period_ranks_[first_row[c.task] + c.shift] : 0u;
  }
  if (!ok) return select_candidates_cpu(peaks_, obs, options);
  return select_candidates_ranked_cpu(std::span<const DmPeak>(expanded_.data(), n), rank, obs,
                                      options);
}

StageTimes GpuSearchPipeline::run() {
  GpuDeviceScope scope(cfg_.device_id);
  gaffa_hip_stream st = stream_.get();
  hip_ok(gaffa_hip_event_record(events_[0], st), "hipEventRecord");
  dedisperse();
  hip_ok(gaffa_hip_event_record(events_[1], st), "hipEventRecord");
  preprocess();
  hip_ok(gaffa_hip_event_record(events_[2], st), "hipEventRecord");
  if (prep_) prep_->synchronize();
  search();  // records events_[3] / events_[4] around the FFA launches
  stream_.synchronize();
  StageTimes t;
  hip_ok(gaffa_hip_event_elapsed_ms(events_[0], events_[1], &t.dedisperse_ms), "elapsed");
  hip_ok(gaffa_hip_event_elapsed_ms(events_[1], events_[2], &t.preprocess_ms), "elapsed");
  hip_ok(gaffa_hip_event_elapsed_ms(events_[3], events_[4], &t.ffa_ms), "elapsed");
  hip_ok(gaffa_hip_event_elapsed_ms(events_[0], events_[4], &t.total_ms), "elapsed");
  for (std::size_t k = 0; k < marks_used_; ++k) {
    for (int j = 0; j < FfaStageMarks::kStages; ++j) {
      float ms = 0.0F;
      hip_ok(gaffa_hip_event_elapsed_ms(marks_[k].at[j], marks_[k].at[j + 1], &ms), "elapsed");
      t.ffa_stage_ms[j] += ms;
    }
  }
  t.ffa_tiles = static_cast<std::uint32_t>(marks_used_);
  return t;
}

void GpuSearchPipeline::download_plane(void* out) const {
  GpuDeviceScope scope(cfg_.device_id);
  std::vector<std::uint32_t> tmp(ndm_ * stride_);
  copy_to_host(tmp.data(), plane_.data(), tmp.size() * 4, stream_.get());
  auto* o = static_cast<std::uint32_t*>(out);
  for (std::size_t i = 0; i < ndm_; ++i)
    std::copy_n(tmp.data() + i * stride_, n_out_, o + i * n_out_);
}

}  // namespace gaffa

namespace gaffa {

void GpuSearchPipeline::enable_periodogram(bool on) {
  GpuDeviceScope scope(cfg_.device_id);
  if (!on) {
    ffa_->set_periodogram(nullptr, 0, 0);
    periodogram_ = GpuDeviceBuffer<float>();
    return;
  }
  if (trial_periods_.empty()) {
    trial_periods_ = ffa_plan_trial_periods(ffa_plan_);
    // rows by period (plan order is ascending for riptide plans, but the
    // lookup must not depend on it)
    trial_order_.resize(trial_periods_.size());
    for (std::size_t i = 0; i < trial_order_.size(); ++i)
      trial_order_[i] = static_cast<std::uint32_t>(i);
    std::stable_sort(trial_order_.begin(), trial_order_.end(),
                     [&](std::uint32_t a, std::uint32_t b) {
                       return trial_periods_[a] < trial_periods_[b];
                     });
  }
  const std::size_t widths = periodogram_widths();
  const std::size_t stride = trial_periods_.size() * widths;
  periodogram_ = GpuDeviceBuffer<float>(ndm_ * stride);
  // entries of widths a task does not evaluate stay 0
  hip_ok(gaffa_hip_fill_u32(reinterpret_cast<std::uint32_t*>(periodogram_.data()), 0,
                            periodogram_.size(), stream_.get()),
         "fill_u32");
  ffa_->set_periodogram(periodogram_.data(), stride, static_cast<std::uint32_t>(widths));
  stream_.synchronize();
}

void GpuSearchPipeline::download_periodogram(std::size_t dm_index, float* out) const {
  check_arg(periodogram_enabled(), "the periodogram is not enabled");
  check_arg(dm_index < ndm_, "dm_index out of range");
  GpuDeviceScope scope(cfg_.device_id);
  const std::size_t stride = trial_periods_.size() * periodogram_widths();
  copy_to_host(out, periodogram_.data() + dm_index * stride, stride * sizeof(float),
               stream_.get());
  stream_.synchronize();
}

void GpuSearchPipeline::candidate_features(std::span<const Candidate> candidates, float* cubes,
                                           float* patches) {
  check_arg(!candidates.empty(), "candidate list is empty");
  check_arg(periodogram_enabled(), "candidate features need the periodogram (enable it "
                                   "before the search)");
  check_arg(candidates.size() < 65536, "at most 65535 candidates per call");
  GpuDeviceScope scope(cfg_.device_id);
  const std::size_t widths = periodogram_widths();
  const std::size_t nsub = GAFFA_HIP_SCORER_SUBINTS, nbin = GAFFA_HIP_SCORER_PHASE_BINS;
  std::vector<FoldJob> jobs;
  std::vector<std::uint32_t> series, row, width;
  for (const Candidate& c : candidates) {
    check_arg(c.dm_index < ndm_, "candidate dm_index out of range");
    const FfaPeak& p = c.best_peak.peak;
    const auto it = std::lower_bound(
        trial_order_.begin(), trial_order_.end(), p.period,
        [&](std::uint32_t r, double v) { return trial_periods_[r] < v; });
    check_arg(it != trial_order_.end() && trial_periods_[*it] == p.period,
              "candidate period is not a trial period of this pipeline's plan");
    check_arg(p.width_index < widths, "candidate width index out of range");
    jobs.push_back(FoldJob{c.dm_index * stride_, p.period, 0});
    series.push_back(static_cast<std::uint32_t>(c.dm_index));
    row.push_back(*it);
    width.push_back(static_cast<std::uint32_t>(p.width_index));
  }
  const GpuFoldInput in{plane_.data(), FoldDtype::F32, n_out_, 1, 1, cfg_.device_id};
  fold_batch_gpu_device(in, jobs, cfg_.tsamp, nbin, (n_out_ + nsub - 1) / nsub, 1, cubes,
                        stream_.get());
  hip_ok(gaffa_hip_scorer_normalise_rows(cubes, static_cast<std::uint32_t>(candidates.size()),
                                         static_cast<std::uint32_t>(nsub * nbin), stream_.get()),
         "scorer_normalise_rows");
  GpuDeviceBuffer<std::uint32_t> d_series = gaffa::upload(series.data(), series.size());
  GpuDeviceBuffer<std::uint32_t> d_row = gaffa::upload(row.data(), row.size());
  GpuDeviceBuffer<std::uint32_t> d_width = gaffa::upload(width.data(), width.size());
  gaffa_hip_scorer_patch_args a{};
  a.periodogram = periodogram_.data();
  a.series_stride = trial_periods_.size() * widths;
  a.rows = trial_periods_.size();
  a.widths = static_cast<std::uint32_t>(widths);
  a.series = d_series.data();
  a.row = d_row.data();
  a.width = d_width.data();
  a.n = static_cast<std::uint32_t>(candidates.size());
  a.out = patches;
  hip_ok(gaffa_hip_scorer_gather_patches(&a, stream_.get()), "scorer_gather_patches");
  stream_.synchronize();
}

double GpuSearchPipeline::fold_candidates_full(std::span<const Candidate> candidates,
                                               std::size_t nbin, std::size_t nsubint,
                                               std::size_t output_channels,
                                               std::vector<float>* cubes) {
  check_arg(nbin > 0 && nsubint > 0, "fold nbin and nsubint must be positive");
  const std::size_t nout = output_channels == 0 ? cfg_.nchans : output_channels;
  check_arg(nout <= cfg_.nchans && cfg_.nchans % nout == 0,
            "fold output_channels must divide the channel count");
  GpuDeviceScope scope(cfg_.device_id);
  const gaffa_hip_stream st = stream_.get();
  GpuDeviceBuffer<float> cube(nsubint * nout * nbin);
  if (cubes != nullptr) cubes->assign(candidates.size() * cube.size(), 0.0f);
  const double fmax = *std::max_element(cfg_.frequency_mhz.begin(), cfg_.frequency_mhz.end());
  const std::size_t factor = cfg_.nchans / nout;
  if (channels_.size() != 0 && factor <= 256) {
    // subband workloads keep a channel-major copy: fold each candidate from
    // its per-output-channel sums at the candidate's DM (the same cube)
    const std::size_t sum_stride = (cfg_.nsamples + 7) / 8 * 8;
    GpuDeviceBuffer<std::uint16_t> sums(nout * sum_stride);
    GpuDeviceBuffer<std::int32_t> d_delays(cfg_.nchans);
    hip_ok(gaffa_hip_event_record(events_[5], st), "hipEventRecord");
    for (std::size_t i = 0; i < candidates.size(); ++i) {
      const Candidate& c = candidates[i];
      const MultiDmDedispersionPlan plan{c.dm, 1.0, 1, fmax, cfg_.tsamp, 0, cfg_.nchans};
      const std::size_t n = multi_dm_output_nsamples(cfg_.nsamples, cfg_.frequency_mhz, plan);
      check_arg(n > 0, "candidate DM leaves no aligned samples");
      const std::vector<std::int32_t> delays = make_multi_dm_delay_table(cfg_.frequency_mhz, plan);
      const std::int32_t dmax = *std::max_element(delays.begin(), delays.end());
      check_arg(*std::min_element(delays.begin(), delays.end()) >= 0,
                "dedispersion requires non-negative delays");
      // the host table stays alive until the copy has run (synchronous copy)
      hip_ok(gaffa_hip_memcpy(d_delays.data(), delays.data(), d_delays.bytes(), nullptr),
             "hipMemcpy");
      gaffa_hip_subband_spectrum_args a{};
      a.channels = channels_.data();
      a.channel_stride = channel_stride_;
      a.nsel = cfg_.nchans;
      a.subband_channels = factor;
      a.delays = d_delays.data();
      a.max_delay = dmax;
      a.out = sums.data();
      a.out_stride = sum_stride;
      a.out_nsamples = n;
      hip_ok(gaffa_hip_subband_spectrum_u8(&a, st), "subband_spectrum");
      const FoldJob job{0, c.period, 0};
      GpuFoldInput in{sums.data(), FoldDtype::U16, n, nout, 1, cfg_.device_id};
      in.channel_stride = sum_stride;
      in.divisor = static_cast<double>(factor);
      fold_batch_gpu_device(in, std::span<const FoldJob>(&job, 1), cfg_.tsamp, nbin,
                            (n + nsubint - 1) / nsubint, nout, cube.data(), st);
      if (cubes != nullptr)
        copy_to_host(cubes->data() + i * cube.size(), cube.data(), cube.bytes(), st);
    }
    hip_ok(gaffa_hip_event_record(events_[6], st), "hipEventRecord");
    stream_.synchronize();
    float ms = 0.0f;
    hip_ok(gaffa_hip_event_elapsed_ms(events_[5], events_[6], &ms), "elapsed");
    return ms;
  }
  GpuDeviceBuffer<std::uint8_t> spectrum(cfg_.nsamples * cfg_.nchans);
  GpuDedispersionOptions o;
  o.device_id = cfg_.device_id;
  o.stream = st;
  hip_ok(gaffa_hip_event_record(events_[5], st), "hipEventRecord");
  for (std::size_t i = 0; i < candidates.size(); ++i) {
    const Candidate& c = candidates[i];
    const SingleDmDedispersionPlan plan{c.dm, fmax, cfg_.tsamp, 0, cfg_.nchans};
    const std::size_t n = spectrum_output_nsamples(cfg_.nsamples, cfg_.frequency_mhz, plan);
    check_arg(n > 0, "candidate DM leaves no aligned samples");
    dedisperse_spectrum_gpu_raw(filterbank_.data(), 1, SampleShape{cfg_.nsamples, 1, cfg_.nchans},
                                cfg_.frequency_mhz, plan, spectrum.data(), n, o);
    const FoldJob job{0, c.period, 0};
    fold_batch_gpu_device(GpuFoldInput{spectrum.data(), FoldDtype::U8, n, cfg_.nchans, 0,
                                       cfg_.device_id},
                          std::span<const FoldJob>(&job, 1), cfg_.tsamp, nbin,
                          (n + nsubint - 1) / nsubint, nout, cube.data(), st);
    if (cubes != nullptr)
      copy_to_host(cubes->data() + i * cube.size(), cube.data(), cube.bytes(), st);
  }
  hip_ok(gaffa_hip_event_record(events_[6], st), "hipEventRecord");
  stream_.synchronize();
  float ms = 0.0f;
  hip_ok(gaffa_hip_event_elapsed_ms(events_[5], events_[6], &ms), "elapsed");
  return ms;
}

}  // namespace gaffa
