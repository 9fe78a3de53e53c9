// Batched FFA search over series with their own lengths and plans.
#include "gaffa/ffa_batch.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <parallel/algorithm>
#include <stdexcept>

namespace gaffa {
namespace {

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

std::size_t align64(std::size_t v) { return (v + 63) / 64 * 64; }

// Workspace floats one member adds to a tile, by GpuFfaExecutionPlan's own
// accounting: one 64-aligned prepared copy per downsample factor != 1, and two
// 64-aligned rows*bins regions per task with rows >= 2.
std::size_t member_workspace_floats(const FfaSearchPlan& p) {
  std::map<double, std::size_t> prepared;
  std::size_t work = 0;
  for (const FfaSearchTask& t : p.tasks) {
    if (t.downsample_factor != 1.0) prepared[t.downsample_factor] = align64(t.prepared_nsamples);
    if (t.rows >= 2) work += align64(t.rows * t.bins);
  }
  std::size_t total = 2 * work;
  for (const auto& kv : prepared) total += kv.second;
  return total;
}

}  // namespace

struct GpuFfaBatchedSearch::Tile {
  std::size_t first = 0, count = 0;
  std::unique_ptr<GpuFfaExecutionPlan> plan;
  GpuDeviceBuffer<gaffa_hip_ffa_group> groups;
  GpuDeviceBuffer<gaffa_hip_ffa_task> tasks;
  GpuDeviceBuffer<gaffa_hip_ffa_item> subtree, level, terminal;
  GpuDeviceBuffer<gaffa_hip_ffa_level_span> level_spans;
  GpuDeviceBuffer<gaffa_hip_ffa_level_row> level_rows;
  GpuDeviceBuffer<gaffa_hip_ffa_merge> merges;
  gaffa_hip_ffa_plan_desc desc{};
};

GpuFfaBatchedSearch::GpuFfaBatchedSearch(std::vector<FfaSearchPlan> plans,
                                         const GpuFfaBatchedOptions& options)
    : opt_(options), plans_(std::move(plans)) {
  check_arg(!plans_.empty(), "FFA batch must contain at least one series");
  for (const FfaSearchPlan& p : plans_) {
    check_arg(!p.tasks.empty(), "FFA executor plan must contain at least one task");
    check_arg(p.width_trials == plans_.front().width_trials,
              "FFA batch members must share one width-trial list");
    offsets_.push_back(row_length_);
    lengths_.push_back(p.tasks.front().input_nsamples);
    row_length_ += align64(p.tasks.front().input_nsamples);
  }
  check_arg(plans_.size() < (1ULL << 31), "FFA batch too large");
  GpuDeviceScope scope(opt_.device_id);
  stream_ = GpuStream::create();
  row_ = GpuDeviceBuffer<float>(row_length_);
  count_ = GpuDeviceBuffer<std::uint32_t>(1);
  peaks_ = GpuDeviceBuffer<gaffa_hip_ffa_peak>(
      std::max<std::size_t>(1, opt_.initial_peak_buffer_bytes / sizeof(gaffa_hip_ffa_peak)));
  for (auto& e : ev_) hip_ok(gaffa_hip_event_create(&e), "hipEventCreate");
  set_series_per_tile(opt_.series_per_tile);
}

GpuFfaBatchedSearch::~GpuFfaBatchedSearch() {
  if (graph_ != nullptr) (void)gaffa_hip_graph_destroy(graph_);
  for (auto e : ev_)
    if (e != nullptr) (void)gaffa_hip_event_destroy(e);
}

void GpuFfaBatchedSearch::upload(const float* host_row, std::size_t count) {
  check_arg(count == row_length_, "batch row length does not match");
  GpuDeviceScope scope(opt_.device_id);
  copy_to_device(row_.data(), host_row, count * sizeof(float), nullptr);
}

void GpuFfaBatchedSearch::set_series_per_tile(std::size_t n) {
  check_arg(n > 0, "series_per_tile must be positive");
  series_per_tile_ = n;
  build_tiles();
}

void GpuFfaBatchedSearch::set_item_shape(FfaItemShape shape) {
  opt_.item_shape = shape;
  build_tiles();
}

void GpuFfaBatchedSearch::build_tiles() {
  GpuDeviceScope scope(opt_.device_id);
  if (graph_ != nullptr) {
    hip_ok(gaffa_hip_graph_destroy(graph_), "hipGraphExecDestroy");
    graph_ = nullptr;
  }
  tiles_.clear();
  tile_of_.assign(plans_.size(), -1);
  std::size_t max_prepared = 1, max_work = 1;
  std::vector<std::size_t> need(plans_.size());
  for (std::size_t i = 0; i < plans_.size(); ++i) need[i] = member_workspace_floats(plans_[i]);
  for (std::size_t i0 = 0; i0 < plans_.size();) {
    std::vector<FfaBatchMember> members;
    std::size_t i = i0, floats = 0;
    // grow the tile while it fits the workspace limit (one series always fits)
    while (i < plans_.size() && members.size() < series_per_tile_) {
      if (!members.empty() && opt_.workspace_bytes_limit != 0 &&
          (floats + need[i]) * sizeof(float) > opt_.workspace_bytes_limit)
        break;
      members.push_back(FfaBatchMember{&plans_[i], offsets_[i], static_cast<std::uint32_t>(i)});
      floats += need[i];
      ++i;
    }
    auto t = std::make_unique<Tile>();
    t->first = i0;
    t->count = i - i0;
    t->plan = std::make_unique<GpuFfaExecutionPlan>(members, opt_.item_shape);
    const GpuFfaExecutionPlan& e = *t->plan;
    t->groups = gaffa::upload(e.groups().data(), e.groups().size());
    t->tasks = gaffa::upload(e.tasks().data(), e.tasks().size());
    t->subtree = gaffa::upload(e.subtree_items().data(), e.subtree_items().size());
    t->merges = gaffa::upload(e.subtree_merges().data(), e.subtree_merges().size());
    t->level = gaffa::upload(e.level_items().data(), e.level_items().size());
    t->level_spans = gaffa::upload(e.level_spans().data(), e.level_spans().size());
    t->level_rows = gaffa::upload(e.level_leaf_rows().data(), e.level_leaf_rows().size());
    t->terminal = gaffa::upload(e.terminal_items().data(), e.terminal_items().size());
    gaffa_hip_ffa_plan_desc& d = t->desc;
    d.groups = t->groups.data();
    d.ngroups = static_cast<std::uint32_t>(e.groups().size());
    d.tasks = t->tasks.data();
    d.ntasks = static_cast<std::uint32_t>(e.tasks().size());
    d.subtree_items = t->subtree.data();
    d.nsubtree_items = static_cast<std::uint32_t>(e.subtree_items().size());
    d.subtree_merges = t->merges.data();
    d.level_items = t->level.data();
    d.level_spans = t->level_spans.data();
    d.level_leaf_rows = t->level_rows.data();
    d.level_item_offsets = e.level_offsets().data();
    d.max_depth = static_cast<std::uint32_t>(e.max_depth());
    d.terminal_items = t->terminal.data();
    d.nterminal_items = static_cast<std::uint32_t>(e.terminal_items().size());
    d.max_bins = e.max_bins();
    d.subtree_lds_bytes = e.subtree_lds_bytes();
    d.subtree_half_floats = e.subtree_half_floats();
    d.max_subtree_desc = e.max_subtree_desc();
    d.terminal_lds_bytes = e.terminal_lds_bytes();
    d.level_lds_bytes = e.level_lds_bytes();
    d.prep_blocks = e.prep_blocks();
    d.max_prepared_nsamples = e.max_prepared();
    d.prepared_stride = e.prepared_stride();
    d.work_stride = e.work_stride();
    max_prepared = std::max(max_prepared, e.prepared_stride());
    max_work = std::max(max_work, e.work_stride());
    for (std::size_t k = i0; k < i; ++k) tile_of_[k] = static_cast<int>(tiles_.size());
    tiles_.push_back(std::move(t));
    i0 = i;
  }
  if (prepared_.size() < max_prepared) prepared_ = GpuDeviceBuffer<float>(max_prepared);
  if (work_.size() < 2 * max_work) work_ = GpuDeviceBuffer<float>(2 * max_work);
}

void GpuFfaBatchedSearch::enqueue(float thr) {
  hip_ok(gaffa_hip_fill_u32(count_.data(), 0, 1, stream_.get()), "fill_u32");
  for (const auto& t : tiles_) {
    gaffa_hip_ffa_search_args a{};
    a.plan = &t->desc;
    a.series = row_.data();
    a.nseries = 1;
    a.nsamples = row_length_;
    a.series_stride = row_length_;
    a.series_index_base = 0;
    a.snr_threshold = thr;
    a.prepared = prepared_.data();
    a.work = work_.data();
    a.peaks = peaks_.data();
    a.peak_capacity = peaks_.size();
    a.peak_count = count_.data();
    hip_ok(gaffa_hip_ffa_search_batch(&a, stream_.get()), "ffa_search_batch");
  }
}

std::uint32_t GpuFfaBatchedSearch::read_count() {
  std::uint32_t n = 0;
  copy_to_host(&n, count_.data(), sizeof(n), stream_.get());
  return n;
}

FfaBatchSearchResult GpuFfaBatchedSearch::run(float thr, FfaBatchTiming* timing) {
  check_arg(std::isfinite(thr), "FFA search S/N threshold must be finite");
  GpuDeviceScope scope(opt_.device_id);
  std::uint32_t count = 0;
  for (;;) {
    if (opt_.use_graph && (graph_ == nullptr || graph_threshold_ != thr)) {
      if (graph_ != nullptr) hip_ok(gaffa_hip_graph_destroy(graph_), "hipGraphExecDestroy");
      graph_ = nullptr;
      // the capture records kernel launches only: the counter reset is a
      // fill kernel and every tile's launch limits are raised beforehand
      for (const auto& t : tiles_) hip_ok(gaffa_hip_ffa_reserve_launch(&t->desc), "ffa_reserve");
      hip_ok(gaffa_hip_graph_begin(stream_.get()), "hipStreamBeginCapture");
      try {
        enqueue(thr);
      } catch (...) {
        gaffa_hip_graph abandoned = nullptr;  // leave capture mode before reporting
        if (gaffa_hip_graph_end(stream_.get(), &abandoned) == GAFFA_HIP_OK && abandoned)
          (void)gaffa_hip_graph_destroy(abandoned);
        throw;
      }
      hip_ok(gaffa_hip_graph_end(stream_.get(), &graph_), "hipStreamEndCapture");
      graph_threshold_ = thr;
    }
    hip_ok(gaffa_hip_event_record(ev_[0], stream_.get()), "hipEventRecord");
    if (opt_.use_graph) {
      hip_ok(gaffa_hip_graph_launch(graph_, stream_.get()), "hipGraphLaunch");
    } else {
      enqueue(thr);
    }
    hip_ok(gaffa_hip_event_record(ev_[1], stream_.get()), "hipEventRecord");
    count = read_count();
    if (count <= peaks_.size()) break;
    const std::size_t max_count = opt_.max_peak_buffer_bytes / sizeof(gaffa_hip_ffa_peak);
    if (count > max_count)
      throw std::runtime_error("FFA compact peak buffer exceeded max_peak_buffer_bytes");
    peaks_ = GpuDeviceBuffer<gaffa_hip_ffa_peak>(
        std::min(max_count, std::max<std::size_t>(count, 2 * peaks_.size())));
    if (graph_ != nullptr) {  // the captured launches hold the old buffer
      hip_ok(gaffa_hip_graph_destroy(graph_), "hipGraphExecDestroy");
      graph_ = nullptr;
    }
  }
  if (timing != nullptr) {
    hip_ok(gaffa_hip_event_elapsed_ms(ev_[0], ev_[1], &timing->gpu_ms), "elapsed");
    timing->tiles = tiles_.size();
    std::size_t launches = 1;  // counter reset
    for (const auto& t : tiles_) {
      std::size_t levels = 0;
      const auto& off = t->plan->level_offsets();
      for (std::size_t d = 1; d + 1 < off.size(); ++d) levels += off[d + 1] > off[d] ? 1 : 0;
      launches += (t->plan->max_prepared() != 0) + (!t->plan->subtree_items().empty()) + levels + 1;
    }
    timing->launches = launches;
  }
  return collect(count);
}

FfaBatchSearchResult GpuFfaBatchedSearch::run_per_series(float thr, FfaBatchTiming* timing) {
  const std::size_t keep = series_per_tile_;
  const bool keep_graph = opt_.use_graph;
  opt_.use_graph = false;
  set_series_per_tile(1);
  FfaBatchSearchResult r;
  try {
    r = run(thr, timing);
  } catch (...) {
    opt_.use_graph = keep_graph;
    set_series_per_tile(keep);
    throw;
  }
  opt_.use_graph = keep_graph;
  set_series_per_tile(keep);
  return r;
}

namespace {

// A batch's shape for the tuning cache: the members' lengths and plan sizes
// (the tuned knobs change the time only, never the results, so two batches
// that merely collide share a configuration and stay correct).
std::uint64_t batch_shape_key(const std::vector<FfaSearchPlan>& plans) {
  std::uint64_t h = 1469598103934665603ULL;
  const auto mix = [&](std::uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
  };
  mix(plans.size());
  for (const FfaSearchPlan& p : plans) {
    mix(p.tasks.front().input_nsamples);
    mix(p.tasks.size());
    mix(p.tasks.front().bins);
    mix(p.tasks.back().rows);
  }
  return h;
}

std::map<std::uint64_t, FfaBatchTuning>& tuning_cache() {
  static std::map<std::uint64_t, FfaBatchTuning> cache;
  return cache;
}

std::mutex& tuning_mutex() {
  static std::mutex m;
  return m;
}

}  // namespace

std::size_t GpuFfaBatchedSearch::autotune(const std::vector<std::size_t>& tile_candidates,
                                          const std::vector<std::uint32_t>& level_rows,
                                          const std::vector<std::uint32_t>& terminal_rows) {
  check_arg(!tile_candidates.empty(), "autotune needs at least one candidate");
  const std::uint64_t key = batch_shape_key(plans_);
  {
    std::lock_guard<std::mutex> lock(tuning_mutex());
    const auto it = tuning_cache().find(key);
    if (it != tuning_cache().end()) {
      tuning_ = it->second;
      tuning_.cached = true;
      opt_.item_shape = tuning_.item_shape;
      set_series_per_tile(tuning_.series_per_tile);
      return tuning_.series_per_tile;
    }
  }
  const auto time_current = [&]() {
    FfaBatchTiming t;
    (void)run(1e30F, &t);  // no peaks: times the transform and detection alone
    (void)run(1e30F, &t);  // second run: graph captured, caches warm
    return t.gpu_ms;
  };
  FfaBatchTuning best;
  for (std::size_t c : tile_candidates) {
    if (c == 0) continue;
    set_series_per_tile(c);
    const float ms = time_current();
    if (best.series_per_tile == 0 || ms < best.gpu_ms) {
      best.series_per_tile = c;
      best.gpu_ms = ms;
    }
  }
  check_arg(best.series_per_tile != 0, "autotune needs a positive candidate");
  series_per_tile_ = best.series_per_tile;
  best.item_shape = opt_.item_shape;
  const auto tune_rows = [&](const std::vector<std::uint32_t>& values, bool level) {
    for (std::uint32_t v : values) {
      FfaItemShape s = best.item_shape;
      (level ? s.level_rows : s.terminal_rows) = v;
      const std::uint32_t cur = level ? best.item_shape.level_rows : best.item_shape.terminal_rows;
      const std::uint32_t dflt = level ? gaffa_hip_ffa_level_rows() : gaffa_hip_ffa_terminal_rows();
      if (v == 0 || v == (cur == 0 ? dflt : cur)) continue;  // the configuration already timed
      set_item_shape(s);
      const float ms = time_current();
      if (ms < best.gpu_ms) {
        best.item_shape = s;
        best.gpu_ms = ms;
      }
    }
  };
  tune_rows(level_rows, true);
  tune_rows(terminal_rows, false);
  opt_.item_shape = best.item_shape;
  set_series_per_tile(best.series_per_tile);
  tuning_ = best;
  std::lock_guard<std::mutex> lock(tuning_mutex());
  tuning_cache()[key] = best;
  return best.series_per_tile;
}

FfaBatchSearchResult GpuFfaBatchedSearch::collect(std::uint32_t count) const {
  std::vector<gaffa_hip_ffa_peak> compact(std::min<std::size_t>(count, peaks_.size()));
  if (!compact.empty())
    copy_to_host(compact.data(), peaks_.data(), compact.size() * sizeof(compact[0]),
                 stream_.get());
  // detection plans by bins (all members share the width trials)
  std::map<std::size_t, FfaDetectionPlan> det;
  for (const auto& t : tiles_)
    for (const FfaSearchTask& task : t->plan->plan().tasks)
      if (det.count(task.bins) == 0)
        det.emplace(task.bins, make_ffa_detection_plan(plans_.front().width_trials, task.bins));
  for (const gaffa_hip_ffa_peak& p : compact) {
    if (p.series >= plans_.size())
      throw std::logic_error("FFA compact peak refers to an unknown series");
    const auto& tasks = tiles_[tile_of_[p.series]]->plan->plan().tasks;
    if (p.task >= tasks.size() ||
        (p.phase_width >> 16) >= det.at(tasks[p.task].bins).boxcar_trials.size())
      throw std::logic_error("FFA compact peak refers to an unknown task or width");
  }
  FfaBatchSearchResult r;
  r.peaks.resize(compact.size());
#pragma omp parallel for schedule(static)
  for (std::int64_t i = 0; i < static_cast<std::int64_t>(compact.size()); ++i) {
    const gaffa_hip_ffa_peak& p = compact[static_cast<std::size_t>(i)];
    const FfaSearchTask& task = tiles_[tile_of_[p.series]]->plan->plan().tasks[p.task];
    const FfaBoxcarTrial& trial = det.at(task.bins).boxcar_trials[p.phase_width >> 16];
    r.peaks[static_cast<std::size_t>(i)] =
        FfaBatchPeak{p.series, make_ffa_peak(task, trial, p.shift, p.phase_width & 0xFFFFu, p.snr)};
  }
  __gnu_parallel::sort(r.peaks.begin(), r.peaks.end(),
                       [](const FfaBatchPeak& a, const FfaBatchPeak& b) {
                         if (a.series_index != b.series_index)
                           return a.series_index < b.series_index;
                         return is_better_ffa_peak(a.peak, b.peak);
                       });
  return r;
}

}  // namespace gaffa
